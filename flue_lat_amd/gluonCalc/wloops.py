"""Gluonic observables with the names of FLUE's Fortran module FLUE_wloops
(src/wloops.F90), callable from Python.

Every function accepts the gauge field as
  * a numpy complex array of shape (NT, NX, NY, NZ, 4, 3, 3) -- handled like the
    f2py module: copied to the GPU, measured there, results returned as numpy/floats;
  * a torch complex tensor of that shape (any strides, CPU or GPU) -- re-laid out
    on the GPU without a host round trip when it is already there;
  * a :class:`flue_lat_amd.GaugeField` -- already resident in the SoA layout.
Lattice dimensions are taken from the array, as f2py does for the Fortran
explicit-shape arguments.  Directions are the reference's: 1 = time, 2..4 = x, y, z.
"""
import numpy as np

from ..compiled import flue_c as fc


def _is_torch(x):
    try:
        import torch
        return isinstance(x, torch.Tensor)
    except ImportError:
        return False


def _device_field(data):
    from ..field import GaugeField
    if isinstance(data, GaugeField):
        return data
    if _is_torch(data):
        return GaugeField.from_complex(data, "fp64")
    return None


def periodCoord(coord, dataShape):
    """src/wloops.F90:546-565"""
    return fc.periodcoord_c(coord, dataShape)


def genericPath(data, coordBase, path):
    """src/wloops.F90:20-57 -- the ordered product along `path`, transposed."""
    f = _device_field(data)
    if f is None:
        return fc.genericpath_c(data, coordBase, path)
    return f.generic_path(coordBase, path)


def genPlaquette(data, muStart, muEnd, nuEnd):
    """src/wloops.F90:59-127 -- (sumTrP, nP, time)."""
    f = _device_field(data)
    if f is None:
        return fc.genplaquette_c(data, muStart, muEnd, nuEnd)
    import time
    t0 = time.perf_counter()
    s, n = f.plaquette(muStart, muEnd, nuEnd)
    return s, n, time.perf_counter() - t0


def plaquette(data, muStart, muEnd, nuEnd):
    """src/wloops.F90:915-1017 -- (sumTrP, nP, time)."""
    f = _device_field(data)
    if f is None:
        return fc.plaquette_c(data, muStart, muEnd, nuEnd)
    return genPlaquette(f, muStart, muEnd, nuEnd)


def polyakov(data):
    """src/wloops.F90:1019-1069 -- (sumTrP, nP, time); NT < 3 raises FlueError."""
    f = _device_field(data)
    if f is None:
        return fc.polyakov_c(data)
    import time
    t0 = time.perf_counter()
    if f.dims[0] < 3:
        from .._lib import FlueError
        raise FlueError(4, "polyakov")
    s, n = f.polyakov()
    return s, n, time.perf_counter() - t0


def cloverLoopMuNuCoord(data, mu, nu):
    """src/wloops.F90:129-199 -- (NT,NX,NY,NZ,3,3) complex."""
    f = _device_field(data)
    if f is None:
        return fc.cloverloopmunucoord_c(data, mu, nu)
    return f.fmunu(mu, nu, clover=True)


def Loop5MuNuCoord(data, mu, nu):
    """src/wloops.F90:201-432 -- (NT,NX,NY,NZ,3,3) complex."""
    f = _device_field(data)
    if f is None:
        return fc.loop5munucoord_c(data, mu, nu)
    return f.fmunu(mu, nu)


def magnetic(data):
    """src/wloops.F90:479-544 -- B."""
    f = _device_field(data)
    if f is None:
        return fc.magnetic_c(data)
    return f.magnetic()


def wmunu(data, mu, nu, nMu, nNu):
    """src/wloops.F90:679-791"""
    f = _device_field(data)
    if f is None:
        return fc.wmunu_c(data, mu, nu, nMu, nNu)
    return f.wilson_loop(mu, nu, nMu, nNu)


def wmunutau(data, tau, mu, nu, nMu, nNu):
    """src/wloops.F90:567-677"""
    f = _device_field(data)
    if f is None:
        return fc.wmunutau_c(data, tau, mu, nu, nMu, nNu)
    return f.wilson_loop(mu, nu, nMu, nNu, tau=tau)


__all__ = ["periodCoord", "genericPath", "genPlaquette", "plaquette", "polyakov", "cloverLoopMuNuCoord",
           "Loop5MuNuCoord", "magnetic", "wmunu", "wmunutau"]
