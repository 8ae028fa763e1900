"""f2py-style front end of libflue_hip.so's host drop-ins.

Mirrors the module numpy.f2py builds from c_wrapper/FLUE_c.f90 with --lower
(python/meson.build:55-67): lattice dimensions are inferred from the array shapes,
non-Fortran-ordered inputs are copied to Fortran order, and the ``intent(out)``
arguments come back as a tuple in argument-list order.  A nonzero status from the
library raises FlueError (the Fortran routines would stop or run into undefined
behaviour instead).

The field ``data`` is a complex array of shape (NT, NX, NY, NZ, 4, 3, 3) indexed
[t, x, y, z, mu, a, b] = (U_mu(x))_ab, mu = 0 time (src/wloops.F90:21).
"""
import ctypes

import numpy as np

from .. import _lib
from .._lib import check

def _c128(a, shape=None):
    a = np.asfortranarray(np.asarray(a), dtype=np.complex128)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _field(data):
    U = _c128(data)
    if U.ndim != 7 or U.shape[4:] != (4, 3, 3):
        raise ValueError("data must have shape (NT,NX,NY,NZ,4,3,3)")
    return U


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ri(v):
    return ctypes.byref(ctypes.c_int32(int(v)))


def _dims(U):
    return [ctypes.byref(ctypes.c_int32(int(n))) for n in U.shape[:4]]


def multiplymatmat_c(left, right):
    """MultiplyMatMat_c (c_wrapper/FLUE_c.f90:31-36): returns left @ right."""
    A, B = _c128(left, (3, 3)), _c128(right, (3, 3))
    MM = np.zeros((3, 3), np.complex128, order="F")
    check(_lib.lib().multiplymatmat_c(_p(MM), _p(A), _p(B)), "multiplymatmat_c")
    return MM


def get_qhat_c(coord, shape, a=None):
    """get_qhat_c (c_wrapper/FLUE_c.f90:49-56)."""
    c = np.ascontiguousarray(coord, dtype=np.float64)
    s = np.ascontiguousarray(shape, dtype=np.int32)
    q = np.zeros(4, np.float64)
    ap = ctypes.byref(ctypes.c_double(a)) if a is not None else None
    check(_lib.lib().get_qhat_c(_p(c), _p(s), _p(q), ap), "get_qhat_c")
    return q


def scalargluonprop_c(U, xi):
    """scalarGluonProp_c (c_wrapper/FLUE_c.f90:9-16): U of shape (NT,NS,NS,NS,4,3,3) with
    NT <= NS -> D of shape (NS,NS,NS,NS), computed on the GPU."""
    U = _field(U)
    NT, NS = U.shape[0], U.shape[1]
    if U.shape[2] != NS or U.shape[3] != NS:
        raise ValueError("U must have shape (NT,NS,NS,NS,4,3,3)")
    D = np.zeros((NS,) * 4, np.float64, order="F")
    check(_lib.lib().scalargluonprop_c(_ri(NS), _ri(NT), _p(U), ctypes.byref(ctypes.c_double(xi)), _p(D)),
          "scalargluonprop_c")
    return D


def calc_mom_space_scalard_c(U, mu_start, xi):
    """calc_mom_space_scalarD_c (c_wrapper/FLUE_c.f90:18-29): U is the Fourier-transformed
    field declared U(0:NT, 0:NS, 0:NS, 0:NS, 0:3, 0:2, 0:2), so NT and NS are one less than
    its extents -> D of shape (NT//2+1, NS//2+1, NS//2+1, NS//2+1), computed on the GPU."""
    U = _field(U)
    NT, NS = U.shape[0] - 1, U.shape[1] - 1
    if U.shape[2] != NS + 1 or U.shape[3] != NS + 1:
        raise ValueError("U must have shape (NT+1,NS+1,NS+1,NS+1,4,3,3)")
    D = np.zeros((NT // 2 + 1,) + (NS // 2 + 1,) * 3, np.complex128, order="F")
    check(_lib.lib().calc_mom_space_scalard_c(_ri(NS), _ri(NT), _p(U), _p(D), _ri(mu_start),
                                              ctypes.byref(ctypes.c_double(xi))), "calc_mom_space_scalard_c")
    return D


def q_average_c(NQ, q, D, q_slices):
    """Q_Average_c (c_wrapper/FLUE_c.f90:38-47) -> (DOUT, QOUT, Qcount), DOUT and QOUT of
    NQ + 1 values."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    D = np.ascontiguousarray(D, dtype=np.float64)
    n = int(NQ)
    if q.size < n or D.size < n:
        raise ValueError("q and D need NQ values")
    dout = np.zeros(max(n, 0) + 1)
    qout = np.zeros(max(n, 0) + 1)
    cnt = ctypes.c_int32(0)
    check(_lib.lib().q_average_c(_ri(n), _p(q), _p(D), _ri(q_slices), _p(dout), _p(qout), ctypes.byref(cnt)),
          "q_average_c")
    return dout, qout, cnt.value


def cone_cut_c(NQ, radius, Q, D, D4, NT, NS, angleIN, xiIN, IRCutIN, IRRadiusIN):
    """cone_cut_c (c_wrapper/FLUE_c.f90:58-71) -> (QOUT, Dout, D4out, qcount)."""
    n = int(NQ)
    Q = np.asfortranarray(Q, dtype=np.float64)
    D = np.ascontiguousarray(D, dtype=np.float64)
    D4 = np.ascontiguousarray(D4, dtype=np.float64)
    if Q.shape != (n, 4) or D.size < n or D4.size < n:
        raise ValueError("Q must be (NQ, 4) and D, D4 hold NQ values")
    qo = np.zeros((n, 4), np.float64, order="F")
    do = np.zeros(n)
    d4o = np.zeros(n)
    cnt = ctypes.c_int32(0)
    dbl = lambda v: ctypes.byref(ctypes.c_double(float(v)))  # noqa: E731
    check(_lib.lib().cone_cut_c(_ri(n), _ri(radius), _p(Q), _p(D), _p(D4), _ri(NT), _ri(NS), dbl(angleIN),
                                dbl(xiIN), dbl(IRCutIN), dbl(IRRadiusIN), _p(qo), _p(do), _p(d4o),
                                ctypes.byref(cnt)), "cone_cut_c")
    return qo, do, d4o, cnt.value


def periodcoord_c(coord, datashape):
    """periodCoord (src/wloops.F90:546-565)."""
    c = np.ascontiguousarray(coord, dtype=np.int32)
    d = np.ascontiguousarray(datashape, dtype=np.int32)
    if d.size < 4:
        raise ValueError("datashape needs at least the 4 lattice extents")
    d7 = np.zeros(7, np.int32)
    d7[:min(7, d.size)] = d[:7]
    out = np.zeros(4, np.int32)
    check(_lib.lib().periodcoord_c(_p(c), _p(d7), _p(out)), "periodcoord_c")
    return out


def genericpath_c(data, coordbase, path):
    """genericPath (src/wloops.F90:20-57): transposed ordered product, (3,3) complex."""
    U = _field(data)
    c = np.ascontiguousarray(coordbase, dtype=np.int32)
    p = np.ascontiguousarray(path, dtype=np.int32)
    out = np.zeros((3, 3), np.complex128, order="F")
    check(_lib.lib().genericpath_c(*_dims(U), _p(U), _p(c), _ri(p.size), _p(p), _p(out)), "genericpath_c")
    return out


def _plaq(fn, name, data, mustart, muend, nuend):
    U = _field(data)
    s = ctypes.c_double()
    n = ctypes.c_int32()
    t = ctypes.c_double()
    check(fn(*_dims(U), _p(U), _ri(mustart), _ri(muend), _ri(nuend), ctypes.byref(s), ctypes.byref(n),
             ctypes.byref(t)), name)
    return s.value, n.value, t.value


def genplaquette_c(data, mustart, muend, nuend):
    """genPlaquette (src/wloops.F90:59-127): (sumTrP, nP, time)."""
    return _plaq(_lib.lib().genplaquette_c, "genplaquette_c", data, mustart, muend, nuend)


def plaquette_c(data, mustart, muend, nuend):
    """plaquette (src/wloops.F90:915-1017): (sumTrP, nP, time)."""
    return _plaq(_lib.lib().plaquette_c, "plaquette_c", data, mustart, muend, nuend)


def polyakov_c(data):
    """polyakov (src/wloops.F90:1019-1069): (sumTrP, nP, time)."""
    U = _field(data)
    s = ctypes.c_double()
    n = ctypes.c_int32()
    t = ctypes.c_double()
    check(_lib.lib().polyakov_c(*_dims(U), _p(U), ctypes.byref(s), ctypes.byref(n), ctypes.byref(t)), "polyakov_c")
    return s.value, n.value, t.value


def _fplane(fn, name, data, mu, nu):
    U = _field(data)
    out = np.zeros(U.shape[:4] + (3, 3), np.complex128, order="F")
    check(fn(*_dims(U), _p(U), _ri(mu), _ri(nu), _p(out)), name)
    return out


def cloverloopmunucoord_c(data, mu, nu):
    """cloverLoopMuNuCoord (src/wloops.F90:129-199): (NT,NX,NY,NZ,3,3)."""
    return _fplane(_lib.lib().cloverloopmunucoord_c, "cloverloopmunucoord_c", data, mu, nu)


def loop5munucoord_c(data, mu, nu):
    """Loop5MuNuCoord (src/wloops.F90:201-432): (NT,NX,NY,NZ,3,3)."""
    return _fplane(_lib.lib().loop5munucoord_c, "loop5munucoord_c", data, mu, nu)


def magnetic_c(data):
    """magnetic (src/wloops.F90:479-544): B."""
    U = _field(data)
    B = ctypes.c_double()
    check(_lib.lib().magnetic_c(*_dims(U), _p(U), ctypes.byref(B)), "magnetic_c")
    return B.value


def wmunu_c(data, mu, nu, nmu, nnu):
    """wmunu (src/wloops.F90:679-791)."""
    U = _field(data)
    v = ctypes.c_double()
    check(_lib.lib().wmunu_c(*_dims(U), _p(U), _ri(mu), _ri(nu), _ri(nmu), _ri(nnu), ctypes.byref(v)), "wmunu_c")
    return v.value


def wmunutau_c(data, tau, mu, nu, nmu, nnu):
    """wmunutau (src/wloops.F90:567-677)."""
    U = _field(data)
    v = ctypes.c_double()
    check(_lib.lib().wmunutau_c(*_dims(U), _p(U), _ri(tau), _ri(mu), _ri(nu), _ri(nmu), _ri(nnu), ctypes.byref(v)),
          "wmunutau_c")
    return v.value
