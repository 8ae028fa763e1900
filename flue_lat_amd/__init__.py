"""flue_lat_amd -- MI355X implementation of FLUE's gluonic-observable path.

The module tree mirrors the reference's Python package (python/FLUE/__init__.py:1-4):
``gluonCalc`` (with the observables in ``gluonCalc.wloops``), ``gluonAna`` and
``compiled.flue_c``; ``gaugeio`` (openQCD / ILDG files, host side) and ``jack``
(jackknife statistics) follow src/gaugeIO/ and src/jack.f90.  A drop-in user writes
``import flue_lat_amd as FLUE``.
The compute path is libflue_hip.so (HIP kernels for gfx950, C ABI in
include/flue_c.h); nothing here falls back to the CPU.
"""
from . import compiled, gaugeio, gluonAna, gluonCalc, jack
from ._lib import FlueError, FlueNotBuilt

__version__ = "0.1.0"  # src/FLUE.f90:15


def __getattr__(name):
    # the device classes import torch; load them on first use
    if name == "GaugeField":
        from .field import GaugeField
        return GaugeField
    raise AttributeError(name)


__all__ = ["gluonAna", "gluonCalc", "compiled", "gaugeio", "jack", "GaugeField", "FlueError", "FlueNotBuilt"]
