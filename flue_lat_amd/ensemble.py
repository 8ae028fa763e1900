"""Ensembles of gauge fields with different extents, processed together on the GPU.

The fields lie back to back in one torch tensor, each in the SoA layout of its own
lattice; `Ensemble.field(c)` is a GaugeField view of configuration c.  Flowing the
ensemble runs one launch per RK3 stage for all configurations and, with
use_graph=True, the whole run (flow steps and the periodic plaquette / B
measurements) is captured in one hipGraph and launched once.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check
from .field import GaugeField, _prec, _ptr, _stream, require_gpu


class Ensemble:
    def __init__(self, dims_list, precision="fp32", device=None):
        require_gpu()
        self.dims = [tuple(int(v) for v in d) for d in dims_list]
        if not self.dims or any(len(d) != 4 or min(d) < 1 for d in self.dims):
            raise ValueError("dims_list must hold (NT, NX, NY, NZ) tuples of positive extents")
        self.n = len(self.dims)
        self.prec = _prec(precision)
        self.precision = "fp32" if self.prec == _lib.FLUE_F32 else "fp64"
        self.dtype = torch.float32 if self.prec == _lib.FLUE_F32 else torch.float64
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        flat = [v for d in self.dims for v in d]
        self._cdims = (ctypes.c_int32 * len(flat))(*flat)
        nbytes = _lib.lib().flue_ensemble_field_bytes(self.n, self._cdims, self.prec)
        if nbytes == 0:
            raise ValueError("invalid ensemble dimensions")
        self.data = torch.empty(nbytes // self.dtype.itemsize, dtype=self.dtype, device=self.device)
        self.volumes = [int(np.prod(d)) for d in self.dims]
        self.offsets = np.concatenate([[0], np.cumsum([72 * v for v in self.volumes])]).astype(np.int64)

    @classmethod
    def random(cls, dims_list, seeds, eps, precision="fp32", device=None):
        e = cls(dims_list, precision, device)
        if len(seeds) != e.n:
            raise ValueError("one seed per configuration")
        cs = (ctypes.c_uint64 * e.n)(*[int(s) for s in seeds])
        with torch.cuda.device(e.device):
            check(_lib.lib().flue_ensemble_random(e.n, e._cdims, cs, float(eps), _ptr(e.data), e.prec, _stream()),
                  "flue_ensemble_random")
        return e

    def field(self, c):
        """Configuration c as a GaugeField sharing the ensemble's storage."""
        lo, hi = int(self.offsets[c]), int(self.offsets[c + 1])
        return GaugeField(self.dims[c], self.precision, self.device, _data=self.data[lo:hi])

    def autotune(self, reps=3):
        """Time the variants of the ensemble flow on this ensemble (persistent grid of 1, 2,
        4 or 8 workgroups per CU, with the staple or the plaquette-first stencil; one flow
        stage of every field into scratch, the fields are only read) and remember the
        fastest for flow() on ensembles of this shape.  Returns (variant, ms per stage)."""
        best = ctypes.c_int32()
        ms = ctypes.c_double()
        nbytes = _lib.lib().flue_ensemble_work_bytes(self.n, self._cdims, self.prec)
        work = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check(_lib.lib().flue_ensemble_autotune(self.n, self._cdims, _ptr(self.data), self.prec, int(reps),
                                                    ctypes.byref(best), ctypes.byref(ms), _ptr(work), nbytes,
                                                    _stream()), "flue_ensemble_autotune")
        return best.value, ms.value

    def flow(self, eps, nsteps, measure_every=10, use_graph=True):
        """nsteps RK3 Wilson-flow steps of every configuration, in place.  Returns
        (plaquette sums, B) as arrays of shape (nsteps // measure_every + 1, n): the sum of
        ReTr P over all 6V plaquettes and the magnetic B, measured at steps
        0, measure_every, 2*measure_every, ..."""
        nmeas = nsteps // measure_every + 1 if measure_every > 0 else 0
        plaq = torch.zeros(max(nmeas, 1) * self.n, dtype=torch.float64, device=self.device)
        B = torch.zeros_like(plaq)
        nbytes = _lib.lib().flue_ensemble_work_bytes(self.n, self._cdims, self.prec)
        work = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            cur = torch.cuda.current_stream()
            side = torch.cuda.Stream(device=self.device) if (use_graph and cur.cuda_stream == 0) else cur
            if side is not cur:
                side.wait_stream(cur)
            with torch.cuda.stream(side):
                check(_lib.lib().flue_ensemble_flow(self.n, self._cdims, _ptr(self.data), self.prec, float(eps),
                                                    int(nsteps), int(measure_every), _ptr(plaq), _ptr(B),
                                                    1 if use_graph else 0, _ptr(work), nbytes, _stream()),
                      "flue_ensemble_flow")
            if side is not cur:
                cur.wait_stream(side)
        if nmeas == 0:
            return None, None
        return (plaq.cpu().numpy().reshape(nmeas, self.n), B.cpu().numpy().reshape(nmeas, self.n))
