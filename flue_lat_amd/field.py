"""Device-resident gauge fields and the device entry points of libflue_hip.so.

A :class:`GaugeField` owns one torch tensor holding the library's SoA layout (36
complex planes, direction mu and colour a, b, of V sites with adjacent real and
imaginary parts; site index x + NX*(y + NY*(z + NZ*t))) in FP64 or FP32.  Its methods call the C ABI on the
current torch stream and keep the results on the device until asked for.

Directions and coordinates follow the reference (src/wloops.F90): mu = 1 is time,
2..4 are x, y, z; coordinates are 1-based (t, x, y, z).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import FLUE_C64, FLUE_C128, FLUE_F32, FLUE_F64, FLUE_F64_SU3, FLUE_OUT_SOA, FLUE_OUT_STRIDED
from ._lib import check as check_status

# "fp64-su3": FP64 links the caller declares SU(3); the magnetic measurement may then carry
# two rows of each product, as FP32 always does
_PREC = {"fp64": FLUE_F64, "f64": FLUE_F64, "double": FLUE_F64, torch.float64: FLUE_F64,
         "fp32": FLUE_F32, "f32": FLUE_F32, "float": FLUE_F32, torch.float32: FLUE_F32,
         "fp64-su3": FLUE_F64_SU3}
_PREC_NAME = {FLUE_F64: "fp64", FLUE_F32: "fp32", FLUE_F64_SU3: "fp64-su3"}


def _prec(p):
    try:
        return _PREC[p]
    except KeyError:
        raise ValueError(f"precision must be 'fp64' or 'fp32', got {p!r}") from None


def _i32(*vals):
    return (ctypes.c_int32 * len(vals))(*vals)


def _i64(vals):
    return (ctypes.c_int64 * len(vals))(*vals)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("flue_lat_amd device path needs a visible GPU")


def tune_save(path):
    """Write the autotuner's table (fastest kernel variant per device, op, precision, extents)."""
    check_status(_lib.lib().flue_tune_save(str(path).encode()), "flue_tune_save")


def tune_load(path):
    """Merge a table written by tune_save into the autotuner."""
    check_status(_lib.lib().flue_tune_load(str(path).encode()), "flue_tune_load")


class GaugeField:
    """SU(3) gauge field on the GPU in the SoA layout.

    dims = (NT, NX, NY, NZ) as in the reference's data(NT,NX,NY,NZ,4,3,3).
    """

    def __init__(self, dims, precision="fp64", device=None, _data=None, check=True):
        """A field of the given extents; with _data, a view of existing links in the SoA
        layout.  Declaring existing links "fp64-su3" runs check_su3() on them first (raises
        FlueError FLUE_ESU3 on links that are not SU(3)); check=False skips that."""
        require_gpu()
        declared = _data is not None   # existing links declared in this precision
        self.dims = tuple(int(d) for d in dims)
        if len(self.dims) != 4 or min(self.dims) < 1:
            raise ValueError(f"dims must be 4 positive extents (NT, NX, NY, NZ), got {dims}")
        self.prec = _prec(precision)
        self.dtype = torch.float32 if self.prec == FLUE_F32 else torch.float64
        self.V = int(np.prod(self.dims))
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if _data is None:
            _data = torch.empty(72 * self.V, dtype=self.dtype, device=self.device)
        self.data = _data
        self._cdims = _i32(*self.dims)
        self._ws = None
        self._mws = None
        if check and self.prec == FLUE_F64_SU3 and declared:
            self.check_su3()

    # ------------------------------------------------------------ construction
    @property
    def precision(self):
        return _PREC_NAME[self.prec]

    def workspace(self):
        if self._ws is None:
            n = _lib.lib().flue_workspace_bytes(self._cdims)
            self._ws = torch.empty(n, dtype=torch.uint8, device=self.device)
        return self._ws

    def magnetic_workspace(self):
        """The workspace of magnetic(): every variant's, including the rectangle-first
        form's footprints (flue_magnetic_work_bytes); kept with the field."""
        if self._mws is None:
            n = _lib.lib().flue_magnetic_work_bytes(self._cdims, self.prec)
            self._mws = torch.empty(n, dtype=torch.uint8, device=self.device)
        return self._mws

    @classmethod
    def random(cls, dims, seed, eps, precision="fp64", device=None):
        """Seeded synthetic field U = fix(exp(i eps sum_a c_a lambda_a/2)), generated on the GPU."""
        f = cls(dims, precision, device)
        with torch.cuda.device(f.device):
            check_status(_lib.lib().flue_random_field(f._cdims, ctypes.c_uint64(seed), float(eps), _ptr(f.data), f.prec,
                                               _stream()), "flue_random_field")
        return f

    @classmethod
    def from_complex(cls, U, precision="fp64", device=None, check=True):
        """From a (NT,NX,NY,NZ,4,3,3) complex array: numpy (any order) or a torch tensor
        (any strides, complex128 or complex64, CPU or GPU).  For "fp32" and "fp64-su3" the
        imported links are checked with check_su3() unless check=False."""
        if isinstance(U, np.ndarray):
            U = torch.from_numpy(np.asarray(U))
        if U.dim() != 7 or tuple(U.shape[4:]) != (4, 3, 3):
            raise ValueError("gauge field must have shape (NT,NX,NY,NZ,4,3,3)")
        if U.dtype not in (torch.complex128, torch.complex64):
            U = U.to(torch.complex128)
        f = cls(U.shape[:4], precision, device)
        src = U.to(f.device) if U.device != f.device else U
        with torch.cuda.device(f.device):
            src_type = FLUE_C64 if src.dtype == torch.complex64 else FLUE_C128
            check_status(_lib.lib().flue_import_field(f._cdims, _ptr(src), src_type, _i64(src.stride()), _ptr(f.data), f.prec,
                                               _stream()), "flue_import_field")
            if src is not U:
                torch.cuda.current_stream().synchronize()
        if check and f.prec in (FLUE_F32, FLUE_F64_SU3):
            f.check_su3()
        return f

    def to_complex(self, dtype=torch.complex128):
        """Export as a (NT,NX,NY,NZ,4,3,3) complex tensor with the reference's Fortran strides."""
        NT, NX, NY, NZ = self.dims
        buf = torch.empty((3, 3, 4, NZ, NY, NX, NT), dtype=dtype, device=self.device)
        out = buf.permute(6, 5, 4, 3, 2, 1, 0)
        with torch.cuda.device(self.device):
            dst_type = FLUE_C64 if dtype == torch.complex64 else FLUE_C128
            check_status(_lib.lib().flue_export_field(self._cdims, _ptr(self.data), self.prec, _ptr(out), dst_type,
                                               _i64(out.stride()), _stream()), "flue_export_field")
        return out

    def astype(self, precision, check=True):
        """The field in another precision.  FP32 links and "fp64-su3" links must be SU(3)
        (their kernels rebuild row 2 of link products from rows 0 and 1): the result is
        checked with check_su3() and FlueError (FLUE_ESU3) is raised on failure, unless
        check=False."""
        p = _prec(precision)
        g = GaugeField(self.dims, _PREC_NAME[p], self.device)
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_convert_field(self._cdims, _ptr(self.data), self.prec, _ptr(g.data), g.prec,
                                                       _stream()), "flue_convert_field")
        if check and p in (FLUE_F32, FLUE_F64_SU3):
            g.check_su3()
        return g

    def check_su3(self, tol=0.0, raise_on_fail=True):
        """Deterministic scan of every link in FP64: returns {"nonfinite": links with a NaN or
        infinite entry, "unitarity": max |(U U^+ - 1)_ab|, "det": max |det U - 1|}.  Raises
        FlueError (FLUE_ESU3) when a link is non-finite or a deviation exceeds tol (0 = the
        precision's bound: 1e-5 for FP32 links, 1e-10 for FP64) unless raise_on_fail=False."""
        stats = (ctypes.c_double * 3)()
        ws = self.workspace()
        with torch.cuda.device(self.device):
            rc = _lib.lib().flue_check_field(self._cdims, _ptr(self.data), self.prec, float(tol), stats, _ptr(ws),
                                             ws.numel(), _stream())
        if rc != _lib.FLUE_ESU3 or raise_on_fail:
            check_status(rc, "flue_check_field")
        return {"nonfinite": int(stats[0]), "unitarity": stats[1], "det": stats[2]}

    # ------------------------------------------------------------ observables
    def _scalar(self):
        return torch.empty(1, dtype=torch.float64, device=self.device)

    def plaquette(self, muStart=1, muEnd=4, nuEnd=4, sync=True):
        """(sum of ReTr P, number of plaquettes) as in genPlaquette / plaquette (src/wloops.F90:59-127, 915-1017)."""
        out = self._scalar()
        ws = self.workspace()
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_plaquette(self._cdims, _ptr(self.data), self.prec, muStart, muEnd, nuEnd, _ptr(out),
                                            _ptr(ws), ws.numel(), _stream()), "flue_plaquette")
        nP = sum(max(0, nuEnd - mu) for mu in range(muStart, muEnd + 1)) * self.V
        return (out.item() if sync else out), nP

    def polyakov(self, sync=True):
        """(sum_xyz ReTr prod_t U_t, NX*NY*NZ) as in polyakov (src/wloops.F90:1019-1069)."""
        out = self._scalar()
        ws = self.workspace()
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_polyakov(self._cdims, _ptr(self.data), self.prec, _ptr(out), _ptr(ws), ws.numel(),
                                           _stream()), "flue_polyakov")
        return (out.item() if sync else out), self.V // self.dims[0]

    def path_trace(self, path, tau=0, sync=True):
        """Sum of ReTr genericPath over all sites (tau=0) or over time slice tau (1-based)."""
        p = _i32(*[int(v) for v in path])
        out = self._scalar()
        ws = self.workspace()
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_path_trace(self._cdims, _ptr(self.data), self.prec, p, len(path), int(tau), _ptr(out),
                                             _ptr(ws), ws.numel(), _stream()), "flue_path_trace")
        return out.item() if sync else out

    def wilson_loop(self, mu, nu, nMu, nNu, tau=0):
        """wmunu (tau=0) or wmunutau: mean ReTr of the nMu x nNu loop (src/wloops.F90:567-791)."""
        path = [mu] * nMu + [nu] * nNu + [-mu] * nMu + [-nu] * nNu
        s = self.path_trace(path, tau)
        return s / (self.V // self.dims[0] if tau else self.V)

    def wilson_loops(self, mu, nu, Rmax=6, Tmax=6, tau=0):
        """Table W[R-1, T-1] of wmunu(mu, nu, R, T) (tau = 0) or wmunutau(tau, mu, nu, R, T) for
        1 <= R <= Rmax, 1 <= T <= Tmax, computed in one pass (src/wloops.F90:567-791)."""
        out = torch.empty(int(Rmax) * int(Tmax), dtype=torch.float64, device=self.device)
        ws = self.workspace()
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_wilson_loops(self._cdims, _ptr(self.data), self.prec, int(mu), int(nu), int(Rmax),
                                               int(Tmax), int(tau), _ptr(out), _ptr(ws), ws.numel(), _stream()),
                  "flue_wilson_loops")
        return out.cpu().numpy().reshape(int(Rmax), int(Tmax))

    def wilson_loops_slices(self, mu, nu, Rmax=6, Tmax=6):
        """W[R-1, T-1, t-1] = wmunutau(t, mu, nu, R, T) for every time slice t, in one pass
        (the static-potential table W(R, T, t); src/wloops.F90:567-677)."""
        NT = self.dims[0]
        out = torch.empty(int(Rmax) * int(Tmax) * NT, dtype=torch.float64, device=self.device)
        ws = self.workspace()
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_wilson_loops_slices(self._cdims, _ptr(self.data), self.prec, int(mu), int(nu),
                                                      int(Rmax), int(Tmax), _ptr(out), _ptr(ws), ws.numel(),
                                                      _stream()), "flue_wilson_loops_slices")
        return out.cpu().numpy().reshape(int(Rmax), int(Tmax), NT)

    def generic_path(self, coordBase, path):
        """genericPath at one 1-based base point: the transposed ordered product (src/wloops.F90:20-57)."""
        out = torch.empty(18, dtype=torch.float64, device=self.device)
        p = _i32(*[int(v) for v in path])
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_generic_path(self._cdims, _ptr(self.data), self.prec, _i32(*coordBase), p, len(path),
                                               _ptr(out), _stream()), "flue_generic_path")
        v = out.cpu().numpy()
        return (v[0::2] + 1j * v[1::2]).reshape(3, 3, order="F")

    def fmunu(self, mu, nu, clover=False):
        """Per-site F for plane (mu, nu): Loop5MuNuCoord, or cloverLoopMuNuCoord with clover=True.
        Returns a complex128 tensor of shape (NT,NX,NY,NZ,3,3) with Fortran strides."""
        NT, NX, NY, NZ = self.dims
        buf = torch.empty((3, 3, NZ, NY, NX, NT), dtype=torch.complex128, device=self.device)
        out = buf.permute(5, 4, 3, 2, 1, 0)
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_fmunu(self._cdims, _ptr(self.data), self.prec, int(mu), int(nu), 1 if clover else 0,
                                        _ptr(out), FLUE_OUT_STRIDED, _i64(out.stride()), _stream()), "flue_fmunu")
        return out

    def momentum_field(self):
        """U_mu(q) = sum_x exp(-i q.x) U_mu(x) for every link element (numpy.fft.fftn's
        convention), as a complex tensor of shape (36, NT, NZ, NY, NX): plane (mu, a, b) first,
        then the momenta in the site order's axes (its strides within a plane may be permuted,
        as the transform leaves them)."""
        NT, NX, NY, NZ = self.dims
        c = torch.view_as_complex(self.data.view(36, NT, NZ, NY, NX, 2))
        return torch.fft.fftn(c, dim=(1, 2, 3, 4))

    def gluon_propagator(self, mu_start=1, xi=1.0):
        """calc_mom_space_scalarD (src/gluonCalc/gluonProp.f90:112-250) of this field: the
        4-D transform on the device, then the per-momentum colour algebra in one kernel, which
        reads the transform in the layout it comes in.  Returns D as a complex128 tensor of
        shape (NT//2+1, NX//2+1, NY//2+1, NZ//2+1), the reference's loop range over q."""
        Uq = self.momentum_field()
        if Uq.stride(0) != self.V:
            Uq = Uq.contiguous()
        st = Uq.stride()
        qstrides = _i64((st[1], st[4], st[3], st[2]))  # (t, x, y, z)
        dext = tuple(d // 2 + 1 for d in self.dims)
        D = torch.empty(dext, dtype=torch.complex128, device=self.device)
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_gluon_prop_mom(self._cdims, _ptr(Uq), self.prec, qstrides, int(mu_start),
                                                 float(xi), _i32(*dext), _ptr(D), _i64(D.stride()), _stream()),
                  "flue_gluon_prop_mom")
        return D

    def scalar_gluon_propagator(self):
        """scalarGluonProp (src/gluonCalc/gluonProp.f90:17-110) of this field (NT <= NX = NY =
        NZ): D(t, x, y, z) as a float64 tensor of shape (NS, NS, NS, NS), 0 for t >= NT."""
        NS = self.dims[1]
        buf = torch.empty(NS ** 4, dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_scalar_gluon_prop(self._cdims, _ptr(self.data), self.prec, _ptr(buf), _stream()),
                  "flue_scalar_gluon_prop")
        # Fortran order: t fastest
        return buf.view(NS, NS, NS, NS).permute(3, 2, 1, 0)

    def magnetic(self, sync=True, variant=0):
        """B = -(1/V) sum_x sum_i ReTr F_i^2 over the spatial planes (src/wloops.F90:479-544).
        variant selects the kernel form and occupancy target: 1-3 site-major, 4-6
        plane-major, at 1-3 waves per SIMD; 7 rectangle-first (NX, NY, NZ divisible by 16;
        the default form otherwise) (0 = tuned or default choice)."""
        out = self._scalar()
        ws = self.magnetic_workspace()
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_magnetic_ex(self._cdims, _ptr(self.data), self.prec, int(variant), _ptr(out),
                                              _ptr(ws), ws.numel(), _stream()), "flue_magnetic_ex")
        return out.item() if sync else out

    def autotune(self, op="magnetic", reps=5):
        """Time the variants of `op` on this field and remember the fastest for its (device,
        precision, extents); later calls of that operation use it.  op: "magnetic" (kernel
        form x waves per SIMD), "b2_grad" (gather or reverse-mode adjoint), "plaquette" (time
        slices per thread, also used by wilson_action), "flow" (stencil visit order, used by
        wilson_flow and wilson_grad).  The field is only read.  Returns (variant, ms per call)."""
        if op not in _lib.OPS or op == "ensemble_flow":
            raise ValueError(f"unknown operation {op!r}")
        code = _lib.OPS[op]
        best = ctypes.c_int32()
        ms = ctypes.c_double()
        n = _lib.lib().flue_autotune_work_bytes(code, self._cdims, self.prec)
        ws = torch.empty(max(n, 1), dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_autotune(code, self._cdims, _ptr(self.data), self.prec, int(reps),
                                                  ctypes.byref(best), ctypes.byref(ms), _ptr(ws), n, _stream()),
                         "flue_autotune")
        return best.value, ms.value

    # ------------------------------------------------------------ gradients and flow
    def _like(self):
        return GaugeField(self.dims, self.precision, self.device)

    def wilson_action(self, beta, sync=True):
        """S = beta * sum_p (1 - ReTr P / 3) over all 6V plaquettes."""
        out = self._scalar()
        ws = self.workspace()
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_wilson_action(self._cdims, _ptr(self.data), self.prec, float(beta), _ptr(out),
                                                _ptr(ws), ws.numel(), _stream()), "flue_wilson_action")
        return out.item() if sync else out

    def wilson_grad(self, beta, variant=0):
        """Lie gradient of the Wilson action, G = (beta/6) P_TA(U Sigma), as a GaugeField-shaped
        su(3) field (same SoA layout and precision).  variant: the stencil (0 = tuned; 1..6
        staple sums in the autotuner's visit orders; 7..13 plaquette-first in storage order,
        walking along t, or in the brick orders of 2..6, used for FP32 and "fp64-su3" fields,
        whose links are SU(3))."""
        g = self._like()
        n = _lib.lib().flue_pf_work_bytes(self._cdims, self.prec)
        ws = torch.empty(n, dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_wilson_grad_ex(self._cdims, _ptr(self.data), self.prec, float(beta),
                                                        int(variant), _ptr(g.data), _ptr(ws), n, _stream()),
                         "flue_wilson_grad_ex")
        return g

    def magnetic_grad(self, variant=0):
        """Lie gradient of the magnetic B (temporal links get 0).  Needs SU(3) links in every
        precision: the library checks them first and FlueError (FLUE_ESU3) is raised for a
        field that fails the check.  variant: 1 = per-link gather, 2 = reverse mode,
        0 = tuned or reverse mode."""
        g = self._like()
        n = _lib.lib().flue_b2_work_bytes(self._cdims, self.prec)
        work = torch.empty(n, dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            check_status(_lib.lib().flue_b2_grad_ex(self._cdims, _ptr(self.data), self.prec, int(variant),
                                                    _ptr(g.data), _ptr(work), n, _stream()), "flue_b2_grad_ex")
        return g

    def wilson_flow(self, eps, nsteps, use_graph=True):
        """nsteps Luescher RK3 Wilson-flow steps of size eps, in place.  With use_graph the
        steps are replayed from a captured hipGraph."""
        n = _lib.lib().flue_flow_work_bytes(self._cdims, self.prec)
        work = torch.empty(n, dtype=torch.uint8, device=self.device)
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream()
            if use_graph and stream.cuda_stream == 0:
                side = torch.cuda.Stream(device=self.device)
                side.wait_stream(stream)
                with torch.cuda.stream(side):
                    check_status(_lib.lib().flue_wilson_flow(self._cdims, _ptr(self.data), self.prec, float(eps), int(nsteps),
                                                      1, _ptr(work), n, _stream()), "flue_wilson_flow")
                stream.wait_stream(side)
            else:
                check_status(_lib.lib().flue_wilson_flow(self._cdims, _ptr(self.data), self.prec, float(eps), int(nsteps),
                                                  1 if use_graph else 0, _ptr(work), n, _stream()), "flue_wilson_flow")
        return self

    def clone(self):
        g = self._like()
        g.data.copy_(self.data)
        return g
