"""ctypes front end of the CPU FP64 oracle (oracle/flue_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, as the checker.  The product package
(flue_lat_amd) never imports it.

Gauge fields are numpy complex128 arrays of shape (NT, NX, NY, NZ, 4, 3, 3) in
Fortran order, indexed U[t, x, y, z, mu, a, b] = (U_mu(x))_ab with mu = 0 the
time direction -- the array FLUE's Fortran routines take (src/wloops.F90:21).
Directions and coordinates passed to the path functions are the reference's
1-based ones.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libflue_oracle.so")
_lib = None


class OracleError(RuntimeError):
    pass


def build(force=False):
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "libflue_oracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_I = ctypes.c_int32
_D = ctypes.c_double


def _declare(L):
    sig = {
        "or_matmul": [_P, _P, _P],
        "or_matdagmatdag": [_P, _P, _P],
        "or_retrace_mult": [_P, _P],
        "or_period_coord": [_P, _P, _P],
        "or_generic_path": [_P, _P, _P, _P, _I, _P],
        "or_gen_plaquette": [_P, _P, _I, _I, _I, _P, _P],
        "or_plaquette": [_P, _P, _I, _I, _I, _P, _P],
        "or_polyakov": [_P, _P, _P, _P],
        "or_clover_munu": [_P, _P, _I, _I, _P],
        "or_loop5_munu": [_P, _P, _I, _I, _P],
        "or_loop5_munu_sites": [_P, _P, _I, _I, _P, ctypes.c_int64, _P],
        "or_magnetic": [_P, _P, _P],
        "or_magnetic_sum_tslices": [_P, _P, _I, _I, _P],
        "or_wmunu": [_P, _P, _I, _I, _I, _I, _P],
        "or_wmunutau": [_P, _P, _I, _I, _I, _I, _I, _P],
        "or_su3_exp": [_P, _P],
        "or_fix_su3": [_P],
        "or_gellmann": [_I, _P],
        "or_random_field": [_P, ctypes.c_uint64, _D, _P],
        "or_philox4x32": [_P, _P, _P],
        "or_wilson_action": [_P, _P, _D, _P],
        "or_wilson_grad": [_P, _P, _D, _P],
        "or_b2_grad": [_P, _P, _P],
        "or_fd_link": [_P, _P, _I, _D, _P, _I, _I, _D, _P],
        "or_wilson_flow": [_P, _P, _D, _I, _I],
        "or_mom_scalar_d": [_P, _P, _P, _I, _D, _P],
        "or_scalar_gluon_prop": [_P, _P, _D, _P],
        "or_q_average": [_I, _P, _P, _I, _P, _P, _P],
        "or_cone_cut": [_I, _P, _P, _P, _D, _D, _D, _D, _P, _P, _P, _P],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _D if name == "or_retrace_mult" else (None if name in (
            "or_matmul", "or_matdagmatdag", "or_period_coord", "or_su3_exp", "or_fix_su3",
            "or_gellmann", "or_philox4x32") else ctypes.c_int)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _field(U):
    U = np.asarray(U)
    if U.ndim != 7 or U.shape[4:] != (4, 3, 3):
        raise ValueError("gauge field must have shape (NT,NX,NY,NZ,4,3,3)")
    return np.asfortranarray(U, dtype=np.complex128)


def _dims(U):
    return np.array(U.shape[:4], dtype=np.int32)


def _mat(M):
    return np.asfortranarray(np.asarray(M, dtype=np.complex128).reshape(3, 3))


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with code {rc}")


def matmul(A, B):
    out = np.zeros((3, 3), np.complex128, order="F")
    A, B = _mat(A), _mat(B)
    lib().or_matmul(_ptr(out), _ptr(A), _ptr(B))
    return out


def matdagmatdag(A, B):
    out = np.zeros((3, 3), np.complex128, order="F")
    A, B = _mat(A), _mat(B)
    lib().or_matdagmatdag(_ptr(out), _ptr(A), _ptr(B))
    return out


def retrace_mult(A, B):
    A, B = _mat(A), _mat(B)
    return lib().or_retrace_mult(_ptr(A), _ptr(B))


def period_coord(coord, dims):
    c = np.asarray(coord, np.int32)
    d = np.asarray(dims, np.int32)[:4].copy()
    out = np.zeros(4, np.int32)
    lib().or_period_coord(_ptr(c), _ptr(d), _ptr(out))
    return out


def generic_path(U, coord, path):
    U = _field(U)
    c = np.asarray(coord, np.int32)
    p = np.asarray(path, np.int32)
    out = np.zeros((3, 3), np.complex128, order="F")
    _check(lib().or_generic_path(_ptr(U), _ptr(_dims(U)), _ptr(c), _ptr(p), len(p), _ptr(out)), "generic_path")
    return out


def _sum_np(fn, U, *args):
    U = _field(U)
    s = np.zeros(1, np.float64)
    n = np.zeros(1, np.int32)
    _check(fn(_ptr(U), _ptr(_dims(U)), *args, _ptr(s), _ptr(n)), fn.__name__)
    return float(s[0]), int(n[0])


def gen_plaquette(U, muStart, muEnd, nuEnd):
    return _sum_np(lib().or_gen_plaquette, U, muStart, muEnd, nuEnd)


def plaquette(U, muStart, muEnd, nuEnd):
    return _sum_np(lib().or_plaquette, U, muStart, muEnd, nuEnd)


def polyakov(U):
    return _sum_np(lib().or_polyakov, U)


def _plane_field(fn, U, mu, nu):
    U = _field(U)
    out = np.zeros(U.shape[:4] + (3, 3), np.complex128, order="F")
    _check(fn(_ptr(U), _ptr(_dims(U)), mu, nu, _ptr(out)), fn.__name__)
    return out


def clover_munu(U, mu, nu):
    return _plane_field(lib().or_clover_munu, U, mu, nu)


def loop5_munu(U, mu, nu):
    return _plane_field(lib().or_loop5_munu, U, mu, nu)


def loop5_munu_sites(U, mu, nu, sites):
    """F at the listed 0-based Fortran site offsets; returns (n, 3, 3)."""
    U = _field(U)
    s = np.ascontiguousarray(sites, dtype=np.int64)
    out = np.zeros((len(s), 3, 3), np.complex128)
    buf = np.zeros(9 * len(s), np.complex128)
    _check(lib().or_loop5_munu_sites(_ptr(U), _ptr(_dims(U)), mu, nu, _ptr(s), len(s), _ptr(buf)), "loop5_sites")
    for i in range(len(s)):
        out[i] = buf[9 * i:9 * i + 9].reshape(3, 3, order="F")
    return out


def magnetic(U):
    U = _field(U)
    B = np.zeros(1, np.float64)
    _check(lib().or_magnetic(_ptr(U), _ptr(_dims(U)), _ptr(B)), "magnetic")
    return float(B[0])


def magnetic_sum_tslices(U, t0, t1):
    """Unnormalised sum of -ReTr F_i^2 over time slices t0..t1 (1-based, inclusive)."""
    U = _field(U)
    s = np.zeros(1, np.float64)
    _check(lib().or_magnetic_sum_tslices(_ptr(U), _ptr(_dims(U)), t0, t1, _ptr(s)), "magnetic_sum_tslices")
    return float(s[0])


def wmunu(U, mu, nu, nMu, nNu):
    U = _field(U)
    v = np.zeros(1, np.float64)
    _check(lib().or_wmunu(_ptr(U), _ptr(_dims(U)), mu, nu, nMu, nNu, _ptr(v)), "wmunu")
    return float(v[0])


def wmunutau(U, tau, mu, nu, nMu, nNu):
    U = _field(U)
    v = np.zeros(1, np.float64)
    _check(lib().or_wmunutau(_ptr(U), _ptr(_dims(U)), tau, mu, nu, nMu, nNu, _ptr(v)), "wmunutau")
    return float(v[0])


def su3_exp(X):
    X = _mat(X)
    out = np.zeros((3, 3), np.complex128, order="F")
    lib().or_su3_exp(_ptr(X), _ptr(out))
    return out


def fix_su3(M):
    M = _mat(M).copy(order="F")
    lib().or_fix_su3(_ptr(M))
    return M


def gellmann(a):
    out = np.zeros((3, 3), np.complex128, order="F")
    lib().or_gellmann(a, _ptr(out))
    return out


def philox4x32(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32(_ptr(c), _ptr(k), _ptr(out))
    return out


def random_field(dims, seed, eps):
    d = np.asarray(dims, np.int32)
    out = np.zeros(tuple(int(x) for x in d) + (4, 3, 3), np.complex128, order="F")
    _check(lib().or_random_field(_ptr(d), ctypes.c_uint64(seed), eps, _ptr(out)), "random_field")
    return out


def wilson_action(U, beta):
    U = _field(U)
    s = np.zeros(1, np.float64)
    _check(lib().or_wilson_action(_ptr(U), _ptr(_dims(U)), beta, _ptr(s)), "wilson_action")
    return float(s[0])


def wilson_grad(U, beta):
    U = _field(U)
    g = np.zeros(U.shape, np.complex128, order="F")
    _check(lib().or_wilson_grad(_ptr(U), _ptr(_dims(U)), beta, _ptr(g)), "wilson_grad")
    return g


def b2_grad(U):
    U = _field(U)
    g = np.zeros(U.shape, np.complex128, order="F")
    _check(lib().or_b2_grad(_ptr(U), _ptr(_dims(U)), _ptr(g)), "b2_grad")
    return g


def fd_link(U, which, coord, mu, a, h=1e-6, beta=6.0):
    """Central difference of the Wilson action (which=0) or B (which=1) along exp(+-h T^a) on U_mu(coord).
    A Fortran-ordered complex128 U is perturbed in place and restored; anything else is copied."""
    if not (isinstance(U, np.ndarray) and U.dtype == np.complex128 and U.flags.f_contiguous and U.flags.writeable):
        U = _field(U).copy(order="F")
    c = np.asarray(coord, np.int32)
    d = np.zeros(1, np.float64)
    _check(lib().or_fd_link(_ptr(U), _ptr(_dims(U)), which, beta, _ptr(c), mu, a, h, _ptr(d)), "fd_link")
    return float(d[0])


def wilson_flow(U, eps, nsteps, reunit=True):
    U = _field(U).copy(order="F")
    _check(lib().or_wilson_flow(_ptr(U), _ptr(_dims(U)), eps, nsteps, 1 if reunit else 0), "wilson_flow")
    return U


T_GEN = None


def generators():
    """T^a = -i lambda_a / 2, a = 1..8, as a (8,3,3) array."""
    global T_GEN
    if T_GEN is None:
        T_GEN = np.stack([-0.5j * np.asarray(gellmann(a)) for a in range(1, 9)])
    return T_GEN


def mom_scalar_d(Uq, mu_start, xi, dext=None):
    """calc_mom_space_scalarD on a Fourier-transformed field Uq of shape (Et,Ex,Ey,Ez,4,3,3);
    D over 0 <= q < dext (default E // 2 + 1 per direction)."""
    Uq = _field(Uq)
    E = _dims(Uq)
    dext = np.asarray(E // 2 + 1 if dext is None else dext, np.int32)
    D = np.zeros(tuple(int(v) for v in dext), np.complex128, order="F")
    _check(lib().or_mom_scalar_d(_ptr(Uq), _ptr(E), _ptr(dext), mu_start, xi, _ptr(D)), "mom_scalar_d")
    return D


def scalar_gluon_prop(U, xi):
    """scalarGluonProp on a position-space field with NT <= NS; D of shape (NS,NS,NS,NS)."""
    U = _field(U)
    NS = U.shape[1]
    D = np.zeros((NS,) * 4, np.float64, order="F")
    _check(lib().or_scalar_gluon_prop(_ptr(U), _ptr(_dims(U)), xi, _ptr(D)), "scalar_gluon_prop")
    return D


def q_average(q, D, q_slices):
    """Q_Average -> (DOUT, QOUT, Qcount) with NQ + 1 entries in DOUT, QOUT."""
    q = np.ascontiguousarray(q, np.float64)
    D = np.ascontiguousarray(D, np.float64)
    n = q.shape[0]
    dout = np.zeros(n + 1)
    qout = np.zeros(n + 1)
    cnt = np.zeros(1, np.int32)
    _check(lib().or_q_average(n, _ptr(q), _ptr(D), q_slices, _ptr(dout), _ptr(qout), _ptr(cnt)), "q_average")
    return dout, qout, int(cnt[0])


def cone_cut(Q, D, D4, angle, xi, IRCut, IRRadius):
    """cone_cut with every optional argument given -> (QOUT, DOUT, D4OUT, qcount)."""
    Q = np.asfortranarray(Q, np.float64)
    D = np.ascontiguousarray(D, np.float64)
    D4 = np.ascontiguousarray(D4, np.float64)
    n = Q.shape[0]
    qo = np.zeros((n, 4), np.float64, order="F")
    do = np.zeros(n)
    d4o = np.zeros(n)
    cnt = np.zeros(1, np.int32)
    _check(lib().or_cone_cut(n, _ptr(Q), _ptr(D), _ptr(D4), angle, xi, IRCut, IRRadius, _ptr(qo), _ptr(do),
                             _ptr(d4o), _ptr(cnt)), "cone_cut")
    return qo, do, d4o, int(cnt[0])
