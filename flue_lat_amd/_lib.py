"""Loader for libflue_hip.so (the C ABI in include/flue_c.h).

The library is built in-tree by ``python -m flue_lat_amd.build``; importing an entry
point without it raises immediately -- there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FLUE_HIP_LIB may point at another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("FLUE_HIP_LIB") or os.path.join(HERE, "libflue_hip.so")

FLUE_F64, FLUE_F32, FLUE_F64_SU3 = 0, 1, 2
FLUE_C128, FLUE_C64 = 0, 1
FLUE_OUT_STRIDED, FLUE_OUT_SOA = 0, 1

STATUS = {
    0: "FLUE_OK", 1: "FLUE_EINVAL", 2: "FLUE_ENULL", 3: "FLUE_EPATH", 4: "FLUE_ENT", 5: "FLUE_EDIM",
    6: "FLUE_EHIP", 7: "FLUE_ENOMEM", 8: "FLUE_ENOTIMPL", 9: "FLUE_ENODEV", 10: "FLUE_EWS", 11: "FLUE_ESU3",
}
FLUE_ESU3 = 11
# tunable operations (flue_autotune)
OPS = {"magnetic": 0, "b2_grad": 1, "plaquette": 2, "flow": 3, "ensemble_flow": 4}


class FlueError(RuntimeError):
    """Nonzero status from libflue_hip.so."""

    def __init__(self, status, what):
        self.status = status
        msg = _lib.flue_strerror(status).decode() if _lib is not None else STATUS.get(status, "?")
        super().__init__(f"{what}: {STATUS.get(status, status)} ({msg})")


class FlueNotBuilt(ImportError):
    pass


_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int32
_D = ctypes.c_double
_SZ = ctypes.c_size_t

# name -> argtypes; every function returns int status unless listed in _RESTYPE
SIGNATURES = {
    "flue_strerror": [ctypes.c_int],
    "flue_version": [_P, _P, _P],
    "multiplymatmat_c": [_P, _P, _P],
    "get_qhat_c": [_P, _P, _P, _P],
    "scalargluonprop_c": [_P, _P, _P, _P, _P],
    "calc_mom_space_scalard_c": [_P, _P, _P, _P, _P, _P],
    "q_average_c": [_P, _P, _P, _P, _P, _P, _P],
    "cone_cut_c": [_P] * 15,
    "periodcoord_c": [_P, _P, _P],
    "genericpath_c": [_P, _P, _P, _P, _P, _P, _P, _P, _P],
    "genplaquette_c": [_P] * 11,
    "plaquette_c": [_P] * 11,
    "polyakov_c": [_P] * 8,
    "cloverloopmunucoord_c": [_P] * 8,
    "loop5munucoord_c": [_P] * 8,
    "magnetic_c": [_P] * 6,
    "wmunu_c": [_P] * 10,
    "wmunutau_c": [_P] * 11,
    "flue_field_bytes": [_P, _I],
    "flue_workspace_bytes": [_P],
    "flue_import_field": [_P, _P, _I, _P, _P, _I, _P],
    "flue_export_field": [_P, _P, _I, _P, _I, _P, _P],
    "flue_convert_field": [_P, _P, _I, _P, _I, _P],
    "flue_random_field": [_P, ctypes.c_uint64, _D, _P, _I, _P],
    "flue_check_field": [_P, _P, _I, _D, _P, _P, _SZ, _P],
    "flue_plaquette": [_P, _P, _I, _I, _I, _I, _P, _P, _SZ, _P],
    "flue_polyakov": [_P, _P, _I, _P, _P, _SZ, _P],
    "flue_path_trace": [_P, _P, _I, _P, _I, _I, _P, _P, _SZ, _P],
    "flue_generic_path": [_P, _P, _I, _P, _P, _I, _P, _P],
    "flue_wilson_loops": [_P, _P, _I, _I, _I, _I, _I, _I, _P, _P, _SZ, _P],
    "flue_wilson_loops_slices": [_P, _P, _I, _I, _I, _I, _I, _P, _P, _SZ, _P],
    "flue_fmunu": [_P, _P, _I, _I, _I, _I, _P, _I, _P, _P],
    "flue_gluon_prop_mom": [_P, _P, _I, _P, _I, _D, _P, _P, _P, _P],
    "flue_scalar_gluon_prop": [_P, _P, _I, _P, _P],
    "flue_magnetic": [_P, _P, _I, _P, _P, _SZ, _P],
    "flue_magnetic_ex": [_P, _P, _I, _I, _P, _P, _SZ, _P],
    "flue_magnetic_noalloc": [_P, _P, _I, _I, _P, _P],
    "flue_b2_work_bytes": [_P, _I],
    "flue_flow_work_bytes": [_P, _I],
    "flue_wilson_action": [_P, _P, _I, _D, _P, _P, _SZ, _P],
    "flue_wilson_grad": [_P, _P, _I, _D, _P, _P],
    "flue_wilson_grad_ex": [_P, _P, _I, _D, _I, _P, _P, _SZ, _P],
    "flue_pf_work_bytes": [_P, _I],
    "flue_magnetic_work_bytes": [_P, _I],
    "flue_b2_grad": [_P, _P, _I, _P, _P, _SZ, _P],
    "flue_b2_grad_ex": [_P, _P, _I, _I, _P, _P, _SZ, _P],
    "flue_flow_stage": [_P, _P, _I, _P, _P, _D, _I, _P],
    "flue_wilson_flow": [_P, _P, _I, _D, _I, _I, _P, _SZ, _P],
    "flue_autotune": [_I, _P, _P, _I, _I, _P, _P, _P, _SZ, _P],
    "flue_autotune_work_bytes": [_I, _P, _I],
    "flue_tuned_variant": [_I, _P, _I, _P],
    "flue_tune_save": [ctypes.c_char_p],
    "flue_tune_load": [ctypes.c_char_p],
    "flue_ensemble_field_bytes": [_I, _P, _I],
    "flue_ensemble_work_bytes": [_I, _P, _I],
    "flue_ensemble_random": [_I, _P, _P, _D, _P, _I, _P],
    "flue_ensemble_flow": [_I, _P, _P, _I, _D, _I, _I, _P, _P, _I, _P, _SZ, _P],
    "flue_ensemble_autotune": [_I, _P, _P, _I, _I, _P, _P, _P, _SZ, _P],
}
_RESTYPE = {"flue_strerror": ctypes.c_char_p, "flue_field_bytes": _SZ, "flue_workspace_bytes": _SZ,
            "flue_b2_work_bytes": _SZ, "flue_flow_work_bytes": _SZ, "flue_ensemble_field_bytes": _SZ,
            "flue_ensemble_work_bytes": _SZ, "flue_autotune_work_bytes": _SZ,
            "flue_pf_work_bytes": _SZ, "flue_magnetic_work_bytes": _SZ}


def lib():
    """The loaded library with declared signatures; raises FlueNotBuilt if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FlueNotBuilt(f"{LIB_PATH} is missing: run `python -m flue_lat_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(status, what):
    if status != 0:
        raise FlueError(status, what)


def exported_symbols():
    return list(SIGNATURES)
