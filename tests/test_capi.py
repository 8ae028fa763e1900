"""CPU-side checks of the C ABI: the library loads, exports every symbol the header
declares, and rejects bad arguments before any device work."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flue_c.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "return")))


@pytest.fixture(scope="module")
def L():
    from flue_lat_amd import _lib
    return _lib.lib()


def test_header_lists_the_hot_path():
    names = header_functions()
    for n in ("genplaquette_c", "plaquette_c", "polyakov_c", "genericpath_c", "loop5munucoord_c",
              "cloverloopmunucoord_c", "magnetic_c", "wmunu_c", "wmunutau_c", "flue_magnetic", "flue_plaquette"):
        assert n in names


def test_library_exports_every_header_symbol(L):
    for n in header_functions():
        assert hasattr(L, n), n


def test_python_signatures_cover_header():
    from flue_lat_amd import _lib
    assert set(header_functions()) <= set(_lib.SIGNATURES)


def test_strerror_and_version(L):
    assert L.flue_strerror(0) == b"ok"
    assert b"NT >= 3" in L.flue_strerror(4)
    a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    assert L.flue_version(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0
    assert (a.value, b.value) == (0, 1)


def test_host_helpers_match_reference_semantics():
    import flue_lat_amd as FL
    A = np.arange(9).reshape(3, 3) + 1j * np.arange(9, 18).reshape(3, 3)
    B = np.eye(3) * (2 - 1j) + 0.5
    np.testing.assert_allclose(FL.gluonCalc.multiplymatmat(A, B), A @ B, atol=1e-13)
    q = FL.gluonAna.get_qhat([1.0, 2.0, 3.0, 4.0], [8, 8, 8, 16], a=0.5)
    np.testing.assert_allclose(q, 2 * np.pi / 0.5 * np.array([1, 2, 3, 4]) / np.array([8, 8, 8, 16]))
    np.testing.assert_array_equal(FL.compiled.flue_c.periodcoord_c([5, 0, 3, 8], [4, 5, 6, 7, 4, 3, 3]),
                                  [1, 5, 3, 1])


def test_propagator_argument_errors_before_device_work():
    from flue_lat_amd._lib import FlueError
    import flue_lat_amd as FL
    # D(NS,NS,NS,NS) cannot hold a field with NT > NS
    with pytest.raises(FlueError) as e:
        FL.gluonCalc.scalarGluonProp(np.zeros((5, 4, 4, 4, 4, 3, 3), complex), 1.0)
    assert e.value.status == 5
    with pytest.raises(FlueError) as e:
        FL.gluonCalc.calc_mom_space_scalarD(np.zeros((5, 4, 4, 4, 4, 3, 3), complex), -1, 1.0)
    assert e.value.status == 1
    with pytest.raises(ValueError):
        FL.gluonCalc.calc_mom_space_scalarD(np.zeros((5, 4, 4, 3, 4, 3, 3), complex), 1, 1.0)
    with pytest.raises(FlueError) as e:
        FL.gluonAna.q_average(3, [1.0, 2.0, 3.0], [0.0, 0.0, 0.0], 0)
    assert e.value.status == 1
    # D(NS^4) past 2^31 entries is refused by the size check, which comes before the pointer
    # checks: NULL pointers get FLUE_EDIM here, never a launch
    from flue_lat_amd import _lib
    dims = (ctypes.c_int32 * 4)(1, 256, 256, 256)
    assert _lib.lib().flue_scalar_gluon_prop(dims, None, 0, None, None) == 5
    dims = (ctypes.c_int32 * 4)(1, 4, 4, 4)
    assert _lib.lib().flue_scalar_gluon_prop(dims, None, 0, None, None) == 2


def test_argument_errors_before_device_work(L):
    from flue_lat_amd._lib import FlueError, check
    import flue_lat_amd as FL
    U = np.zeros((4, 4, 4, 4, 4, 3, 3), complex)
    with pytest.raises(FlueError) as e:
        FL.compiled.flue_c.genericpath_c(U, [1, 1, 1, 1], [1, 5])
    assert e.value.status == 3
    with pytest.raises(FlueError) as e:
        FL.compiled.flue_c.genplaquette_c(U, 0, 4, 4)
    assert e.value.status == 1
    d = (ctypes.c_int32 * 4)(4, 4, 0, 4)
    assert L.flue_field_bytes(d, 0) == 0
    d = (ctypes.c_int32 * 4)(16, 8, 8, 8)
    assert L.flue_field_bytes(d, 0) == 72 * 8 * 8192
    assert L.flue_field_bytes(d, 1) == 72 * 4 * 8192
    assert L.flue_workspace_bytes(d) > 0


def test_no_gpu_is_reported_not_faked(L):
    import flue_lat_amd as FL
    from flue_lat_amd._lib import FlueError
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    with pytest.raises(FlueError) as e:
        FL.compiled.flue_c.magnetic_c(np.zeros((3, 3, 3, 3, 4, 3, 3), complex))
    assert e.value.status == 9


def test_autotune_table_roundtrip(L, tmp_path):
    import ctypes
    p = tmp_path / "tune.txt"
    p.write_text("# header\n0 0 1 4 5 6 7 3 0.25\n0 0 0 4 5 6 7 9 0.5\n"
                 "0 1 0 4 5 6 7 2 9.0\n0 1 1 4 5 6 7 3 9.0\n0 3 0 4 5 6 7 5 1.0\n0 4 1 4 5 6 7 4 2.0\n"
                 "0 9 0 4 5 6 7 1 1.0\n")
    assert L.flue_tune_load(str(p).encode()) == 0
    v = ctypes.c_int32(-1)
    d = (ctypes.c_int32 * 4)(4, 5, 6, 7)
    assert L.flue_tuned_variant(0, d, 1, ctypes.byref(v)) == 0 and v.value == 3
    # out-of-range variants (per operation) and unknown operations are ignored on load
    assert L.flue_tuned_variant(0, d, 0, ctypes.byref(v)) == 0 and v.value == 0
    assert L.flue_tuned_variant(1, d, 0, ctypes.byref(v)) == 0 and v.value == 2
    assert L.flue_tuned_variant(1, d, 1, ctypes.byref(v)) == 0 and v.value == 0
    assert L.flue_tuned_variant(3, d, 0, ctypes.byref(v)) == 0 and v.value == 5
    assert L.flue_tuned_variant(4, d, 1, ctypes.byref(v)) == 0 and v.value == 4
    assert L.flue_tuned_variant(9, d, 0, ctypes.byref(v)) == 0 and v.value == 0
    # work sizes of the tunable operations (no device needed)
    for op in range(4):
        assert L.flue_autotune_work_bytes(op, d, 0) > 0
    assert L.flue_autotune_work_bytes(4, d, 0) == 0 and L.flue_autotune_work_bytes(7, d, 0) == 0
    q = tmp_path / "out.txt"
    assert L.flue_tune_save(str(q).encode()) == 0
    assert "0 0 1 4 5 6 7 3" in q.read_text()
