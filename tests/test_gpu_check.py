"""The SU(3) check of link fields (flue_check_field) and the entry points that rely on it:
FP32 conversion, the "fp64-su3" declaration and the B^2 gradient get a status code back,
FLUE_ESU3, for links that are not SU(3), instead of silently wrong results.

The reference reunitarises every link it reads (src/gaugeIO/ILDG_bin.f90:138-151,
openqcdfileio_standalone.f90:80-83), so its kernels only ever see SU(3) links; the library's
FP32 two-row chains and the B^2 gather assume the same and check it.
"""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

ESU3 = 11


@pytest.fixture(scope="module")
def FL():
    import flue_lat_amd as FL
    return FL


def spoiled(kind):
    """A seeded SU(3) field with one link spoiled: scaled (not unitary), given a U(1)
    phase (unitary, det != 1) or holding a NaN."""
    U = O.random_field((4, 4, 3, 5), 808, 0.3)
    c = (2, 1, 0, 3, 2)
    if kind == "scaled":
        U[c] *= 1.001
    elif kind == "phase":
        U[c] *= np.exp(0.01j)
    elif kind == "nan":
        U[c + (1, 2)] = complex(np.nan, 0.0)
    return U


def test_su3_fields_pass_with_small_deviations(FL):
    U = O.random_field((4, 4, 3, 5), 808, 0.3)
    g = FL.GaugeField.from_complex(U)
    st = g.check_su3()
    assert st["nonfinite"] == 0 and st["unitarity"] < 1e-14 and st["det"] < 1e-14
    st32 = g.astype("fp32").check_su3()
    assert st32["nonfinite"] == 0 and 0 < st32["unitarity"] < 1e-6 and st32["det"] < 1e-6
    # the 48^3x96 generator output in both precisions
    big = FL.GaugeField.random((96, 48, 48, 48), 1003, 0.3)
    assert big.check_su3()["unitarity"] < 1e-14
    assert big.astype("fp32").check_su3()["unitarity"] < 1e-6


@pytest.mark.parametrize("kind", ["scaled", "phase", "nan"])
def test_spoiled_links_get_the_status_code(FL, kind):
    from flue_lat_amd._lib import FlueError, lib
    U = spoiled(kind)
    g = FL.GaugeField.from_complex(U)                       # plain FP64 assumes nothing
    st = g.check_su3(raise_on_fail=False)
    if kind == "nan":
        assert st["nonfinite"] == 1
    elif kind == "scaled":
        assert st["unitarity"] > 1e-3
    else:
        assert st["unitarity"] < 1e-14 and abs(st["det"] - abs(np.exp(0.03j) - 1)) < 1e-12
    # the C entry returns the status as a normal error return
    stats = (ctypes.c_double * 3)()
    ws = g.workspace()
    rc = lib().flue_check_field(g._cdims, ctypes.c_void_p(g.data.data_ptr()), 0, ctypes.c_double(0.0), stats,
                                ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == ESU3
    for make in (lambda: g.astype("fp32"),
                 lambda: g.astype("fp64-su3"),
                 lambda: FL.GaugeField(g.dims, "fp64-su3", _data=g.data),
                 lambda: FL.GaugeField.from_complex(U, "fp32"),
                 lambda: g.magnetic_grad()):
        with pytest.raises(FlueError) as e:
            make()
        assert e.value.status == ESU3
    # the C gradient entry returns the code and leaves the gradient untouched
    grad = torch.full_like(g.data, 7.0)
    n = lib().flue_b2_work_bytes(g._cdims, 0)
    work = torch.empty(n, dtype=torch.uint8, device=g.data.device)
    rc = lib().flue_b2_grad(g._cdims, ctypes.c_void_p(g.data.data_ptr()), 0, ctypes.c_void_p(grad.data_ptr()),
                            ctypes.c_void_p(work.data_ptr()), n,
                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == ESU3
    assert bool((grad == 7.0).all())
    # the plain FP64 measurements still run and the opt-outs are honoured
    if kind != "nan":
        assert abs(g.magnetic() - O.magnetic(U)) < 1e-12 * O.magnetic(U)
    g.astype("fp32", check=False)


def test_check_tolerance_argument(FL):
    U = spoiled("scaled")
    g = FL.GaugeField.from_complex(U)
    st = g.check_su3(tol=1e-2)             # a looser bound accepts the field
    assert 1e-3 < st["unitarity"] < 1e-2
    from flue_lat_amd._lib import lib
    rc = lib().flue_check_field(g._cdims, ctypes.c_void_p(g.data.data_ptr()), 0, ctypes.c_double(-1.0), None,
                                None, 0, None)
    assert rc == 1                          # a negative bound is an argument error
