"""The rectangle-first magnetic form (variant 7 of flue_magnetic_ex).

Each rectangle of the clover is formed once at its lower-left corner and its leaves at the
other three corners follow by conjugation; the per-tile footprints are summed in a fixed
order.  Checked against the oracle (Loop5MuNuCoord and magnetic, src/wloops.F90:201-544), the
abelian closed form in each spatial plane, gauge invariance, determinism, and the fallback to
the default form where the extents do not tile or the workspace is too small.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from fields import abelian_flux, f5, gauge_transform, random_su3

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.fixture(scope="module")
def FL():
    import flue_lat_amd as FL
    return FL


# tiles of 16 x 16 base sites: one tile per plane row (16), several (32, 48), and planes
# whose second direction has a different tile count
@pytest.mark.parametrize("dims", [(3, 16, 16, 16), (2, 32, 16, 48), (1, 48, 32, 16)])
def test_rect_matches_oracle(FL, dims):
    print(f"[rect] oracle {dims}", flush=True)
    U = O.random_field(dims, 7100 + sum(dims), 0.3)
    want = O.magnetic(U)
    g = FL.GaugeField.from_complex(U)
    b7 = g.magnetic(variant=7)
    assert rel(b7, want) < 1e-12
    assert rel(b7, g.magnetic(variant=4)) < 1e-12
    assert rel(g.astype("fp64-su3").magnetic(variant=7), want) < 1e-12
    assert rel(g.astype("fp32").magnetic(variant=7), want) < 1e-6
    # fixed summation order: repeated calls are bit-identical
    assert all(g.magnetic(variant=7) == b7 for _ in range(3))


@pytest.mark.parametrize("plane", [(2, 3), (2, 4), (3, 4)])
def test_rect_abelian_flux(FL, plane):
    # constant flux in one spatial plane: B = 2 f5(theta)^2 (the clover of every site sees
    # the same phases), and the unit field gives 0
    dims = (2, 16, 16, 32)
    U, th = abelian_flux(dims, plane[0], plane[1], 3)
    g = FL.GaugeField.from_complex(U)
    assert abs(g.magnetic(variant=7) - 2 * f5(th) ** 2) < 1e-13
    assert FL.GaugeField.from_complex(abelian_flux(dims, 2, 3, 0)[0]).magnetic(variant=7) == 0.0


def test_rect_gauge_invariance(FL):
    dims = (2, 16, 32, 16)
    U = O.random_field(dims, 7300, 0.35)
    G = random_su3(np.random.default_rng(7301), tuple(dims))
    b = FL.GaugeField.from_complex(U).magnetic(variant=7)
    bg = FL.GaugeField.from_complex(gauge_transform(U, G)).magnetic(variant=7)
    assert rel(bg, b) < 1e-12


def test_rect_falls_back_where_it_does_not_apply(FL):
    from flue_lat_amd._lib import lib
    # extents that do not tile: the default form, bit for bit
    U = O.random_field((4, 6, 5, 4), 7400, 0.3)
    g = FL.GaugeField.from_complex(U)
    assert g.magnetic(variant=7) == g.magnetic(variant=4)
    g32 = g.astype("fp32")
    assert g32.magnetic(variant=7) == g32.magnetic(variant=5)
    # a tileable lattice with a workspace too small for the footprints
    U = O.random_field((2, 16, 16, 16), 7401, 0.3)
    g = FL.GaugeField.from_complex(U)
    ws = g.workspace()
    assert ws.numel() < lib().flue_magnetic_work_bytes(g._cdims, 0)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    rc = lib().flue_magnetic_ex(g._cdims, ctypes.c_void_p(g.data.data_ptr()), 0, 7, ctypes.c_void_p(out.data_ptr()),
                                ctypes.c_void_p(ws.data_ptr()), ws.numel(),
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0 and out.item() == g.magnetic(variant=4)


def test_rect_benchmark_lattice(FL):
    # 48^3x96, the benchmark field: every precision against the plane-major forms, and the
    # tuner's choice among all seven variants
    dims = (96, 48, 48, 48)
    f64 = FL.GaugeField.random(dims, 1003, 0.3)
    want = f64.magnetic(variant=4)
    b7 = f64.magnetic(variant=7)
    print(f"[rect] 48^3x96 fp64 rel {rel(b7, want):.2e}", flush=True)
    assert rel(b7, want) < 1e-12
    su3 = FL.GaugeField(dims, "fp64-su3", _data=f64.data)
    assert rel(su3.magnetic(variant=7), want) < 1e-12
    f32 = f64.astype("fp32")
    assert rel(f32.magnetic(variant=7), want) < 1e-6
    assert rel(f32.magnetic(variant=7), f32.magnetic(variant=5)) < 1e-6
    v, ms = f32.autotune("magnetic", reps=2)
    assert 1 <= v <= 7 and ms > 0
    assert rel(f32.magnetic(), want) < 1e-6
    del f64, su3, f32
    torch.cuda.empty_cache()


def test_rect_clover_sums_feed_the_b2_adjoint(FL):
    # on tileable extents the B^2 adjoint takes its per-site clover sums from the rectangle-first
    # tiles (plane (y,x) as the tiles' (x,y) reversed): both adjoint forms against the oracle
    dims = (2, 16, 16, 16)
    U = O.random_field(dims, 7500, 0.3)
    Go = O.b2_grad(U)
    scale = np.abs(Go).max()
    g = FL.GaugeField.from_complex(U)
    for v in (1, 2):
        G = g.magnetic_grad(variant=v).to_complex().cpu().numpy()
        assert np.abs(G - Go).max() < 1e-12 * scale, v
    G32 = g.astype("fp32").magnetic_grad().to_complex().cpu().numpy()
    assert np.abs(G32 - Go).max() < 1e-5 * scale
