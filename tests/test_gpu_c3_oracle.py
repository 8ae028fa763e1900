"""Parity with the FP64 oracle on the 48^3x96 benchmark field itself (seed 1003, eps 0.3,
10,616,832 sites): the configuration the benchmark's Gsites/s figure is quoted on.

The field is generated on the GPU in FP64 and exported to the host unchanged, so the
oracle sees exactly the links the FP64 path reads; the FP32 path reads the same field
rounded to complex64 and is compared with the FP64 oracle, as the precision workload
prescribes.  Checked against the oracle:
  * per-site F of the three magnetic planes (3,2), (2,4), (4,3) at 96 sampled sites,
    including sites on every boundary face (src/wloops.F90:201-432);
  * the full-lattice plaquette sums (1,4,4), (2,4,4), (1,1,4) (src/wloops.F90:915-1017),
    the Polyakov sum (:1019-1069) and B over all 96 time slices (:479-544).
Bounds: FP64 1e-12 relative (per-site F relative to max|F|); FP32 links 1e-6 relative on
B and the plaquette sums, 2e-5 of max|F| per site, and 1e-6 per spatial site on the
Polyakov sum (a product of 96 links whose sum nearly cancels).
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

DIMS = (96, 48, 48, 48)
SEED, EPS = 1003, 0.3
MAG_PLANES = ((3, 2), (2, 4), (4, 3))
PLAQ_SETS = ((1, 4, 4), (2, 4, 4), (1, 1, 4))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.fixture(scope="module")
def FL():
    import flue_lat_amd as FL
    return FL


def sample_sites(n=96, seed=31):
    """(t, x, y, z) of n sites: random ones plus corners and faces where every wrap fires."""
    NT, NX, NY, NZ = DIMS
    rng = np.random.default_rng(seed)
    pts = [(0, 0, 0, 0), (NT - 1, NX - 1, NY - 1, NZ - 1), (0, NX - 1, 0, NZ - 1), (NT - 1, 0, NY - 1, 0),
           (5, 1, NY - 2, 2), (50, NX - 3, 2, NZ - 3)]
    while len(pts) < n:
        pts.append((int(rng.integers(NT)), int(rng.integers(NX)), int(rng.integers(NY)), int(rng.integers(NZ))))
    return np.array(pts, dtype=np.int64)


@pytest.fixture(scope="module")
def bench_field(FL):
    say("[c3] generating the 48^3x96 field (FP64, seed 1003, eps 0.3) on the GPU")
    g64 = FL.GaugeField.random(DIMS, SEED, EPS, "fp64")
    U = g64.to_complex().cpu().numpy()
    torch.cuda.synchronize()
    say("[c3] oracle: plaquette sums, Polyakov sum, B over the full lattice")
    ref = {"plaq": {s: O.plaquette(U, *s) for s in PLAQ_SETS}, "poly": O.polyakov(U)}
    ref["B"] = O.magnetic(U)
    pts = sample_sites()
    t, x, y, z = pts.T
    NT, NX, NY, NZ = DIMS
    fortran = t + NT * (x + NX * (y + NY * z))
    ref["F"] = {pl: O.loop5_munu_sites(U, pl[0], pl[1], fortran) for pl in MAG_PLANES}
    say(f"[c3] oracle done: B = {ref['B']:.15g}")
    del U
    return g64, pts, ref


def observables(g):
    out = {"plaq": {s: g.plaquette(*s) for s in PLAQ_SETS}, "poly": g.polyakov(), "B": g.magnetic()}
    return out


def check_sites(g, pts, ref, tol):
    t, x, y, z = (torch.from_numpy(c).cuda() for c in pts.T)
    errs = {}
    for pl in MAG_PLANES:
        F = g.fmunu(*pl)
        got = F[t, x, y, z].cpu().numpy()
        del F
        want = ref["F"][pl]
        err = np.abs(got - want).max() / np.abs(want).max()
        errs[pl] = err
        assert err < tol, (pl, err)
    return errs


def test_c3_fp64_against_oracle(FL, bench_field):
    g64, pts, ref = bench_field
    say("[c3] FP64 observables")
    got = observables(g64)
    for s in PLAQ_SETS:
        assert got["plaq"][s][1] == ref["plaq"][s][1], s
        assert rel(got["plaq"][s][0], ref["plaq"][s][0]) < 1e-12, (s, got["plaq"][s], ref["plaq"][s])
    assert got["poly"][1] == ref["poly"][1]
    assert rel(got["poly"][0], ref["poly"][0]) < 1e-12, (got["poly"], ref["poly"])
    assert rel(got["B"], ref["B"]) < 1e-12, (got["B"], ref["B"])
    say("[c3] FP64 per-site F")
    errs = check_sites(g64, pts, ref, 1e-12)
    say(f"[c3] FP64 rel err: B {rel(got['B'], ref['B']):.2e}, plaquette "
        f"{rel(got['plaq'][(1, 4, 4)][0], ref['plaq'][(1, 4, 4)][0]):.2e}, Polyakov "
        f"{rel(got['poly'][0], ref['poly'][0]):.2e}, F {max(errs.values()):.2e}")


@pytest.fixture(scope="module")
def field32(bench_field):
    return bench_field[0].astype("fp32")


def test_c3_fp32_against_oracle(FL, bench_field, field32):
    _, pts, ref = bench_field
    g32 = field32
    say("[c3] FP32-link observables")
    got = observables(g32)
    for s in PLAQ_SETS:
        assert got["plaq"][s][1] == ref["plaq"][s][1], s
        assert rel(got["plaq"][s][0], ref["plaq"][s][0]) < 1e-6, (s, got["plaq"][s], ref["plaq"][s])
    assert abs(got["poly"][0] - ref["poly"][0]) < 1e-6 * ref["poly"][1], (got["poly"], ref["poly"])
    assert rel(got["B"], ref["B"]) < 1e-6, (got["B"], ref["B"])
    say("[c3] FP32 per-site F")
    errs = check_sites(g32, pts, ref, 2e-5)
    say(f"[c3] FP32 rel err: B {rel(got['B'], ref['B']):.2e}, plaquette "
        f"{rel(got['plaq'][(1, 4, 4)][0], ref['plaq'][(1, 4, 4)][0]):.2e}, Polyakov "
        f"{rel(got['poly'][0], ref['poly'][0]):.2e}, F {max(errs.values()):.2e}")


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6])
def test_c3_every_magnetic_variant(FL, bench_field, field32, variant):
    # every form and occupancy target of the headline kernel at the headline size
    g64, _, ref = bench_field
    assert rel(g64.magnetic(variant=variant), ref["B"]) < 1e-12, variant
    assert rel(field32.magnetic(variant=variant), ref["B"]) < 1e-6, variant
    su3 = FL.GaugeField(DIMS, "fp64-su3", _data=g64.data)
    assert rel(su3.magnetic(variant=variant), ref["B"]) < 1e-12, variant
