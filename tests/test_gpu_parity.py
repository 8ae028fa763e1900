"""Parity of the HIP path (libflue_hip.so) with the CPU FP64 oracle.

Inputs are the oracle's seeded fields (complex128), imported to the GPU, so both
sides see identical links.  Tolerances:
  * FP64 path: 1e-12 relative on lattice sums and per-site F (F scaled by its max);
  * FP32 links with FP64 sums: 1e-6 relative on the magnetic B and the plaquette sums,
    2e-5 relative on per-site F and 1e-5 on the Polyakov sum (a product of NT links).
"""
import numpy as np
import pytest

import oracle as O
from tests.fields import abelian_flux, f5, unit_field

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.fixture(scope="module")
def FL():
    import flue_lat_amd as FL
    return FL


@pytest.fixture(scope="module")
def c0():
    # the small CPU-runnable workload: 8^3x16, eps = 0.25
    return O.random_field((16, 8, 8, 8), 1000, 0.25)


@pytest.fixture(scope="module")
def small():
    # uneven extents exercise every wrap and stride
    return O.random_field((6, 5, 7, 4), 1234, 0.35)


def gfield(FL, U, prec="fp64"):
    return FL.GaugeField.from_complex(U, prec)


# ------------------------------------------------------------------ field plumbing

def test_import_export_roundtrip_exact(FL, small):
    g = gfield(FL, small)
    back = g.to_complex().cpu().numpy()
    assert np.array_equal(back, small)
    # a C-ordered torch tensor on the GPU imports identically
    t = torch.from_numpy(np.ascontiguousarray(small)).cuda()
    g2 = FL.GaugeField.from_complex(t)
    assert torch.equal(g2.data, g.data)


def test_device_generator_matches_oracle(FL):
    dims = (6, 5, 4, 3)
    g = FL.GaugeField.random(dims, 4242, 0.3)
    U = O.random_field(dims, 4242, 0.3)
    assert np.abs(g.to_complex().cpu().numpy() - U).max() < 1e-13


# ------------------------------------------------------------------ plaquette family

@pytest.mark.parametrize("args", [(1, 4, 4), (2, 4, 4), (1, 1, 4), (1, 2, 4), (3, 3, 4)])
def test_plaquette_fp64(FL, c0, args):
    want, n = O.plaquette(c0, *args)
    got, nP = gfield(FL, c0).plaquette(*args)
    assert nP == n
    assert rel(got, want) < 1e-12


def test_plaquette_host_dropins(FL, small):
    fc = FL.compiled.flue_c
    for args in ((1, 4, 4), (2, 4, 4), (1, 1, 4)):
        want, n = O.gen_plaquette(small, *args)
        s, nP, t = fc.genplaquette_c(small, *args)
        assert nP == n and rel(s, want) < 1e-12 and t >= 0
        s2, nP2, _ = fc.plaquette_c(small, *args)
        assert nP2 == n and rel(s2, want) < 1e-12


def test_plaquette_unit_and_flux(FL):
    dims = (4, 6, 5, 4)
    s, n = gfield(FL, unit_field(dims)).plaquette()
    assert s == 3.0 * n
    U, th = abelian_flux(dims, 2, 3, 1)
    V = np.prod(dims)
    s, _ = gfield(FL, U).plaquette()
    assert rel(s, 15 * V + (2 * np.cos(th) + 1) * V) < 1e-13


def test_plaquette_fp32(FL, c0):
    want, _ = O.plaquette(c0, 1, 4, 4)
    got, _ = gfield(FL, c0, "fp32").plaquette()
    assert rel(got, want) < 1e-6


# ------------------------------------------------------------------ Polyakov

def test_polyakov_fp64(FL, c0, small):
    for U in (c0, small):
        want, n = O.polyakov(U)
        got, nP = gfield(FL, U).polyakov()
        assert nP == n
        assert abs(got - want) < 1e-12 * max(1.0, abs(want)) * n


def test_polyakov_host_and_short_time(FL, small):
    want, n = O.polyakov(small)
    s, nP, _ = FL.compiled.flue_c.polyakov_c(small)
    assert nP == n and abs(s - want) < 1e-12 * n
    with pytest.raises(FL.FlueError):
        FL.compiled.flue_c.polyakov_c(unit_field((2, 3, 3, 3)))


def test_polyakov_constant_diagonal(FL):
    dims = (7, 3, 2, 4)
    U = unit_field(dims)
    p1, p2 = 0.3, -1.1
    U[..., 0, 0, 0] = np.exp(1j * p1)
    U[..., 0, 1, 1] = np.exp(1j * p2)
    U[..., 0, 2, 2] = np.exp(-1j * (p1 + p2))
    s, n = gfield(FL, U).polyakov()
    assert abs(s - n * (np.cos(7 * p1) + np.cos(7 * p2) + np.cos(7 * (p1 + p2)))) < 1e-12


def test_polyakov_fp32(FL, c0):
    want, n = O.polyakov(c0)
    got, _ = gfield(FL, c0, "fp32").polyakov()
    assert abs(got - want) < 1e-5 * n


# ------------------------------------------------------------------ generic path / Wilson loops

@pytest.mark.parametrize("path", [[1], [3], [-2], [1, 2, -1, -2], [2, 2, 3, -4, -1, -1, 1, 4],
                                  [-3, -3, -3, -3, -3, 1, 1, 1, 1], [4, 4, 4, 4, 4, 4, 4, 4]])
def test_generic_path(FL, small, path):
    g = gfield(FL, small)
    for coord in ([1, 1, 1, 1], [6, 5, 7, 4], [2, 3, 1, 4]):
        want = O.generic_path(small, coord, path)
        np.testing.assert_allclose(g.generic_path(coord, path), want, rtol=0, atol=1e-13)
    np.testing.assert_allclose(FL.compiled.flue_c.genericpath_c(small, [2, 3, 1, 4], path),
                               O.generic_path(small, [2, 3, 1, 4], path), rtol=0, atol=1e-13)


def test_generic_path_rejects_bad_direction(FL, small):
    with pytest.raises(FL.FlueError):
        FL.compiled.flue_c.genericpath_c(small, [1, 1, 1, 1], [1, 5])
    with pytest.raises(FL.FlueError):
        gfield(FL, small).generic_path([1, 1, 1, 1], [0, 1])


@pytest.mark.parametrize("mu,nu,R,T", [(2, 1, 1, 1), (3, 1, 2, 3), (4, 2, 3, 1), (1, 4, 6, 6), (2, 3, 5, 2)])
def test_wilson_loops(FL, small, mu, nu, R, T):
    g = gfield(FL, small)
    assert rel(g.wilson_loop(mu, nu, R, T), O.wmunu(small, mu, nu, R, T)) < 1e-12
    assert rel(g.wilson_loop(mu, nu, R, T, tau=3), O.wmunutau(small, 3, mu, nu, R, T)) < 1e-12
    fc = FL.compiled.flue_c
    assert rel(fc.wmunu_c(small, mu, nu, R, T), O.wmunu(small, mu, nu, R, T)) < 1e-12
    assert rel(fc.wmunutau_c(small, 2, mu, nu, R, T), O.wmunutau(small, 2, mu, nu, R, T)) < 1e-12


@pytest.mark.parametrize("mu,nu", [(2, 1), (1, 3), (4, 2)])
def test_wilson_loop_table_fp64(FL, small, mu, nu):
    # R, T up to 6 on extents 4..7: loops wrap the lattice more than once
    g = gfield(FL, small)
    W = g.wilson_loops(mu, nu, 6, 6)
    Wt = g.wilson_loops(mu, nu, 4, 5, tau=2)
    for R in range(1, 7):
        for T in range(1, 7):
            assert abs(W[R - 1, T - 1] - O.wmunu(small, mu, nu, R, T)) < 1e-13, (R, T)
            if R <= 4 and T <= 5:
                assert abs(Wt[R - 1, T - 1] - O.wmunutau(small, 2, mu, nu, R, T)) < 1e-13, (R, T)


def test_wilson_loop_slices(FL, small):
    g = gfield(FL, small)
    W = g.wilson_loops_slices(2, 3, 4, 5)
    assert W.shape == (4, 5, small.shape[0])
    for (R, T) in ((1, 1), (4, 5), (2, 3)):
        for t in (1, 3, 6):
            assert abs(W[R - 1, T - 1, t - 1] - O.wmunutau(small, t, 2, 3, R, T)) < 1e-13, (R, T, t)
    # the slice averages average to the whole-lattice table
    assert np.abs(W.mean(axis=2) - g.wilson_loops(2, 3, 4, 5)).max() < 1e-13
    W1 = g.wilson_loops_slices(1, 4, 6, 6)      # loops reaching along time too
    assert abs(W1[5, 5, 4] - O.wmunutau(small, 5, 1, 4, 6, 6)) < 1e-13


def test_wilson_loop_tables_ragged_slices(FL):
    # 630 sites per slice: two full blocks of 256 and one of 118 per slice, in the
    # slice-fastest order of the time planes, the storage order of (y, z) and the
    # along-z slices of (x, z)
    U = O.random_field((6, 10, 9, 7), 1017, 0.3)
    g = gfield(FL, U)
    for mu, nu in ((2, 1), (1, 4), (3, 4), (2, 4)):
        W = g.wilson_loops(mu, nu, 6, 6)
        S = g.wilson_loops_slices(mu, nu, 6, 6)
        for R, T in ((1, 1), (6, 6), (3, 5), (6, 2)):
            assert abs(W[R - 1, T - 1] - O.wmunu(U, mu, nu, R, T)) < 1e-13, (mu, nu, R, T)
            for t in (1, 4, 6):
                assert abs(S[R - 1, T - 1, t - 1] - O.wmunutau(U, t, mu, nu, R, T)) < 1e-13, (mu, nu, R, T, t)
        assert np.abs(S.mean(axis=2) - W).max() < 1e-13


def test_path_trace_ragged_slices_brute_force(FL):
    # whole-lattice and one-slice sums of ReTr genericPath against the path walked at every
    # site; paths with a time step take the slice-fastest block order (630 sites per slice)
    U = O.random_field((6, 10, 9, 7), 1021, 0.3)
    g = gfield(FL, U)
    NT, NX, NY, NZ = U.shape[:4]
    for path in ([2, 1, 3, -2, -1, -3], [1, 1, 1, 1, 1, 1, 1], [2, 3, -2, -3, 4, -4]):
        per_t = np.zeros(NT)
        for t in range(NT):
            for x in range(NX):
                for y in range(NY):
                    for z in range(NZ):
                        per_t[t] += np.trace(O.generic_path(U, [t + 1, x + 1, y + 1, z + 1], path)).real
        assert abs(g.path_trace(path) - per_t.sum()) < 1e-12 * U[..., 0, 0, 0].size, path
        assert abs(g.path_trace(path, tau=4) - per_t[3]) < 1e-12 * NX * NY * NZ, path


def test_wilson_loop_table_fp32_and_errors(FL, small):
    W = gfield(FL, small, "fp32").wilson_loops(3, 2, 3, 6)
    for R in range(1, 4):
        for T in range(1, 7):
            assert abs(W[R - 1, T - 1] - O.wmunu(small, 3, 2, R, T)) < 1e-6
    g = gfield(FL, small)
    for bad in ((2, 2, 6, 6), (2, 1, 7, 6), (2, 1, 6, 0), (0, 1, 2, 2)):
        with pytest.raises(FL.FlueError):
            g.wilson_loops(*bad)


@pytest.fixture(scope="module")
def c1():
    # 24^3 x 48, eps = 0.35 (the plain FP64 drop-in workload)
    return O.random_field((48, 24, 24, 24), 1001, 0.35)


def test_wilson_loop_table_24x48(FL, c1):
    # R along each spatial direction, T along time, R, T in 1..6
    g = gfield(FL, c1)
    for mu in (2, 3, 4):
        W = g.wilson_loops(mu, 1, 6, 6)
        for R, T in ((1, 1), (2, 3), (6, 6), (5, 2)):
            assert rel(W[R - 1, T - 1], O.wmunu(c1, mu, 1, R, T)) < 1e-12, (mu, R, T)
        # every entry against the loop walked on its own
        for R in range(1, 7):
            for T in range(1, 7):
                assert rel(W[R - 1, T - 1], g.wilson_loop(mu, 1, R, T)) < 1e-12, (mu, R, T)


def test_loop5_sampled_sites_24x48(FL, c1):
    # per-site F on all six planes at 64 sampled sites (Fortran site offsets)
    NT, NX, NY, NZ = c1.shape[:4]
    rng = np.random.default_rng(11)
    sites = rng.choice(NT * NX * NY * NZ, size=64, replace=False)
    t, r = sites % NT, sites // NT
    x, r = r % NX, r // NX
    y, z = r % NY, r // NY
    g = gfield(FL, c1)
    for mu in range(1, 5):
        for nu in range(mu + 1, 5):
            want = O.loop5_munu_sites(c1, mu, nu, sites)
            got = g.fmunu(mu, nu)[t, x, y, z].cpu().numpy()
            assert np.abs(got - want).max() < 1e-12 * np.abs(want).max(), (mu, nu)


# ------------------------------------------------------------------ F_munu and B

ALL_PLANES = [(mu, nu) for mu in range(1, 5) for nu in range(1, 5) if mu != nu]


@pytest.mark.parametrize("mu,nu", ALL_PLANES)
def test_loop5_per_site_fp64(FL, small, mu, nu):
    want = O.loop5_munu(small, mu, nu)
    got = gfield(FL, small).fmunu(mu, nu).cpu().numpy()
    scale = np.abs(want).max()
    assert np.abs(got - want).max() < 1e-12 * scale


def test_loop5_per_site_ragged_slices(FL):
    # 630 sites per slice (blocks of 256, 256 and 118 per slice): planes with time take the
    # blocks slice by slice, the others in storage order; every site and the FP32 form
    U = O.random_field((6, 10, 9, 7), 1019, 0.3)
    g, g32 = gfield(FL, U), gfield(FL, U, "fp32")
    for mu, nu in ((1, 2), (4, 1), (1, 3), (2, 3), (4, 2)):
        want = O.loop5_munu(U, mu, nu)
        scale = np.abs(want).max()
        assert np.abs(g.fmunu(mu, nu).cpu().numpy() - want).max() < 1e-12 * scale, (mu, nu)
        assert np.abs(g32.fmunu(mu, nu).cpu().numpy() - want).max() < 2e-5 * scale, (mu, nu)
        clov = O.clover_munu(U, mu, nu)
        assert np.abs(g.fmunu(mu, nu, clover=True).cpu().numpy() - clov).max() < 1e-12 * np.abs(clov).max()


@pytest.mark.parametrize("mu,nu", [(1, 2), (3, 2), (4, 1)])
def test_clover_per_site_fp64(FL, small, mu, nu):
    want = O.clover_munu(small, mu, nu)
    got = gfield(FL, small).fmunu(mu, nu, clover=True).cpu().numpy()
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


def test_loop5_host_dropins(FL, small):
    fc = FL.compiled.flue_c
    want = O.loop5_munu(small, 2, 4)
    assert np.abs(fc.loop5munucoord_c(small, 2, 4) - want).max() < 1e-12 * np.abs(want).max()
    want = O.clover_munu(small, 3, 1)
    assert np.abs(fc.cloverloopmunucoord_c(small, 3, 1) - want).max() < 1e-12 * np.abs(want).max()


def test_loop5_fp32(FL, small):
    want = O.loop5_munu(small, 2, 3)
    got = gfield(FL, small, "fp32").fmunu(2, 3).cpu().numpy()
    assert np.abs(got - want).max() < 2e-5 * np.abs(want).max()


def test_magnetic_fp64(FL, c0, small):
    for U in (c0, small):
        want = O.magnetic(U)
        assert rel(gfield(FL, U).magnetic(), want) < 1e-12
        assert rel(FL.compiled.flue_c.magnetic_c(U), want) < 1e-12


def test_magnetic_fp32(FL, c0):
    want = O.magnetic(c0)
    assert rel(gfield(FL, c0, "fp32").magnetic(), want) < 1e-6


def test_magnetic_fp64_su3(FL, c0, small):
    # FP64 links declared SU(3): two-row chains in variants 4 and 5, full products elsewhere;
    # every variant within the FP64 bound of the oracle on gauge fields
    for U in (c0, small):
        want = O.magnetic(U)
        g = gfield(FL, U).astype("fp64-su3")
        assert g.precision == "fp64-su3"
        for v in range(1, 7):
            assert rel(g.magnetic(variant=v), want) < 1e-12, v
        best, ms = g.autotune("magnetic", reps=2)
        assert 1 <= best <= 6 and rel(g.magnetic(), want) < 1e-12
        # everything else treats the field as FP64
        s, n = g.plaquette()
        so, no = O.plaquette(U, 1, 4, 4)
        assert n == no and rel(s, so) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 5])
def test_magnetic_abelian_flux(FL, n):
    dims = (2, 7, 6, 3)
    U, th = abelian_flux(dims, 2, 3, n)
    assert abs(gfield(FL, U).magnetic() - 2 * f5(th) ** 2) < 1e-13
    assert gfield(FL, unit_field(dims)).magnetic() == 0.0


def test_sums_are_deterministic(FL, c0):
    g = gfield(FL, c0)
    a = [g.magnetic() for _ in range(3)]
    b = [g.plaquette()[0] for _ in range(3)]
    assert a[0] == a[1] == a[2] and b[0] == b[1] == b[2]


def test_wloops_module_accepts_all_field_kinds(FL, small):
    W = FL.gluonCalc.wloops
    want = O.magnetic(small)
    assert rel(W.magnetic(small), want) < 1e-12
    assert rel(W.magnetic(torch.from_numpy(small).cuda()), want) < 1e-12
    assert rel(W.magnetic(gfield(FL, small)), want) < 1e-12
    s, n, _ = W.genPlaquette(torch.from_numpy(small), 2, 4, 4)
    ws, wn = O.gen_plaquette(small, 2, 4, 4)
    assert n == wn and rel(s, ws) < 1e-12


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_every_magnetic_variant_matches_oracle(FL, c0, small, prec):
    tol = 1e-12 if prec == "fp64" else 1e-6
    for U in (c0, small):
        want = O.magnetic(U)
        g = gfield(FL, U, prec)
        got = [g.magnetic(variant=v) for v in range(1, 7)]
        for v, b in zip(range(1, 7), got):
            assert rel(b, want) < tol, (v, b, want)
        assert g.magnetic(variant=5) == got[4]


@pytest.mark.parametrize("prec,dims", [("fp64", (3, 32, 5, 6)), ("fp32", (2, 64, 4, 5)), ("fp64", (2, 48, 4, 4))])
def test_plane_major_chunked_x_order(FL, prec, dims):
    # NX with a divisor spanning >= 128 bytes: the (z,y) plane visits x in chunks
    U = O.random_field(dims, 77, 0.3)
    want = O.magnetic(U)
    tol = 1e-12 if prec == "fp64" else 1e-6
    g = gfield(FL, U, prec)
    for v in (4, 5, 6):
        assert rel(g.magnetic(variant=v), want) < tol


def test_autotune_picks_a_variant_and_keeps_results(FL, c0):
    g = gfield(FL, c0, "fp32")
    B0 = g.magnetic(variant=2)
    v, ms = g.autotune("magnetic", reps=2)
    assert v in range(1, 7) and ms > 0
    assert g.magnetic() == g.magnetic(variant=v)
    assert abs(g.magnetic() - B0) <= 1e-12 * B0


# ------------------------------------------------------------------ degenerate extents

@pytest.mark.parametrize("dims", [(3, 1, 2, 5), (4, 2, 1, 1), (1, 3, 2, 2), (5, 2, 2, 1)])
def test_extents_of_one_and_two(FL, dims):
    # every 2- and 3-link step wraps one or more times around such directions
    print(f"degenerate extents {dims}", flush=True)
    U = O.random_field(dims, 5151, 0.4)
    g = gfield(FL, U)
    for args in ((1, 4, 4), (2, 4, 4), (1, 1, 4)):
        want, n = O.plaquette(U, *args)
        got, nP = g.plaquette(*args)
        assert nP == n and abs(got - want) <= 1e-12 * max(1.0, abs(want)), args
    for v in range(1, 7):
        assert abs(g.magnetic(variant=v) - O.magnetic(U)) <= 1e-12 * O.magnetic(U), v
    for (mu, nu) in ((2, 3), (1, 4)):
        F = g.fmunu(mu, nu).cpu().numpy()
        Fo = O.loop5_munu(U, mu, nu)
        assert np.abs(F - Fo).max() <= 1e-12 * max(1.0, np.abs(Fo).max()), (mu, nu)
    W = g.wilson_loops(2, 1, 3, 3)
    for R in range(1, 4):
        for T in range(1, 4):
            assert abs(W[R - 1, T - 1] - O.wmunu(U, 2, 1, R, T)) < 1e-12, (R, T)
    W5 = O.wilson_flow(U, 0.02, 1)
    g.wilson_flow(0.02, 1, use_graph=False)
    assert np.abs(g.to_complex().cpu().numpy() - W5).max() < 1e-12
