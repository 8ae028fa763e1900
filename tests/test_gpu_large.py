"""Fields too large for the buffer-resource loads (18 V sizeof(real) >= 4 GiB per direction):
the kernels then take the plain global-load view.  Checked through a size-independent
property: a field tiled k times along t repeats every per-site value, so B and the
Wilson-loop averages are unchanged, plaquette sums scale by k, and one flow step acts
slice by slice.  Fields are 17 GB on the device.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SMALL = (8, 16, 16, 16)
VS = 16 ** 3


def tiled(FL, g, k):
    NT = g.dims[0]
    # 36 complex planes of V (re, im) pairs; a time slice is 2 VS contiguous reals of a plane
    data = g.data.view(36, NT, 2 * VS).repeat(1, k, 1).reshape(-1)
    return FL.GaugeField((NT * k,) + g.dims[1:], g.precision, _data=data)


@pytest.fixture(scope="module")
def FL():
    import flue_lat_amd as FL
    return FL


def dir_bytes(prec, k):
    """bytes of the 9 complex planes of one direction of the tiled field"""
    return 18 * SMALL[0] * k * VS * (8 if prec == "fp64" else 4)


@pytest.mark.parametrize("prec,k", [("fp64", 920), ("fp32", 1830)])
def test_unbuffered_view_matches_small_field(FL, prec, k):
    assert dir_bytes(prec, k) >= 2 ** 32   # past the buffer-resource range
    check_tiled(FL, prec, k)


@pytest.mark.parametrize("prec,k", [("fp64", 600), ("fp32", 1200)])
def test_buffer_records_above_2gib(FL, prec, k):
    # buffer resources whose record count has bit 31 set (2 GiB <= 9 planes < 4 GiB): the
    # common 64^4 FP64 lattice is in this window
    assert 2 ** 31 <= dir_bytes(prec, k) < 2 ** 32
    check_tiled(FL, prec, k)


def check_tiled(FL, prec, k):
    print(f"[large] {prec} tiled x{k}: {dir_bytes(prec, k) / 2 ** 30:.2f} GiB per direction", flush=True)
    g = FL.GaugeField.random(SMALL, 77, 0.3, "fp64").astype(prec)
    big = tiled(FL, g, k)
    # past the buffer-resource range there is one magnetic form (site-major, variant 2 in FP32
    # and 1 in FP64): the small field runs the same per-site arithmetic for the comparison
    v = (2 if prec == "fp32" else 1) if dir_bytes(prec, k) >= 2 ** 32 else 0
    try:
        B, Bb = g.magnetic(variant=v), big.magnetic(variant=v)
        assert abs(Bb - B) < 1e-12 * B, (B, Bb)
        s, n = g.plaquette(1, 4, 4)
        sb, nb = big.plaquette(1, 4, 4)
        assert nb == k * n and abs(sb - k * s) < 1e-12 * abs(k * s)
        W = g.wilson_loops(2, 1, 3, 6)
        Wb = big.wilson_loops(2, 1, 3, 6)
        assert np.abs(Wb - W).max() < 1e-12
        # per-site F at sites of the last copy against the small field
        F = g.fmunu(3, 1)
        Fb = big.fmunu(3, 1)
        off = (k - 1) * SMALL[0]
        tol = (1e-13 if prec == "fp64" else 1e-6) * float(F.abs().max())
        for t in (0, 3, 7):
            assert float((Fb[off + t] - F[t]).abs().max()) <= tol
        del F, Fb
    finally:
        del big
        torch.cuda.empty_cache()


@pytest.mark.parametrize("k", [920, 600])
def test_unbuffered_flow_step_acts_slice_by_slice(FL, k):
    # k = 920: past the buffer-resource range; k = 600: record count above 2 GiB
    g = FL.GaugeField.random(SMALL, 78, 0.3, "fp64")
    big = tiled(FL, g, k)
    try:
        g.wilson_flow(0.02, 1, use_graph=False)
        big.wilson_flow(0.02, 1, use_graph=False)
        a = g.data.view(36, SMALL[0], 2 * VS)
        b = big.data.view(36, SMALL[0] * k, 2 * VS)
        for c in (0, k // 2, k - 1):
            assert float((b[:, c * SMALL[0]:(c + 1) * SMALL[0]] - a).abs().max()) < 1e-13
    finally:
        del big
        torch.cuda.empty_cache()


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
def test_benchmark_size_symmetries(FL, prec, tol):
    # the 48^3x96 benchmark field: B and the plaquette are invariant under a cyclic shift of
    # the lattice and under exchanging x and y (with the link directions), which reverses the
    # orientation of every magnetic plane and maps them onto each other
    dims = (96, 48, 48, 48)
    g = FL.GaugeField.random(dims, 1003, 0.3, "fp64").astype(prec)
    B0 = g.magnetic()
    P0, _ = g.plaquette()
    L0, _ = g.polyakov()
    W0 = g.wilson_loops(2, 1, 6, 6)
    NT, NX, NY, NZ = dims
    v = g.data.view(4, 9, NT, NZ, NY, NX, 2)
    try:
        shifted = FL.GaugeField(dims, prec, _data=torch.roll(v, shifts=(1, 5, 2), dims=(2, 3, 5)).reshape(-1))
        assert abs(shifted.magnetic() - B0) <= tol * B0
        assert abs(shifted.plaquette()[0] - P0) <= tol * abs(P0)
        assert abs(shifted.polyakov()[0] - L0) <= tol * 1e3 * max(1.0, abs(L0))
        assert np.abs(shifted.wilson_loops(2, 1, 6, 6) - W0).max() <= tol
        del shifted
        # x <-> y: sites (t, z, y, x) -> (t, z, x, y), directions (t, x, y, z) -> (t, y, x, z)
        swapped = v[[0, 2, 1, 3]].transpose(4, 5).contiguous().reshape(-1)
        sw = FL.GaugeField(dims, prec, _data=swapped)
        assert abs(sw.magnetic() - B0) <= tol * B0
        assert abs(sw.plaquette()[0] - P0) <= tol * abs(P0)
        # the (x, t) loops of the field are the (y, t) loops of the exchanged one
        assert np.abs(sw.wilson_loops(3, 1, 6, 6) - W0).max() <= tol
        del sw, swapped
    finally:
        torch.cuda.empty_cache()
