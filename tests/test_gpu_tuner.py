"""The autotuner's operations and their variants.

Every variant of an operation computes the same quantity: the flow visit orders and the
ensemble grids run the same per-link code (bit-identical results), the B^2 adjoint forms
and plaquette walk lengths sum in different orders (agreement to 1e-12 in FP64, and each
against the oracle).  Variants are forced by loading a tuning table, the way a saved
table steers later calls.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

OPS = {"magnetic": 0, "b2_grad": 1, "plaquette": 2, "flow": 3, "ensemble_flow": 4}


@pytest.fixture(scope="module")
def FL():
    import flue_lat_amd as FL
    return FL


def force(FL, tmp_path, op, key, prec, variant):
    """make `variant` the tuned choice of op for (device 0, prec, key)"""
    p = tmp_path / f"tune_{op}_{variant}.txt"
    p.write_text(f"0 {OPS[op]} {prec} {key[0]} {key[1]} {key[2]} {key[3]} {variant} 1.0\n")
    FL.field.tune_load(p)


def grad_np(g):
    return g.to_complex().cpu().numpy()


@pytest.mark.parametrize("dims", [(2, 4, 4, 5), (3, 5, 7, 6), (2, 1, 2, 3), (1, 3, 2, 2)])
def test_b2_grad_forms_agree_and_match_oracle(FL, dims):
    # extents of 1 and 2 make every 2- and 3-link side of the rectangles wrap
    print(f"[b2] {dims}", flush=True)
    U = O.random_field(dims, 3100 + sum(dims), 0.35)
    Go = O.b2_grad(U)
    scale = np.abs(Go).max()
    g = FL.GaugeField.from_complex(U)
    G1 = grad_np(g.magnetic_grad(variant=1))
    G2 = grad_np(g.magnetic_grad(variant=2))
    assert np.abs(G1 - Go).max() < 1e-12 * scale
    assert np.abs(G2 - Go).max() < 1e-12 * scale
    assert np.abs(G2 - G1).max() < 1e-12 * scale
    g32 = g.astype("fp32")
    assert np.abs(grad_np(g32.magnetic_grad(variant=2)) - Go).max() < 1e-5 * scale
    assert np.abs(grad_np(g32.magnetic_grad(variant=1)) - Go).max() < 1e-5 * scale


def test_b2_grad_reverse_form_slabs(FL):
    # a field whose slot buffer takes several slabs of time slices would need over 4 GiB of
    # slots per slab; at 48^3x96 FP64 the slab is 3 slices: check the slab hand-over there
    # against the gather form on the full benchmark field
    dims = (96, 48, 48, 48)
    g = FL.GaugeField.random(dims, 1003, 0.3)
    G1 = g.magnetic_grad(variant=1).data
    G2 = g.magnetic_grad(variant=2).data
    err = float((G2 - G1).abs().max()) / float(G1.abs().max())
    print(f"[b2] 48^3x96 reverse vs gather: {err:.2e}", flush=True)
    assert err < 1e-12
    del G1, G2, g
    torch.cuda.empty_cache()


def test_plaquette_walk_lengths(FL, tmp_path):
    dims = (13, 6, 5, 4)  # walks that do not divide NT leave short last walks
    U = O.random_field(dims, 2024, 0.3)
    g = FL.GaugeField.from_complex(U)
    want = {s: O.plaquette(U, *s)[0] for s in ((1, 4, 4), (2, 4, 4), (1, 1, 4))}
    S0 = O.wilson_action(U, 6.0)
    for v in range(1, 7):
        force(FL, tmp_path, "plaquette", dims, 0, v)
        for s, w in want.items():
            got, _ = g.plaquette(*s)
            assert abs(got - w) < 1e-12 * abs(w), (v, s)
        assert abs(g.wilson_action(6.0) - S0) < 1e-12 * abs(S0), v
    b, ms = g.autotune("plaquette", reps=2)
    assert 1 <= b <= 6 and ms > 0
    assert abs(g.plaquette()[0] - want[(1, 4, 4)]) < 1e-12 * abs(want[(1, 4, 4)])


@pytest.mark.parametrize("variant", [7, 8])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("dims", [(6, 5, 4, 7), (13, 4, 3, 5), (1, 3, 4, 2), (2, 5, 3, 3)])
def test_plaquette_first_stencil(FL, tmp_path, prec, tol, variant, dims):
    # variants 7 and 8 form each plaquette once per site (8 walking along t, with walks of 6
    # slices: 13 leaves a short last walk, 1 and 2 wrap within a walk); they agree with the
    # staple sums to rounding
    U = O.random_field(dims, 6060 + sum(dims), 0.3)
    code = 0 if prec == "fp64" else 1
    g = FL.GaugeField.from_complex(U, prec)
    Go = O.wilson_grad(U, 6.0)
    scale = np.abs(Go).max()
    if prec == "fp64":
        # plain FP64 assumes nothing about the links: the gradient keeps the staple form
        assert torch.equal(g.wilson_grad(6.0, variant=variant).data, g.wilson_grad(6.0, variant=1).data)
        gs = FL.GaugeField(dims, "fp64-su3", _data=g.data)
        assert np.abs(grad_np(gs.wilson_grad(6.0, variant=variant)) - Go).max() < tol * scale
    else:
        assert np.abs(grad_np(g.wilson_grad(6.0, variant=variant)) - Go).max() < tol * scale
    force(FL, tmp_path, "flow", dims, code, variant)
    W = g.clone()
    W.wilson_flow(0.02, 3, use_graph=True)
    Wo = O.wilson_flow(U, 0.02, 3)
    assert np.abs(W.to_complex().cpu().numpy() - Wo).max() < tol
    force(FL, tmp_path, "flow", dims, code, 1)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_flow_visit_orders_bit_identical(FL, tmp_path, prec):
    # (x, y) planes of 256 sites, NZ and NT divisible by every brick: all six orders apply,
    # to the staple sums (1..6) and to the plaquette-first stencil (7, 9..13)
    dims = (8, 16, 16, 4)
    code = 0 if prec == "fp64" else 1
    base = FL.GaugeField.random(dims, 909, 0.3).astype(prec)
    Wo = O.wilson_flow(base.to_complex().cpu().numpy(), 0.02, 3)
    tol = 1e-12 if prec == "fp64" else 1e-5
    for group in ((1, 2, 3, 4, 5, 6), (7, 9, 10, 11, 12, 13)):
        want_grad = want_flow = None
        for v in group:
            force(FL, tmp_path, "flow", dims, code, v)
            G = base.wilson_grad(6.0).data.clone()
            W = base.clone()
            W.wilson_flow(0.02, 3, use_graph=False)
            if want_grad is None:
                want_grad, want_flow = G, W.data.clone()
                assert np.abs(W.to_complex().cpu().numpy() - Wo).max() < tol
            else:
                assert torch.equal(G, want_grad), v
                assert torch.equal(W.data, want_flow), v
    b, ms = base.autotune("flow", reps=2)
    assert 1 <= b <= 13 and ms > 0


def test_magnetic_and_b2_tuning_keep_results(FL):
    U = O.random_field((4, 6, 5, 4), 4242, 0.3)
    g = FL.GaugeField.from_complex(U)
    want = O.magnetic(U)
    v, _ = g.autotune("magnetic", reps=2)
    assert abs(g.magnetic() - want) < 1e-12 * want and g.magnetic() == g.magnetic(variant=v)
    Go = O.b2_grad(U)
    v, _ = g.autotune("b2_grad", reps=1)
    assert v in (1, 2)
    assert np.abs(grad_np(g.magnetic_grad()) - Go).max() < 1e-12 * np.abs(Go).max()


def test_ensemble_grids_bit_identical(FL, tmp_path):
    from flue_lat_amd.ensemble import Ensemble
    dims = [(16, 8, 8, 8), (24, 8, 8, 8), (5, 6, 7, 3), (32, 8, 8, 8)]
    seeds = [4100 + c for c in range(4)]
    nblk = sum((int(np.prod(d)) + 255) // 256 for d in dims)
    key = (4, nblk, max(int(np.prod(d)) for d in dims), 0)
    # grids 1..4 (staple sums) are bit-identical, as are 5..8 (plaquette-first); the two
    # stencils agree to rounding and each matches the oracle's flow
    ref = {}
    for v in range(1, 9):
        force(FL, tmp_path, "ensemble_flow", key, 0, v)
        e = Ensemble.random(dims, seeds, 0.4, "fp64")
        plaq, B = e.flow(0.01, 4, measure_every=2, use_graph=True)
        form = (v - 1) // 4
        if form not in ref:
            ref[form] = (e.data.clone(), plaq, B)
            for c in (0, 2):
                W = O.wilson_flow(O.random_field(dims[c], seeds[c], 0.4), 0.01, 4)
                assert np.abs(e.field(c).to_complex().cpu().numpy() - W).max() < 1e-12, (v, c)
        else:
            assert torch.equal(e.data, ref[form][0]), v
            assert np.array_equal(plaq, ref[form][1]) and np.array_equal(B, ref[form][2]), v
    assert float((ref[0][0] - ref[1][0]).abs().max()) < 1e-12
    assert np.abs(ref[0][2] - ref[1][2]).max() < 1e-12 * np.abs(ref[0][2]).max()
    e = Ensemble.random(dims, seeds, 0.4, "fp64")
    b, ms = e.autotune(reps=2)
    assert 1 <= b <= 8 and ms > 0
    plaq, B = e.flow(0.01, 4, measure_every=2, use_graph=True)
    assert float((e.data - ref[0][0]).abs().max()) < 1e-12
