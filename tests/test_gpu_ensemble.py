"""Ensembles with uneven time extents in one launch per operation, against the
single-field path and the oracle.

The ensemble kernels run the same per-site code as the single-field kernels, so the
flowed fields and the measurements must agree with one-field-at-a-time runs to 1e-12
(FP64); the FP64 single-field path is itself checked against the oracle.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def FL():
    import flue_lat_amd as FL
    return FL


def test_ensemble_generator_matches_single_fields(FL):
    from flue_lat_amd.ensemble import Ensemble
    dims = [(4, 3, 3, 3), (7, 3, 4, 2), (3, 5, 3, 3)]
    seeds = [11, 12, 13]
    e = Ensemble.random(dims, seeds, 0.4, "fp64")
    for c in range(3):
        U = O.random_field(dims[c], seeds[c], 0.4)
        assert np.abs(e.field(c).to_complex().cpu().numpy() - U).max() < 1e-13


@pytest.mark.parametrize("use_graph", [False, True])
def test_ensemble_flow_matches_single_fields_fp64(FL, use_graph):
    from flue_lat_amd.ensemble import Ensemble
    dims = [(4, 4, 4, 4), (6, 4, 4, 4), (3, 5, 4, 4), (8, 4, 4, 4)]
    seeds = [4000 + c for c in range(len(dims))]
    e = Ensemble.random(dims, seeds, 0.4, "fp64")
    torch.cuda.synchronize()
    singles = [e.field(c).clone() for c in range(len(dims))]
    plaq, B = e.flow(0.01, 5, measure_every=2, use_graph=use_graph)
    torch.cuda.synchronize()
    assert plaq.shape == (3, len(dims)) and B.shape == (3, len(dims))
    for c, g in enumerate(singles):
        p0, _ = g.plaquette()
        b0 = g.magnetic()
        assert abs(plaq[0, c] - p0) < 1e-12 * abs(p0) and abs(B[0, c] - b0) < 1e-12 * abs(b0)
        g.wilson_flow(0.01, 2, use_graph=False)
        p2, _ = g.plaquette()
        b2 = g.magnetic()
        assert abs(plaq[1, c] - p2) < 1e-12 * abs(p2) and abs(B[1, c] - b2) < 1e-12 * abs(b2)
        g.wilson_flow(0.01, 3, use_graph=False)
        torch.cuda.synchronize()
        assert torch.allclose(e.field(c).data, g.data, rtol=0, atol=1e-12)


def test_ensemble_flow_against_oracle(FL):
    from flue_lat_amd.ensemble import Ensemble
    dims = [(4, 3, 3, 3), (5, 3, 3, 3)]
    seeds = [7, 8]
    e = Ensemble.random(dims, seeds, 0.3, "fp64")
    plaq, B = e.flow(0.02, 4, measure_every=4, use_graph=True)
    for c in range(2):
        W = O.wilson_flow(O.random_field(dims[c], seeds[c], 0.3), 0.02, 4)
        assert np.abs(e.field(c).to_complex().cpu().numpy() - W).max() < 1e-12
        assert abs(B[1, c] - O.magnetic(W)) < 1e-12 * O.magnetic(W)
        assert abs(plaq[1, c] - O.plaquette(W, 1, 4, 4)[0]) < 1e-12 * abs(plaq[1, c])


def c4_dims():
    # the batching workload: 32 configurations of 16^3 x Nt, Nt drawn (seeded) from {16,24,32,48,64,128}
    rng = np.random.default_rng(4000)
    nts = rng.choice([16, 24, 32, 48, 64, 128], size=32)
    return [(int(nt), 16, 16, 16) for nt in nts]


def test_ensemble_c4_shape_fp32(FL):
    # the batching workload in FP32 links: eps = 0.4, flow dt = 0.01 with plaquette / B every 10 steps,
    # here 20 of the 100 steps bench.py's ensemble leg runs
    from flue_lat_amd.ensemble import Ensemble
    dims = c4_dims()
    e = Ensemble.random(dims, [4000 + j for j in range(32)], 0.4, "fp32")
    plaq, B = e.flow(0.01, 20, measure_every=10, use_graph=True)
    vols = np.array([np.prod(d) for d in dims])
    P = plaq / (18.0 * vols)  # mean ReTr P / 3
    assert np.all(np.diff(P, axis=0) > 0)   # the flow smooths: plaquette rises
    assert np.all(np.diff(B, axis=0) < 0)   # and B falls
    # spot-check two members against the FP64 single-field path
    for c in (0, 17):
        g = FL.GaugeField.random(dims[c], 4000 + c, 0.4, "fp64")
        g.wilson_flow(0.01, 20, use_graph=False)
        assert abs(B[-1, c] - g.magnetic()) < 1e-5 * abs(g.magnetic())


def test_ensemble_c4_member_against_oracle_fp64(FL):
    # the whole 32-configuration ensemble in FP64, two RK3 steps in one graph; the members
    # with Nt = 16 and the first member of each other extent against the oracle's flow
    from flue_lat_amd.ensemble import Ensemble
    dims = c4_dims()
    seeds = [4000 + j for j in range(32)]
    e = Ensemble.random(dims, seeds, 0.4, "fp64")
    plaq, B = e.flow(0.01, 2, measure_every=2, use_graph=True)
    firsts = {}
    for c, d in enumerate(dims):
        firsts.setdefault(d[0], c)
    checked = sorted(set(firsts.values()) | {c for c, d in enumerate(dims) if d[0] == 16})
    for c in checked:
        if dims[c][0] > 32:
            continue   # the oracle's serial flow on Nt >= 48 takes too long for a test
        print(f"[c4] member {c} {dims[c]} against the oracle", flush=True)
        W = O.wilson_flow(O.random_field(dims[c], seeds[c], 0.4), 0.01, 2)
        assert np.abs(e.field(c).to_complex().cpu().numpy() - W).max() < 1e-12, c
        assert abs(B[1, c] - O.magnetic(W)) < 1e-12 * O.magnetic(W), c
        assert abs(plaq[1, c] - O.plaquette(W, 1, 4, 4)[0]) < 1e-12 * abs(plaq[1, c]), c
