"""Committed fixtures (tests/golden/observables.json, written by tests/golden/make_golden.py
from the oracle): the seeded generator and the observables on fixed inputs, checked for
the oracle on the CPU and for the HIP path on the GPU."""
import json
import os

import numpy as np
import pytest

import oracle as O
from tests.golden.make_golden import checksums

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "observables.json")) as fh:
    GOLD = json.load(fh)
PROP = GOLD.pop("scalar_gluon_prop")


def mom_summary(D):
    D = np.asarray(D).real
    return {"sum": float(D.sum()), "q0000": float(D[0, 0, 0, 0]), "q1111": float(D[1, 1, 1, 1]),
            "qmax": float(D[-1, -1, -1, -1])}


def close(a, b, tol):
    return abs(a - b) <= tol * max(1.0, abs(b))


@pytest.mark.parametrize("name", sorted(GOLD))
def test_oracle_reproduces_fixtures(name):
    e = GOLD[name]
    U = O.random_field(tuple(e["dims"]), e["seed"], e["eps"])
    for k, v in checksums(U).items():
        assert close(v, e["checksums"][k], 1e-13), k
    s, n = O.plaquette(U, 1, 4, 4)
    assert n == e["plaq_total"][1] and close(s, e["plaq_total"][0], 1e-13)
    assert close(O.magnetic(U), e["magnetic"], 1e-13)
    s, n = O.polyakov(U)
    assert n == e["polyakov"][1] and close(s, e["polyakov"][0], 1e-13)
    Uq = np.fft.fftn(U, axes=(0, 1, 2, 3))
    for key, (mu_start, xi) in (("gluon_mom_d_mu1", (1, 1.0)), ("gluon_mom_d_mu0_xi15", (0, 1.5))):
        got = mom_summary(O.mom_scalar_d(Uq, mu_start, xi))
        for k, v in e[key].items():
            assert close(got[k], v, 1e-12), (key, k)


def test_oracle_reproduces_scalar_prop_fixture():
    U = O.random_field(tuple(PROP["dims"]), PROP["seed"], PROP["eps"])
    D = O.scalar_gluon_prop(U, 1.0)
    assert close(float(D.sum()), PROP["sum"], 1e-12)
    assert close(float(D[0, 0, 0, 0]), PROP["d0000"], 1e-12) and close(float(D[1, 2, 3, 1]), PROP["d1231"], 1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(GOLD))
def test_hip_path_matches_fixtures(name):
    import flue_lat_amd as FL
    e = GOLD[name]
    g = FL.GaugeField.random(tuple(e["dims"]), e["seed"], e["eps"], "fp64")
    U = g.to_complex().cpu().numpy()
    for k, v in checksums(U).items():
        assert close(v, e["checksums"][k], 1e-12), k
    for key, args in (("plaq_total", (1, 4, 4)), ("plaq_spatial", (2, 4, 4)), ("plaq_temporal", (1, 1, 4))):
        s, n = g.plaquette(*args)
        assert n == e[key][1] and close(s, e[key][0], 1e-12), key
    s, n = g.polyakov()
    assert n == e["polyakov"][1] and close(s, e["polyakov"][0], 1e-12)
    assert close(g.magnetic(), e["magnetic"], 1e-12)
    for key, v in e["wmunu"].items():
        mu, nu = int(key[0]), int(key[1])
        R, T = (int(x) for x in key[3:].split("x"))
        assert close(g.wilson_loop(mu, nu, R, T), v, 1e-12), key
    F = g.fmunu(2, 3)[0, 0, 0, 0].cpu().numpy()
    want = np.array([complex(r, i) for r, i in e["loop5_23_site0"]]).reshape(3, 3)
    assert np.abs(F - want).max() < 1e-12 * max(1.0, np.abs(want).max())
    assert close(g.wilson_action(6.0), e["wilson_action_beta6"], 1e-12)
    for key, (mu_start, xi) in (("gluon_mom_d_mu1", (1, 1.0)), ("gluon_mom_d_mu0_xi15", (0, 1.5))):
        got = mom_summary(g.gluon_propagator(mu_start, xi).cpu().numpy())
        scale = max(abs(v) for v in e[key].values())
        for k, v in e[key].items():
            assert abs(got[k] - v) <= 1e-12 * scale, (key, k)
    # FP32 links: within 1e-6 of the FP64 fixture
    assert close(g.astype("fp32").magnetic(), e["magnetic"], 1e-6)


@pytest.mark.gpu
def test_hip_scalar_prop_matches_fixture():
    import flue_lat_amd as FL
    U = FL.GaugeField.random(tuple(PROP["dims"]), PROP["seed"], PROP["eps"], "fp64").to_complex().cpu().numpy()
    D = FL.gluonCalc.scalarGluonProp(U, 1.0)
    scale = abs(PROP["sum"])
    assert abs(float(D.sum()) - PROP["sum"]) <= 1e-12 * scale
    assert abs(float(D[0, 0, 0, 0]) - PROP["d0000"]) <= 1e-12 * scale
    assert abs(float(D[1, 2, 3, 1]) - PROP["d1231"]) <= 1e-12 * scale
