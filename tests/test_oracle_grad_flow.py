"""Pins the oracle's gradients and Wilson flow (no reference counterpart).

Gradient convention: D^a f = d/ds f(U_mu(x) -> exp(s T^a) U_mu(x)) at s = 0 with
T^a = -i lambda_a / 2, and grad = sum_a T^a D^a f, so D^a f = -2 Tr(T^a grad).
The analytic forms are pinned by central finite differences of the *local*
action (a global difference loses ~1e-5 to cancellation at h = 1e-6).
"""
import numpy as np
import pytest

import oracle as O
from tests.fields import gauge_transform, random_su3, unit_field

RNG = np.random.default_rng(777)
T = None


def gens():
    return O.generators()


def components(G):
    """D^a = -2 Tr(T^a G) for the 8 generators."""
    Tg = gens()
    return np.array([-2 * np.trace(Tg[a] @ G) for a in range(8)])


def assert_su3_algebra(G, tol=1e-13):
    assert np.abs(G + np.conj(np.swapaxes(G, -1, -2))).max() < tol
    assert np.abs(np.trace(G, axis1=-2, axis2=-1)).max() < tol


@pytest.fixture(scope="module")
def field():
    return O.random_field((4, 4, 4, 4), 2002, 0.3)


def test_wilson_grad_matches_fd(field):
    beta = 6.0
    G = O.wilson_grad(field, beta)
    assert_su3_algebra(G)
    for trial in range(6):
        c = [int(RNG.integers(1, 5)) for _ in range(4)]
        mu = int(RNG.integers(1, 5))
        comp = components(G[c[0] - 1, c[1] - 1, c[2] - 1, c[3] - 1, mu - 1])
        assert np.abs(comp.imag).max() < 1e-13
        for a in range(1, 9):
            fd = O.fd_link(field, 0, c, mu, a, h=1e-6, beta=beta)
            assert abs(fd - comp[a - 1].real) < 2e-8, (c, mu, a, fd, comp[a - 1])


def test_wilson_grad_scales_with_beta_and_vanishes_on_unit(field):
    G1 = O.wilson_grad(field, 1.0)
    G6 = O.wilson_grad(field, 6.0)
    assert np.abs(G6 - 6 * G1).max() < 1e-13
    assert np.abs(O.wilson_grad(unit_field((3, 3, 3, 3)), 6.0)).max() < 1e-15


def test_wilson_grad_gauge_covariant(field):
    G = random_su3(RNG, field.shape[:4])
    Ug = gauge_transform(field, G)
    g0 = O.wilson_grad(field, 6.0)
    g1 = O.wilson_grad(Ug, 6.0)
    # left-multiplicative derivative at U_mu(x) transforms with G(x) on both sides
    for mu in range(4):
        want = G @ g0[..., mu, :, :] @ np.conj(np.swapaxes(G, -1, -2))
        assert np.abs(g1[..., mu, :, :] - want).max() < 1e-12


@pytest.fixture(scope="module")
def small_field():
    return O.random_field((2, 4, 4, 5), 3003, 0.3)


@pytest.mark.slow
def test_b2_grad_matches_fd(small_field):
    U = small_field
    G = O.b2_grad(U)
    assert_su3_algebra(G, 1e-15)
    # temporal links do not enter the magnetic field
    assert np.abs(G[..., 0, :, :]).max() == 0.0
    rng = np.random.default_rng(5)
    for trial in range(5):
        c = [int(rng.integers(1, n + 1)) for n in U.shape[:4]]
        mu = int(rng.integers(2, 5))
        comp = components(G[c[0] - 1, c[1] - 1, c[2] - 1, c[3] - 1, mu - 1])
        scale = np.abs(comp).max()
        for a in range(1, 9):
            fd = O.fd_link(U, 1, c, mu, a, h=1e-6)
            assert abs(fd - comp[a - 1].real) < 1e-7 * scale + 1e-13, (c, mu, a, fd, comp[a - 1])


def test_b2_grad_gauge_covariant(small_field):
    U = small_field
    G = random_su3(RNG, U.shape[:4])
    g0 = O.b2_grad(U)
    g1 = O.b2_grad(gauge_transform(U, G))
    for mu in range(4):
        want = G @ g0[..., mu, :, :] @ np.conj(np.swapaxes(G, -1, -2))
        assert np.abs(g1[..., mu, :, :] - want).max() < 1e-14


# ---------------------------------------------------------------- Wilson flow

def test_flow_fixes_unit_field():
    U = unit_field((3, 3, 3, 3))
    W = O.wilson_flow(U, 0.05, 3)
    assert np.abs(W - U).max() < 1e-15


def test_flow_smooths_and_stays_su3(field):
    S0 = O.wilson_action(field, 6.0)
    W = O.wilson_flow(field, 0.02, 5)
    S1 = O.wilson_action(W, 6.0)
    assert S1 < 0.9 * S0
    M = W.reshape(-1, 4, 3, 3, order="F")
    M = np.moveaxis(W, (4,), (4,)).reshape(-1, 3, 3) if False else W[..., 1, :, :].reshape(-1, 3, 3)
    assert np.abs(M @ np.conj(np.swapaxes(M, -1, -2)) - np.eye(3)).max() < 1e-13


def test_flow_rate_is_minus_gradient_norm(field):
    # dS/dt = -(beta/6)^{-1} ... with Z = -g0^2 dS, g0^2 = 6/beta: dS/dt = sum_a D^a S * (Z^a)
    beta = 6.0
    eps = 1e-4
    S0 = O.wilson_action(field, beta)
    W = O.wilson_flow(field, eps, 1, reunit=False)
    dS = (O.wilson_action(W, beta) - S0) / eps
    G = O.wilson_grad(field, beta)
    # Z = -(6/beta) G and dS/dt = sum_links sum_a D^a S Z^a with Z^a = -2 Tr(T^a Z);
    # sum_a (-2Tr T^a G)(-2 Tr T^a Z) = -2 Tr(G Z) for G, Z in su(3)
    Z = -(6.0 / beta) * G
    rate = np.real(-2 * np.sum(np.trace(G @ Z, axis1=-2, axis2=-1)))
    assert abs(dS - rate) < 2e-3 * abs(rate)


def test_flow_gauge_covariant(field):
    G = random_su3(RNG, field.shape[:4])
    W0 = O.wilson_flow(field, 0.03, 2)
    W1 = O.wilson_flow(gauge_transform(field, G), 0.03, 2)
    np.testing.assert_allclose(W1, gauge_transform(W0, G), atol=1e-12)


def test_flow_rk3_is_third_order(field):
    # global error after fixed time t behaves like eps^3
    t = 0.12
    ref = O.wilson_flow(field, t / 48, 48)
    errs = []
    for n in (3, 6):
        W = O.wilson_flow(field, t / n, n)
        errs.append(np.abs(W - ref).max())
    ratio = errs[0] / errs[1]
    assert 6.0 < ratio < 10.0, errs
