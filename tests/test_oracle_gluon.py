"""The gluon-propagator row on the CPU: the oracle's restatements of calc_mom_space_scalarD,
scalarGluonProp, Q_Average and cone_cut pinned by closed forms worked out by hand from the
reference's text (src/gluonCalc/gluonProp.f90, src/gluonAna/gpManip.f90,
src/SU3MatrixOps.f90:94-122), and the library's host implementations of Q_Average and
cone_cut checked against them.

The closed forms are chosen so that the reference's quirks show: the trace third is taken
off all nine entries (not the diagonal), the phase of direction mu uses qhat(mu) of a 1-based
qhat, the prefactor counts qx twice, and scalarGluonProp's phase is coord * qhat with a
1-based coordinate.
"""
import math

import numpy as np
import pytest

import oracle as O

LAM1 = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], complex)


def single_momentum_field(E, q0, mu, M):
    """Uq with the matrix M at momentum q0 in direction mu and zero elsewhere."""
    U = np.zeros(tuple(E) + (4, 3, 3), complex, order="F")
    U[tuple(q0) + (mu,)] = M
    return U


def prefac(E, q, mu_start):
    V = float(np.prod(E))
    four = (mu_start == 1 and q[1] + q[2] + q[1] < 1) or (mu_start == 0 and sum(q) + q[1] < 1)
    return 2.0 / (8.0 * V * (4.0 if four else 3.0))


@pytest.mark.parametrize("mu,mu_start,xi", [(1, 1, 1.0), (2, 0, 1.3), (3, 1, 1.0), (0, 0, 1.7)])
def test_mom_scalar_single_lambda1(mu, mu_start, xi):
    # U(q0) = i c lambda_1: TCS(U(q0), 0) = i c lambda_1, A+ = -(c/2) e^{-i th/2} lambda_1, so
    # c_1(A+) = c_1(A-) = -c cos(th/2) and D(q0) = prefac * w * c^2 cos^2(th/2)
    E = (7, 6, 5, 6)
    q0 = (2, 1, 2, 3)
    c = 0.8
    U = single_momentum_field(E, q0, mu, 1j * c * LAM1)
    D = O.mom_scalar_d(U, mu_start, xi)
    th_dir = mu - 1 if mu >= 1 else 0
    th = 2 * np.pi * q0[th_dir] / E[th_dir]
    w = xi * xi if mu == 0 else 1.0
    expect = np.zeros(D.shape)
    expect[q0] = prefac(E, q0, mu_start) * w * c * c * math.cos(th / 2) ** 2
    np.testing.assert_allclose(D.real, expect, rtol=0, atol=1e-15 * abs(expect).max())
    assert np.abs(D.imag).max() == 0.0


def test_mom_scalar_trace_taken_from_all_entries():
    # U(q0) = i c I: Tr = 3ic comes off all nine entries, leaving -i c on the off-diagonal,
    # A+ = (c/2) e^{-i th/2} (J - I) and c_1 = c_4 = c_6 = c cos(th/2): D = 3 c^2 cos^2(th/2)
    # (a projection on the diagonal only would give 0)
    E = (5, 4, 4, 4)
    q0 = (1, 2, 1, 1)
    c = 1.1
    U = single_momentum_field(E, q0, 2, 1j * c * np.eye(3))
    D = O.mom_scalar_d(U, 1, 1.0)
    th = 2 * np.pi * q0[1] / E[1]
    assert D[q0].real == pytest.approx(prefac(E, q0, 1) * 3 * c * c * math.cos(th / 2) ** 2, rel=1e-14)
    D[q0] = 0
    assert np.abs(D).max() == 0.0


def test_mom_scalar_negative_momentum_and_prefactor():
    # U(-q0) = i c lambda_1 with q0 in range: D(q0) sees it as U_neg; TCS(0, i c l1) = i c l1.
    # q0 = (0, 0, 0, 2) checks the prefactor with qz left out of the mu_start = 1 test
    E = (4, 4, 4, 6)
    q0 = (0, 0, 0, 2)
    mq = tuple((e - q) % e for e, q in zip(E, q0))
    c = 0.6
    U = single_momentum_field(E, mq, 1, 1j * c * LAM1)
    D = O.mom_scalar_d(U, 1, 1.0)
    th = 2 * np.pi * q0[0] / E[0]
    assert prefac(E, q0, 1) == 2.0 / (8.0 * np.prod(E) * 4.0)
    assert D[q0].real == pytest.approx(prefac(E, q0, 1) * c * c * math.cos(th / 2) ** 2, rel=1e-14)


def test_mom_scalar_of_unit_field_vanishes():
    # for U = 1 everywhere, U(q) = U(-q)^+ and every TCS is 0
    U = np.zeros((4, 4, 4, 4, 4, 3, 3), complex, order="F")
    U[..., :, :] = np.eye(3)
    Uq = np.fft.fftn(U, axes=(0, 1, 2, 3))
    assert np.abs(O.mom_scalar_d(Uq, 0, 2.0)).max() < 1e-12


def test_scalar_gluon_prop_diagonal_field():
    # U_mu(x) = diag(e^{ia}, e^{-ia}, 1): T2 - T1^+ = V e2 (U - U^+), A = -(1/4)(T2 - T1^+),
    # c_3 = -i V sin(a) e2 and D(x) = -(V^3 sin^2 a / 24) sum_mu cos(2 phi_mu) with
    # phi_mu = c_mu * 2 pi c_mu / N_mu for the 1-based coordinate c
    NT, NS = 3, 4
    a = 0.37
    U = np.zeros((NT, NS, NS, NS, 4, 3, 3), complex, order="F")
    U[...] = np.diag([np.exp(1j * a), np.exp(-1j * a), 1.0])
    D = O.scalar_gluon_prop(U, 1.0)
    V = NT * NS ** 3
    expect = np.zeros((NS,) * 4)
    for t in range(NT):
        for x in range(NS):
            for y in range(NS):
                for z in range(NS):
                    c = (t + 1, x + 1, y + 1, z + 1)
                    N = (NT, NS, NS, NS)
                    s = sum(math.cos(2 * (c[m] * 2 * math.pi * c[m] / N[m])) for m in range(4))
                    expect[t, x, y, z] = -(V ** 3) * math.sin(a) ** 2 / 24.0 * s
    np.testing.assert_allclose(D, expect, rtol=0, atol=1e-12 * abs(expect).max())
    assert np.all(D[NT:] == 0)


def test_q_average_worked_example():
    # maxq = 3, dq = 1: bins (0,1) and (1,2) are kept; the loop stops before (2,3) since
    # 2 + 1 >= 3, and values on the bin edges fall in no bin
    q = np.array([0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 0.25])
    D = np.arange(7.0) + 10
    dout, qout, cnt = O.q_average(q, D, 3)
    assert cnt == 3
    np.testing.assert_array_equal(qout[:2], [0.375, 1.5])
    np.testing.assert_array_equal(dout[:2], [10.0, 10.0])
    assert np.all(qout[2:] == 0) and np.all(dout[2:] == 0)


def test_cone_cut_worked_example():
    # (1,1,1,1): on the diagonal, r = theta = 0; (1,0,0,0): theta = pi/3, r = sin(pi/3);
    # (0,0,0,0): all equal; (3,0,0,0): |q| = 3 > IRCut, theta = pi/3 > angle = 0.9 is cut;
    # (1,1,1,1) with xi = 2 first component leaves the diagonal
    Q = np.array([[1, 1, 1, 1], [1, 0, 0, 0], [0, 0, 0, 0], [3, 0, 0, 0], [0.5, 1, 1, 1]], float)
    D = np.array([1.0, 2, 3, 4, 5])
    D4 = D * 10
    qo, do, d4o, n = O.cone_cut(Q, D, D4, 0.9, 1.0, 1.5, 0.9)
    assert n == 4
    np.testing.assert_array_equal(do[:4], [1, 2, 3, 5])
    np.testing.assert_array_equal(d4o[:4], [10, 20, 30, 50])
    np.testing.assert_array_equal(qo[:4], Q[[0, 1, 2, 4]])
    # r(1,0,0,0) = sin(pi/3) = 0.866 > 0.8 drops it
    _, do, _, n = O.cone_cut(Q, D, D4, 0.9, 1.0, 1.5, 0.8)
    assert n == 3 and list(do[:3]) == [1, 3, 5]
    # xi = 2 scales the first component after the norm is taken: (0.5,1,1,1) lands on the
    # diagonal; (1,1,1,1) leaves it with acos(2.5/2) = NaN and is dropped; (1,0,0,0) and
    # (3,0,0,0) reach acos(1) = 0 and are kept
    _, do, _, n = O.cone_cut(Q, D, D4, 0.9, 2.0, 1.5, 0.9)
    assert list(do[:n]) == [2, 3, 4, 5]


def test_host_q_average_matches_oracle():
    import flue_lat_amd as FL
    rng = np.random.default_rng(5)
    for n, slices in ((1, 1), (7, 3), (200, 17), (500, 1000)):
        q = rng.uniform(0, 4, n)
        q[rng.integers(0, n, n // 5)] = 0.0
        D = rng.normal(size=n)
        dout, qout, cnt = FL.gluonAna.q_average(n, q, D, slices)
        ed, eq, ec = O.q_average(q, D, slices)
        assert cnt == ec
        np.testing.assert_array_equal(qout, eq)
        np.testing.assert_array_equal(dout, ed)


def test_host_cone_cut_matches_oracle():
    import flue_lat_amd as FL
    rng = np.random.default_rng(6)
    n = 400
    Q = rng.integers(0, 8, (n, 4)).astype(float)
    Q[::7] = Q[::7, :1]          # points on the diagonal
    D, D4 = rng.normal(size=n), rng.normal(size=n)
    for angle, xi, ircut, irr in ((np.pi / 2, 1.0, 0.0, 3.0), (0.3, 1.0, 4.0, 2.0), (0.6, 1.5, 2.0, 2.5)):
        qo, do, d4o, cnt = FL.gluonAna.cone_cut(n, 3, Q, D, D4, 8, 8, angle, xi, ircut, irr)
        eq, ed, ed4, ec = O.cone_cut(Q, D, D4, angle, xi, ircut, irr)
        assert cnt == ec and cnt > 0
        np.testing.assert_array_equal(qo, eq)
        np.testing.assert_array_equal(do, ed)
        np.testing.assert_array_equal(d4o, ed4)


def test_host_analysis_empty_and_edge_inputs():
    import flue_lat_amd as FL
    from flue_lat_amd._lib import FlueError
    # no momenta: nothing kept
    qo, do, d4o, n = FL.gluonAna.cone_cut(0, 1, np.zeros((0, 4)), np.zeros(0), np.zeros(0), 4, 4, 1.0, 1.0, 0.0, 1.0)
    assert n == 0 and qo.shape == (0, 4)
    # a single value: maxq = q, one slice; the bin (0, q) excludes q itself
    dout, qout, cnt = FL.gluonAna.q_average(1, [2.0], [5.0], 1)
    ed, eq, ec = O.q_average(np.array([2.0]), np.array([5.0]), 1)
    assert cnt == ec == 1 and np.array_equal(qout, eq) and np.array_equal(dout, ed)
    # all zero momenta: dq = 0, the loop ends after its first pass
    dout, qout, cnt = FL.gluonAna.q_average(3, [0.0, 0.0, 0.0], [1.0, 2.0, 3.0], 4)
    assert cnt == 1 and not qout.any()
    # a NaN, or a negative maximum with one slice, would keep the reference's loop going
    for q in ([1.0, float("nan")], [-1.0, -2.0]):
        with pytest.raises(FlueError):
            FL.gluonAna.q_average(2, q, [0.0, 0.0], 1)
    # with two slices its first exit test holds: one empty pass
    dout, qout, cnt = FL.gluonAna.q_average(2, [-1.0, -2.0], [0.0, 0.0], 2)
    ed, eq, ec = O.q_average(np.array([-1.0, -2.0]), np.array([0.0, 0.0]), 2)
    assert cnt == ec == 1 and not qout.any()
