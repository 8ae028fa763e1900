"""Parity of the gluon-propagator kernels (k_gluon.hip) with the oracle.

calc_mom_space_scalarD and scalarGluonProp through the reference-style host entries and
through the device path (GaugeField.gluon_propagator: 4-D FFT on the device, then the
momentum-space kernel).  Tolerances: 1e-12 relative to max |D| for FP64 inputs; 1e-5 for a
field held in FP32 on the device path (the transform then runs in complex64).  scalarGluonProp is compared at
volumes where the reference's term-by-term volume sum, restated literally by the oracle,
stays within 1e-12 of V times the term.
"""
import numpy as np
import pytest

import oracle as O
from tests.fields import random_su3

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def FL():
    import flue_lat_amd as FL
    return FL


def close(got, want, tol):
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(np.asarray(got) - want).max() / scale
    assert err <= tol, err


@pytest.mark.parametrize("shape,mu_start,xi", [
    ((5, 4, 4, 4), 1, 1.0),      # NT = 4, NS = 3 in the reference's bounds: odd and even extents
    ((6, 5, 5, 5), 0, 1.7),
    ((4, 7, 7, 7), 2, 0.5),
    ((9, 6, 6, 6), 0, 2.0),
])
def test_host_mom_space_matches_oracle(FL, shape, mu_start, xi):
    print(f"calc_mom_space_scalard_c {shape} mu_start={mu_start}", flush=True)
    rng = np.random.default_rng(sum(shape) + mu_start)
    U = rng.normal(size=shape + (4, 3, 3)) + 1j * rng.normal(size=shape + (4, 3, 3))
    D = FL.gluonCalc.calc_mom_space_scalarD(U, mu_start, xi)
    NT, NS = shape[0] - 1, shape[1] - 1
    assert D.shape == (NT // 2 + 1,) + (NS // 2 + 1,) * 3
    want = O.mom_scalar_d(U, mu_start, xi, dext=D.shape)
    close(D, want, 1e-12)
    assert np.abs(D.imag).max() == 0.0


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("mu_start", [0, 1])
def test_device_propagator_matches_oracle_on_fft(FL, prec, tol, mu_start):
    dims = (8, 6, 5, 4)
    print(f"gluon_propagator {dims} {prec} mu_start={mu_start}", flush=True)
    U = O.random_field(dims, 4321, 0.6)
    g = FL.GaugeField.from_complex(U, "fp64").astype(prec)
    D = g.gluon_propagator(mu_start=mu_start, xi=1.4).cpu().numpy()
    # the oracle sees the field the device holds, transformed by numpy
    Uh = g.to_complex().cpu().numpy()
    want = O.mom_scalar_d(np.fft.fftn(Uh, axes=(0, 1, 2, 3)), mu_start, 1.4)
    assert D.shape == want.shape == (5, 4, 3, 3)
    close(D, want, tol)


def test_device_propagator_of_unit_field_vanishes(FL):
    U = np.zeros((4, 4, 4, 4, 4, 3, 3), complex)
    U[...] = np.eye(3)
    D = FL.GaugeField.from_complex(U).gluon_propagator(mu_start=0, xi=2.0)
    assert float(D.abs().max()) < 1e-10


@pytest.mark.parametrize("NT,NS", [(3, 4), (4, 4), (2, 5)])
def test_host_scalar_gluon_prop_matches_oracle(FL, NT, NS):
    print(f"scalargluonprop_c NT={NT} NS={NS}", flush=True)
    rng = np.random.default_rng(NT * 10 + NS)
    U = np.asfortranarray(random_su3(rng, (NT, NS, NS, NS, 4)))
    D = FL.gluonCalc.scalarGluonProp(U, 1.3)
    want = O.scalar_gluon_prop(U, 1.3)
    assert D.shape == (NS,) * 4
    close(D, want, 1e-12)
    assert np.all(D[NT:] == 0)


def test_device_scalar_gluon_prop_fp32_field(FL):
    NT, NS = 4, 4
    U = O.random_field((NT, NS, NS, NS), 99, 0.4)
    g = FL.GaugeField.from_complex(U).astype("fp32")
    D = g.scalar_gluon_propagator().cpu().numpy()
    # the oracle sees the same FP32-rounded links, so the FP64 bound applies
    want = O.scalar_gluon_prop(g.to_complex().cpu().numpy(), 1.0)
    close(D, want, 1e-12)
    with pytest.raises(FL.FlueError):
        FL.GaugeField.random((5, 4, 4, 4), 1, 0.3).scalar_gluon_propagator()


def test_device_propagator_of_time_tiled_field(FL):
    # a field repeated k times along t transforms to k U(q') at q_t = k q'_t and to 0 at the
    # other t momenta; with the 1/V of the prefactor, D(k q'_t, q_s) = k D_small(q'_t, q_s)
    NT, NS, k = 6, 8, 8
    U = O.random_field((NT, NS, NS, NS), 2468, 0.5)
    g = FL.GaugeField.from_complex(U)
    VS = NS ** 3
    big = FL.GaugeField((NT * k, NS, NS, NS), "fp64", _data=g.data.view(36, NT, 2 * VS).repeat(1, k, 1).reshape(-1))
    for mu_start in (0, 1):
        Db = big.gluon_propagator(mu_start=mu_start, xi=1.2).cpu().numpy()
        Ds = O.mom_scalar_d(np.fft.fftn(U, axes=(0, 1, 2, 3)), mu_start, 1.2)
        assert Db.shape == (NT * k // 2 + 1, NS // 2 + 1, NS // 2 + 1, NS // 2 + 1)
        scale = np.abs(Ds).max() * k
        close(Db[::k], k * Ds, 1e-12)
        mask = np.ones(Db.shape[0], bool)
        mask[::k] = False
        assert np.abs(Db[mask]).max() <= 1e-12 * scale


def test_device_scalar_gluon_prop_closed_form_at_size(FL):
    # U_mu(x) = diag(e^{ia}, e^{-ia}, 1): D(x) = -(V^3 sin^2 a / 24) sum_mu cos(2 phi_mu) with
    # phi_mu = c_mu * 2 pi c_mu / N_mu for the 1-based coordinate c (NT = 12 < NS = 16)
    NT, NS, a = 12, 16, 0.21
    U = np.zeros((NT, NS, NS, NS, 4, 3, 3), complex)
    U[...] = np.diag([np.exp(1j * a), np.exp(-1j * a), 1.0])
    D = FL.GaugeField.from_complex(U).scalar_gluon_propagator().cpu().numpy()
    V = NT * NS ** 3
    c = [np.arange(1, n + 1, dtype=float) for n in (NT, NS, NS, NS)]
    N = (NT, NS, NS, NS)
    cs = [np.cos(2 * (ci * 2 * np.pi * ci / n)) for ci, n in zip(c, N)]
    s = cs[0][:, None, None, None] + cs[1][None, :, None, None] + cs[2][None, None, :, None] + cs[3][None, None, None, :]
    want = np.zeros((NS,) * 4)
    want[:NT] = -(V ** 3) * np.sin(a) ** 2 / 24.0 * s
    close(D, want, 1e-12)
