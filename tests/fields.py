"""Synthetic gauge fields shared by the tests (numpy, complex128, Fortran order)."""
import numpy as np


def unit_field(dims):
    U = np.zeros(tuple(dims) + (4, 3, 3), np.complex128, order="F")
    for a in range(3):
        U[..., a, a] = 1.0
    return U


def abelian_flux(dims, mu, nu, n):
    """Constant U(1) flux in plane (mu, nu) (1-based), embedded as diag(e^{i phi}, e^{-i phi}, 1).

    Every counter-clockwise a x b loop in the (mu, nu) plane has phase theta*a*b with
    theta = 2 pi n / (N_mu N_nu); all other loops are trivial."""
    U = unit_field(dims)
    Nmu, Nnu = dims[mu - 1], dims[nu - 1]
    theta = 2 * np.pi * n / (Nmu * Nnu)
    idx = np.indices(tuple(dims))
    s = idx[mu - 1]
    r = idx[nu - 1]
    phi_nu = theta * s
    phi_mu = np.where(s == Nmu - 1, -theta * Nmu * r, 0.0)
    for d, phi in ((nu, phi_nu), (mu, phi_mu)):
        U[..., d - 1, 0, 0] = np.exp(1j * phi)
        U[..., d - 1, 1, 1] = np.exp(-1j * phi)
    return U, theta


def random_su3(rng, shape):
    """Haar-ish random SU(3) matrices via QR of complex Gaussians."""
    Z = rng.normal(size=shape + (3, 3)) + 1j * rng.normal(size=shape + (3, 3))
    Q, R = np.linalg.qr(Z)
    d = np.diagonal(R, axis1=-2, axis2=-1)
    Q = Q * (d / np.abs(d))[..., None, :]
    det = np.linalg.det(Q)
    Q = Q / (det ** (1.0 / 3.0))[..., None, None]
    return Q


def gauge_transform(U, G):
    """U_mu(x) -> G(x) U_mu(x) G(x+mu)^dagger; G has shape (NT,NX,NY,NZ,3,3)."""
    out = np.empty_like(U)
    for mu in range(4):
        Gs = np.roll(G, -1, axis=mu)
        out[..., mu, :, :] = G @ U[..., mu, :, :] @ np.conj(np.swapaxes(Gs, -1, -2))
    return np.asfortranarray(out)


def f5(theta):
    """F_11 / i of the 5-loop clover on an abelian flux theta per plaquette."""
    k5 = 1 / 180
    k1 = 19 / 9 - 55 * k5
    k2 = 1 / 36 - 16 * k5
    k3 = 64 * k5 - 32 / 45
    k4 = 1 / 15 - 6 * k5
    return (k1 * np.sin(theta) + k2 * np.sin(4 * theta) + k5 * np.sin(9 * theta)
            + k3 * np.sin(2 * theta) + k4 * np.sin(3 * theta))
