"""torch.autograd front end of the hand-written adjoint kernels.

`magnetic(U)` and `wilson_action(U, beta)` take a complex gauge field tensor of shape
(NT, NX, NY, NZ, 4, 3, 3) (links U[t,x,y,z,mu] as in the reference's data array) and
return a 0-d float64 tensor whose backward runs the adjoint kernel on the GPU.

The kernels produce the Lie gradient G (traceless anti-Hermitian per link, with
D^a f = -2 Tr(T^a G) the derivative along exp(s T^a) U, T^a = -i lambda_a / 2).
torch's complex gradient of a real f is E = df/dRe U + i df/dIm U, with
df = Re sum conj(E) dU; restricted to the group (dU = X U, X in su(3)) the matching
tangent vector is E = 2 G U, which is what backward returns.  Derivatives along the
group are therefore exact; the components of E normal to SU(3) are zero.
"""
import numpy as np
import torch

from .field import GaugeField

SQRT3 = float(np.sqrt(3.0))


def _grad_complex(g: GaugeField):
    return g.to_complex()


def tangent_from_lie(G, U):
    """E = 2 G U for link arrays of shape (..., 3, 3)."""
    return 2.0 * torch.matmul(G.contiguous(), U.contiguous())


def lie_components(G):
    """D^a = -2 Tr(T^a G) for a = 1..8 from an su(3) array (..., 3, 3) -> (..., 8)."""
    g00, g11, g22 = G[..., 0, 0].imag, G[..., 1, 1].imag, G[..., 2, 2].imag
    out = torch.stack([
        -2 * G[..., 0, 1].imag, -2 * G[..., 0, 1].real, g11 - g00,
        -2 * G[..., 0, 2].imag, -2 * G[..., 0, 2].real,
        -2 * G[..., 1, 2].imag, -2 * G[..., 1, 2].real,
        -(g00 + g11 - 2 * g22) / SQRT3,
    ], dim=-1)
    return out


def generator_matrix(omega):
    """X = sum_a omega_a T^a for omega of shape (..., 8), T^a = -i lambda_a / 2."""
    o = omega.to(torch.float64)
    X = torch.zeros(o.shape[:-1] + (3, 3), dtype=torch.complex128, device=o.device)
    h = 0.5
    X[..., 0, 1] = h * (-1j * o[..., 0] - o[..., 1])
    X[..., 1, 0] = h * (-1j * o[..., 0] + o[..., 1])
    X[..., 0, 2] = h * (-1j * o[..., 3] - o[..., 4])
    X[..., 2, 0] = h * (-1j * o[..., 3] + o[..., 4])
    X[..., 1, 2] = h * (-1j * o[..., 5] - o[..., 6])
    X[..., 2, 1] = h * (-1j * o[..., 5] + o[..., 6])
    d8 = o[..., 7] / SQRT3
    X[..., 0, 0] = -1j * h * (o[..., 2] + d8)
    X[..., 1, 1] = -1j * h * (-o[..., 2] + d8)
    X[..., 2, 2] = -1j * h * (-2 * d8)
    return X


class _Magnetic(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, precision):
        g = GaugeField.from_complex(U.detach(), precision)
        ctx.save_for_backward(U)
        ctx.precision = precision
        B = g.magnetic(sync=False)
        return B.reshape(())

    @staticmethod
    def backward(ctx, grad_out):
        (U,) = ctx.saved_tensors
        g = GaugeField.from_complex(U.detach(), ctx.precision)
        G = _grad_complex(g.magnetic_grad()).to(U.device)
        E = tangent_from_lie(G, U.detach().to(torch.complex128))
        return (grad_out.to(torch.float64) * E).to(U.dtype), None


class _WilsonAction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, beta, precision):
        g = GaugeField.from_complex(U.detach(), precision)
        ctx.save_for_backward(U)
        ctx.beta = float(beta)
        ctx.precision = precision
        return g.wilson_action(beta, sync=False).reshape(())

    @staticmethod
    def backward(ctx, grad_out):
        (U,) = ctx.saved_tensors
        g = GaugeField.from_complex(U.detach(), ctx.precision)
        G = _grad_complex(g.wilson_grad(ctx.beta)).to(U.device)
        E = tangent_from_lie(G, U.detach().to(torch.complex128))
        return (grad_out.to(torch.float64) * E).to(U.dtype), None, None


def magnetic(U, precision="fp64"):
    """B of the reference's `magnetic` (src/wloops.F90:479-544), differentiable in U."""
    return _Magnetic.apply(U, precision)


def wilson_action(U, beta=6.0, precision="fp64"):
    """S = beta sum_p (1 - ReTr P/3), differentiable in U."""
    return _WilsonAction.apply(U, beta, precision)
