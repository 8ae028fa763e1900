"""GPU gradients (Wilson action, B^2), their torch.autograd front end, and the Wilson flow,
against the CPU oracle.

Tolerances: FP64 gradient fields 1e-12 relative to the largest entry; finite
differences (h = 1e-6, central, on the local action) 2e-8 absolute for the Wilson
action and 1e-7 relative for B; FP64 flow 1e-12 absolute per link entry after 5 steps;
graph and stream replays bit-identical.
"""
import numpy as np
import pytest

import oracle as O
from tests.fields import gauge_transform, random_su3, unit_field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def FL():
    import flue_lat_amd as FL
    return FL


def comps(G):
    Tg = O.generators()
    return np.array([(-2 * np.trace(Tg[a] @ G)).real for a in range(8)])


def gfield(FL, U, prec="fp64"):
    return FL.GaugeField.from_complex(U, prec)


def grad_np(g):
    return g.to_complex().cpu().numpy()


# ------------------------------------------------------------------ Wilson action

def test_wilson_action_matches_oracle(FL):
    U = O.random_field((4, 3, 5, 4), 2024, 0.3)
    S = gfield(FL, U).wilson_action(6.0)
    So = O.wilson_action(U, 6.0)
    assert abs(S - So) < 1e-12 * abs(So)


def test_wilson_grad_matches_oracle(FL):
    U = O.random_field((4, 3, 5, 4), 2025, 0.3)
    G = grad_np(gfield(FL, U).wilson_grad(6.0))
    Go = O.wilson_grad(U, 6.0)
    assert np.abs(G - Go).max() < 1e-12 * np.abs(Go).max()
    G32 = grad_np(gfield(FL, U, "fp32").wilson_grad(6.0))
    assert np.abs(G32 - Go).max() < 1e-5 * np.abs(Go).max()


def test_wilson_grad_fd_c2(FL):
    # gradient check workload: 16^3x32, eps = 0.3, beta = 6, 64 random links, 8 Gell-Mann directions
    dims = (32, 16, 16, 16)
    g = FL.GaugeField.random(dims, 1002, 0.3)
    U = np.asfortranarray(g.to_complex().cpu().numpy())
    G = grad_np(g.wilson_grad(6.0))
    rng = np.random.default_rng(64)
    for _ in range(64):
        c = [int(rng.integers(0, n)) for n in dims]
        mu = int(rng.integers(0, 4))
        cg = comps(G[c[0], c[1], c[2], c[3], mu])
        for a in range(1, 9):
            fd = O.fd_link(U, 0, [x + 1 for x in c], mu + 1, a, h=1e-6, beta=6.0)
            assert abs(fd - cg[a - 1]) < 2e-8, (c, mu, a, fd, cg[a - 1])


# ------------------------------------------------------------------ B^2

def test_b2_grad_matches_oracle(FL):
    U = O.random_field((2, 4, 4, 5), 3003, 0.3)
    Go = O.b2_grad(U)
    G = grad_np(gfield(FL, U).magnetic_grad())
    scale = np.abs(Go).max()
    assert np.abs(G - Go).max() < 1e-12 * scale
    G32 = grad_np(gfield(FL, U, "fp32").magnetic_grad())
    assert np.abs(G32 - Go).max() < 1e-5 * scale


def test_b2_grad_matches_oracle_uneven(FL):
    U = O.random_field((3, 5, 7, 6), 3004, 0.35)
    Go = O.b2_grad(U)
    G = grad_np(gfield(FL, U).magnetic_grad())
    assert np.abs(G - Go).max() < 1e-12 * np.abs(Go).max()


def test_b2_grad_fd_c2(FL):
    dims = (32, 16, 16, 16)
    g = FL.GaugeField.random(dims, 1002, 0.3)
    U = np.asfortranarray(g.to_complex().cpu().numpy())
    G = grad_np(g.magnetic_grad())
    rng = np.random.default_rng(65)
    for _ in range(64):
        c = [int(rng.integers(0, n)) for n in dims]
        mu = int(rng.integers(0, 4))
        cg = comps(G[c[0], c[1], c[2], c[3], mu])
        if mu == 0:
            assert np.abs(cg).max() == 0.0
            continue
        scale = np.abs(cg).max()
        for a in range(1, 9):
            fd = O.fd_link(U, 1, [x + 1 for x in c], mu + 1, a, h=1e-6)
            assert abs(fd - cg[a - 1]) < 1e-7 * scale + 1e-15, (c, mu, a, fd, cg[a - 1])


def test_b2_grad_gauge_covariant(FL):
    U = O.random_field((2, 5, 4, 4), 77, 0.3)
    Gt = random_su3(np.random.default_rng(3), U.shape[:4])
    g0 = grad_np(gfield(FL, U).magnetic_grad())
    g1 = grad_np(gfield(FL, gauge_transform(U, Gt)).magnetic_grad())
    for mu in range(4):
        want = Gt @ g0[..., mu, :, :] @ np.conj(np.swapaxes(Gt, -1, -2))
        assert np.abs(g1[..., mu, :, :] - want).max() < 1e-13


# ------------------------------------------------------------------ autograd

def test_autograd_magnetic_directional_derivative(FL):
    from flue_lat_amd import autograd as AG
    U = O.random_field((2, 4, 5, 4), 99, 0.3)
    Ut = torch.from_numpy(np.ascontiguousarray(U)).cuda().requires_grad_(True)
    B = AG.magnetic(Ut)
    assert abs(B.item() - O.magnetic(U)) < 1e-12 * O.magnetic(U)
    (E,) = torch.autograd.grad(B, Ut)
    # derivative along U -> exp(s X) U for a random su(3) field X
    rng = np.random.default_rng(5)
    om = torch.from_numpy(rng.normal(size=U.shape[:5] + (8,))).cuda()
    X = AG.generator_matrix(om)
    Uc = Ut.detach()
    dir_grad = torch.sum((torch.conj(E) * torch.matmul(X, Uc)).real).item()
    h = 1e-5
    Up = torch.matrix_exp(h * X) @ Uc
    Um = torch.matrix_exp(-h * X) @ Uc
    fd = (AG.magnetic(Up).item() - AG.magnetic(Um).item()) / (2 * h)
    assert abs(dir_grad - fd) < 1e-6 * abs(fd)


def test_autograd_wilson_components(FL):
    from flue_lat_amd import autograd as AG
    U = O.random_field((4, 3, 4, 3), 11, 0.3)
    Ut = torch.from_numpy(np.ascontiguousarray(U)).cuda().requires_grad_(True)
    S = AG.wilson_action(Ut, 5.5)
    S.backward()
    E = Ut.grad
    # D^a = Re Tr(T^a U E^+) must equal -2 Tr(T^a G) of the oracle's Lie gradient
    Go = O.wilson_grad(U, 5.5)
    T = torch.from_numpy(O.generators()).cuda()
    UE = torch.matmul(Ut.detach(), torch.conj(E.transpose(-1, -2)).contiguous())
    # Tr(T^a UE) = sum_ji UE_ji T^a_ij
    TT = T.transpose(-1, -2).contiguous().reshape(8, 9).transpose(0, 1).contiguous()
    D = torch.matmul(UE.reshape(-1, 9), TT).real.reshape(U.shape[:4] + (4, 8)).cpu().numpy()
    Dp = np.stack([[comps(Go[idx][mu]) for mu in range(4)] for idx in np.ndindex(*U.shape[:4])]).reshape(
        U.shape[:4] + (4, 8))
    assert np.abs(D - Dp).max() < 1e-11


# ------------------------------------------------------------------ Wilson flow

def test_flow_matches_oracle(FL):
    U = O.random_field((4, 4, 3, 5), 404, 0.3)
    W = O.wilson_flow(U, 0.02, 5)
    for use_graph in (False, True):
        g = gfield(FL, U)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())  # the import ran on the current stream
        with torch.cuda.stream(s):
            g.wilson_flow(0.02, 5, use_graph=use_graph)
        torch.cuda.synchronize()
        assert np.abs(g.to_complex().cpu().numpy() - W).max() < 1e-12


def test_flow_graph_replays_match_eager(FL):
    U = O.random_field((4, 4, 4, 4), 405, 0.35)
    ref = gfield(FL, U)
    ref.wilson_flow(0.01, 4, use_graph=False)
    torch.cuda.synchronize()
    want = ref.data.clone()
    for rep in range(3):
        g = gfield(FL, U)
        _ = torch.ones(1000, device="cuda").sum()  # unrelated eager work between replays
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            g.wilson_flow(0.01, 4, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(g.data, want), rep


def test_flow_unit_field_and_smoothing(FL):
    g = gfield(FL, unit_field((3, 3, 3, 3)))
    g.wilson_flow(0.05, 3, use_graph=False)
    assert np.abs(g.to_complex().cpu().numpy() - unit_field((3, 3, 3, 3))).max() < 1e-14
    g = FL.GaugeField.random((8, 6, 6, 6), 7, 0.4)
    S0 = g.wilson_action(6.0)
    g.wilson_flow(0.01, 10, use_graph=False)
    assert g.wilson_action(6.0) < 0.8 * S0


def test_flow_fp32_close_to_fp64(FL):
    U = O.random_field((4, 4, 4, 4), 406, 0.3)
    W = O.wilson_flow(U, 0.02, 5)
    g = gfield(FL, U, "fp32")
    g.wilson_flow(0.02, 5, use_graph=False)
    torch.cuda.synchronize()
    assert np.abs(g.to_complex().cpu().numpy() - W).max() < 2e-6


def test_flow_stage_entry_composes_to_one_step(FL):
    # three calls of the single-stage entry (buffers ping-ponged, one shared accumulator)
    # make one RK3 step, as the oracle's flow does
    import ctypes
    from flue_lat_amd import _lib
    U = O.random_field((4, 5, 4, 3), 407, 0.3)
    W = O.wilson_flow(U, 0.03, 1)
    g = gfield(FL, U)
    bufs = [g.data, torch.empty_like(g.data)]
    acc = torch.zeros(32 * g.V, dtype=torch.float64, device=g.device)
    dims = (ctypes.c_int32 * 4)(*g.dims)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for st in range(3):
        src, dst = bufs[st % 2], bufs[(st + 1) % 2]
        _lib.check(_lib.lib().flue_flow_stage(dims, ctypes.c_void_p(src.data_ptr()), _lib.FLUE_F64,
                                              ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(acc.data_ptr()),
                                              0.03, st, stream), "flue_flow_stage")
    torch.cuda.synchronize()
    out = FL.GaugeField(g.dims, "fp64", _data=bufs[1])
    assert np.abs(out.to_complex().cpu().numpy() - W).max() < 1e-12
    # Win == Wout is refused
    assert _lib.lib().flue_flow_stage(dims, ctypes.c_void_p(bufs[0].data_ptr()), _lib.FLUE_F64,
                                      ctypes.c_void_p(bufs[0].data_ptr()), ctypes.c_void_p(acc.data_ptr()),
                                      0.03, 0, stream) == 1
