// k_grad.hip -- gradient of the Wilson plaquette action with respect to every link and
// the Wilson-flow (RK3) stage built on it.
//
// Convention (no reference counterpart): for a link U = U_mu(x) the Lie derivative is
// D^a f = d/ds f(exp(s T^a) U)|_0 with T^a = -i lambda_a / 2, and the gradient stored
// per link is G = sum_a T^a D^a f, traceless anti-Hermitian; D^a f = -2 Tr(T^a G).
//   Wilson action S = beta sum_p (1 - ReTr P/3):   G = (beta/6) P_TA(U Sigma)
// with Sigma the staple sum (U Sigma = the six plaquettes through U).
#include "flue_flow.hpp"
#include "flue_launch.hpp"

namespace flue {

namespace {

constexpr int kBS = 256;

// One thread per link.  The direction is the fastest block index: the four blocks of the
// same 256 sites run together and read the same neighbour links, so they share them in L2
// (a thread walking the four directions of its site read them again long after).
template <typename FV>
__global__ void __launch_bounds__(kBS) wilson_grad_kernel(Lat L, FV f, double beta,
                                                          typename FV::real* __restrict__ grad) {
    using R = typename FV::real;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(blk & 3u);
    const uint32_t s = (blk >> 2) * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
    {
        const Mat<R> M = mul(load_link(f, mu, s), staple_sum(L, f, s, c, mu));
        Mat<R> G = proj_ta(M);
        const R sc = (R)(beta / 6.0);
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            G.r[e] *= sc;
            G.i[e] *= sc;
        }
        store_link(grad, L.V, mu, s, G);
    }
}

template <typename FV>
__global__ void __launch_bounds__(kBS) flow_stage_kernel(Lat L, FV f, typename FV::real* __restrict__ wout,
                                                         double* __restrict__ acc, double eps, double ca, double cb,
                                                         int reunit) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(blk & 3u);
    const uint32_t s = (blk >> 2) * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
    flow_link(L, f, wout, acc, s, c, mu, eps, ca, cb, reunit);
}

}  // namespace

// four blocks (one per direction) per 256 sites
uint32_t grad_blocks(const Lat& L) { return 4u * ((L.V + kBS - 1) / kBS); }

hipError_t launch_wilson_grad(const Lat& L, const void* U, int is_f32, double beta, void* grad, hipStream_t st) {
    const uint32_t g = grad_blocks(L);
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        wilson_grad_kernel<F><<<g, kBS, 0, st>>>(L, f, beta, (typename F::real*)grad);
    });
    return hipGetLastError();
}

hipError_t launch_flow_stage(const Lat& L, const void* Win, int is_f32, void* Wout, double* acc, double eps,
                             double ca, double cb, int reunit, hipStream_t st) {
    const uint32_t g = grad_blocks(L);
    with_view(L.V, Win, is_f32, [&](auto f) {
        using F = decltype(f);
        flow_stage_kernel<F><<<g, kBS, 0, st>>>(L, f, (typename F::real*)Wout, acc, eps, ca, cb, reunit);
    });
    return hipGetLastError();
}

}  // namespace flue
