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

// Visit orders of the 256-site blocks (the autotuner's variants of the flow and gradient
// stencils).  Every order runs the same per-link code, so results are bit-identical.
//   bz = 0:  storage order (x, y, z, t).
//   bz > 0:  bricks of bz z-planes x bt t-slices (whole (x, y) planes of 256k sites):
//            within a brick t runs fastest, then z, then the blocks of a plane, so the
//            blocks in flight on an XCD hold the z and t neighbours of what they read.
struct BlockOrder {
    uint32_t bz, bt;    // brick extents (0 = storage order)
    uint32_t nbp;       // 256-site blocks per (z, t) plane
    uint32_t ntb;       // bricks along t
};

__device__ __forceinline__ uint32_t site_block(const Lat& L, const BlockOrder& bo, uint32_t lb) {
    if (bo.bz == 0) return lb;
    const uint32_t k = lb % bo.nbp, q = lb / bo.nbp;
    const uint32_t per = bo.bz * bo.bt, pib = q % per, b = q / per;
    const uint32_t t = (b % bo.ntb) * bo.bt + pib % bo.bt;
    const uint32_t z = (b / bo.ntb) * bo.bz + pib / bo.bt;
    return (t * (uint32_t)L.n[3] + z) * bo.nbp + k;
}

// One thread per link.  The direction is the fastest block index: the four blocks of the
// same 256 sites run together and read the same neighbour links, so they share them in L2
// (a thread walking the four directions of its site read them again long after).
template <typename FV>
__global__ void __launch_bounds__(kBS) wilson_grad_kernel(Lat L, FV f, BlockOrder bo, double beta,
                                                          typename FV::real* __restrict__ grad) {
    using R = typename FV::real;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(blk & 3u);
    const uint32_t s = site_block(L, bo, blk >> 2) * kBS + threadIdx.x;
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
        store_link_stream(grad, L.V, mu, s, G);
    }
}

template <typename FV>
__global__ void __launch_bounds__(kBS) flow_stage_kernel(Lat L, FV f, BlockOrder bo, typename FV::real* __restrict__ wout,
                                                         typename FV::real* __restrict__ acc, double eps, double ca,
                                                         double cb, int reunit) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(blk & 3u);
    const uint32_t s = site_block(L, bo, blk >> 2) * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
    flow_link(L, f, wout, acc, s, c, mu, eps, ca, cb, reunit);
}

// plaquette-first stencil (kFlowPF): the plaquette parts of every site, then per link
template <typename FV>
__global__ void __launch_bounds__(kBS) plaq_ta_kernel(Lat L, FV f, BlockOrder bo, typename FV::real* __restrict__ T) {
    const uint32_t s = site_block(L, bo, xcd_block(blockIdx.x, gridDim.x)) * kBS + threadIdx.x;
    if (s >= L.V) return;
    plaquette_ta_site(L, f, s, T);
}

template <typename FV>
__global__ void __launch_bounds__(kBS) wilson_grad_pf_kernel(Lat L, FV f, BlockOrder bo,
                                                             const typename FV::real* __restrict__ T, double beta,
                                                             typename FV::real* __restrict__ grad) {
    using R = typename FV::real;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(blk & 3u);
    const uint32_t s = site_block(L, bo, blk >> 2) * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
    Mat<R> G = ta_force(L, f, T, s, c, mu);
    const R sc = (R)(beta / 6.0);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        G.r[e] *= sc;
        G.i[e] *= sc;
    }
    store_link_stream(grad, L.V, mu, s, G);
}

template <typename FV>
__global__ void __launch_bounds__(kBS) flow_stage_pf_kernel(Lat L, FV f, BlockOrder bo,
                                                            const typename FV::real* __restrict__ T,
                                                            typename FV::real* __restrict__ wout,
                                                            typename FV::real* __restrict__ acc, double eps, double ca,
                                                            double cb, int reunit) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(blk & 3u);
    const uint32_t s = site_block(L, bo, blk >> 2) * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
    flow_link_pf(L, f, T, wout, acc, s, c, mu, eps, ca, cb, reunit);
}

// Plaquette-first with walks along t (kFlowPFWalk).  Pass 1: one thread per spatial site walks
// kPFWalk slices, keeping the spatial links at x + t in registers as the next slice's own
// links (13 link loads per site and slice, none from another slice).  Pass 2: one thread per
// (spatial site, direction) walks the same slices, carrying the previous slice's T_{0 mu} and
// U_0 for the t term of a spatial link, so no thread reads another slice after its first.
constexpr int kPFWalk = 6;

template <typename FV>
__global__ void __launch_bounds__(kBS) plaq_ta_walk_kernel(Lat L, FV f, uint32_t nbs, typename FV::real* __restrict__ T) {
    using R = typename FV::real;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t tc = blk / nbs, pb = blk - tc * nbs;
    const uint32_t p = pb * kBS + threadIdx.x;
    if (p >= L.Vs) return;
    const int NT = L.n[0];
    const int t0 = (int)tc * kPFWalk, t1 = min(t0 + kPFWalk, NT);
    int c[4];
    site_coords(L, p, c);
    uint32_t sf[4];  // forward spatial neighbours within a slice
#pragma unroll
    for (int d = 1; d < 4; ++d) sf[d] = c[d] == L.n[d] - 1 ? p - (uint32_t)(L.n[d] - 1) * L.st[d] : p + L.st[d];
    Mat<R> Us[3];
#pragma unroll
    for (int b = 1; b < 4; ++b) Us[b - 1] = load_link(f, b, p + L.Vs * (uint32_t)t0);
    for (int t = t0; t < t1; ++t) {
        const uint32_t ts = L.Vs * (uint32_t)t, s = p + ts;
        const uint32_t sn = t == NT - 1 ? p : s + L.Vs;  // x + t
        const Mat<R> Ut = load_link(f, 0, s);
        Mat<R> N[3];
#pragma unroll
        for (int b = 1; b < 4; ++b) {
            N[b - 1] = load_link(f, b, sn);
            store_plaq_ta(T, L.V, plaq_index(0, b), s, mul(Ut, N[b - 1]), mul(Us[b - 1], load_link(f, 0, sf[b] + ts)));
        }
#pragma unroll
        for (int a = 1; a < 3; ++a)
#pragma unroll
            for (int b = a + 1; b < 4; ++b)
                store_plaq_ta(T, L.V, plaq_index(a, b), s, mul(Us[a - 1], load_link(f, b, sf[a] + ts)),
                              mul(Us[b - 1], load_link(f, a, sf[b] + ts)));
#pragma unroll
        for (int b = 0; b < 3; ++b) Us[b] = N[b];
    }
}

// pass 2 of kFlowPFWalk for the flow (GRAD = false) or the Wilson gradient (GRAD = true)
template <bool GRAD, typename FV>
__global__ void __launch_bounds__(kBS) pf_walk_kernel(Lat L, FV f, const typename FV::real* __restrict__ T,
                                                      uint32_t nbs, typename FV::real* __restrict__ out,
                                                      typename FV::real* __restrict__ acc, double eps, double ca,
                                                      double cb, int reunit, double beta) {
    using R = typename FV::real;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(blk & 3u);
    const uint32_t rest = blk >> 2, tc = rest / nbs, pb = rest - tc * nbs;
    const uint32_t p = pb * kBS + threadIdx.x;
    if (p >= L.Vs) return;
    const int NT = L.n[0];
    const int t0 = (int)tc * kPFWalk, t1 = min(t0 + kPFWalk, NT);
    int c[4];
    site_coords(L, p, c);
    Mat<R> Tprev, U0prev;
    if (mu != 0) {
        const uint32_t sp = p + L.Vs * (uint32_t)(t0 == 0 ? NT - 1 : t0 - 1);
        Tprev = load_ta(T, L.V, plaq_index(0, mu), sp);
        U0prev = load_link(f, 0, sp);
    }
    for (int t = t0; t < t1; ++t) {
        const uint32_t s = p + L.Vs * (uint32_t)t;
        c[0] = t;
        Mat<R> Z = ta_force_walk(L, f, T, s, c, mu, Tprev, U0prev);
        if constexpr (GRAD) {
            const R sc = (R)(beta / 6.0);
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                Z.r[e] *= sc;
                Z.i[e] *= sc;
            }
            store_link_stream(out, L.V, mu, s, Z);
        } else {
            flow_update(L, f, out, acc, s, mu, eps, ca, cb, reunit, Z);
        }
    }
}

}  // namespace

// brick extents (z-planes, t-slices) of the staple-form variants; variant 1 is the storage order
constexpr uint32_t kBrickZ[kFlowPF - 1] = {0, 1, 2, 1, 2, 4};  // variants 1..6
constexpr uint32_t kBrickT[kFlowPF - 1] = {0, 2, 2, 4, 4, 4};

// the order of a variant, or the storage order where the bricks do not tile the lattice
// (planes of a whole number of blocks, extents divisible by the brick).  Plaquette-first
// variant kFlowPF + k (k = 2..6) takes the order of staple variant k; kFlowPF and the walks
// (which order their own blocks) take the storage order.
static BlockOrder block_order(const Lat& L, int variant) {
    BlockOrder bo{0, 0, 1, 1};
    if (variant > kFlowPFWalk && variant <= kFlowVariants) variant -= kFlowPF;
    if (variant < 1 || variant >= kFlowPF) variant = kFlowDefault;
    const uint32_t bz = kBrickZ[variant - 1], bt = kBrickT[variant - 1];
    const uint32_t plane = (uint32_t)L.n[1] * (uint32_t)L.n[2];
    if (bz == 0 || plane % kBS != 0 || L.n[3] % bz != 0 || L.n[0] % bt != 0) return bo;
    return BlockOrder{bz, bt, plane / kBS, (uint32_t)L.n[0] / bt};
}

// four blocks (one per direction) per 256 sites
uint32_t grad_blocks(const Lat& L) { return 4u * ((L.V + kBS - 1) / kBS); }

size_t plaq_ta_bytes(const Lat& L, int is_f32) { return (size_t)48 * L.V * (is_f32 ? 4 : 8); }

hipError_t launch_wilson_grad(const Lat& L, const void* U, int is_f32, double beta, void* grad, hipStream_t st,
                              int variant, void* T) {
    const uint32_t g = grad_blocks(L);
    if (flow_is_pf(variant) && !T) return hipErrorInvalidValue;
    const BlockOrder bo = block_order(L, variant);
    const uint32_t nbs = (L.Vs + kBS - 1) / kBS, ntc = ((uint32_t)L.n[0] + kPFWalk - 1) / kPFWalk;
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        using R = typename F::real;
        if (variant == kFlowPFWalk) {
            plaq_ta_walk_kernel<F><<<nbs * ntc, kBS, 0, st>>>(L, f, nbs, (R*)T);
            pf_walk_kernel<true, F><<<4 * nbs * ntc, kBS, 0, st>>>(L, f, (const R*)T, nbs, (R*)grad, nullptr, 0.0, 0.0,
                                                                  0.0, 0, beta);
        } else if (flow_is_pf(variant)) {
            plaq_ta_kernel<F><<<(L.V + kBS - 1) / kBS, kBS, 0, st>>>(L, f, bo, (R*)T);
            wilson_grad_pf_kernel<F><<<g, kBS, 0, st>>>(L, f, bo, (const R*)T, beta, (R*)grad);
        } else {
            wilson_grad_kernel<F><<<g, kBS, 0, st>>>(L, f, bo, beta, (R*)grad);
        }
    });
    return hipGetLastError();
}

hipError_t launch_flow_stage(const Lat& L, const void* Win, int is_f32, void* Wout, void* acc_, double eps,
                             double ca, double cb, int reunit, hipStream_t st, int variant, void* T) {
    const uint32_t g = grad_blocks(L);
    if (flow_is_pf(variant) && !T) return hipErrorInvalidValue;
    const BlockOrder bo = block_order(L, variant);
    const uint32_t nbs = (L.Vs + kBS - 1) / kBS, ntc = ((uint32_t)L.n[0] + kPFWalk - 1) / kPFWalk;
    with_view(L.V, Win, is_f32, [&](auto f) {
        using F = decltype(f);
        using R = typename F::real;
        R* acc = (R*)acc_;
        if (variant == kFlowPFWalk) {
            plaq_ta_walk_kernel<F><<<nbs * ntc, kBS, 0, st>>>(L, f, nbs, (R*)T);
            pf_walk_kernel<false, F><<<4 * nbs * ntc, kBS, 0, st>>>(L, f, (const R*)T, nbs, (R*)Wout, acc, eps, ca, cb,
                                                                   reunit, 0.0);
        } else if (flow_is_pf(variant)) {
            plaq_ta_kernel<F><<<(L.V + kBS - 1) / kBS, kBS, 0, st>>>(L, f, bo, (R*)T);
            flow_stage_pf_kernel<F><<<g, kBS, 0, st>>>(L, f, bo, (const R*)T, (R*)Wout, acc, eps, ca, cb, reunit);
        } else {
            flow_stage_kernel<F><<<g, kBS, 0, st>>>(L, f, bo, (R*)Wout, acc, eps, ca, cb, reunit);
        }
    });
    return hipGetLastError();
}

}  // namespace flue
