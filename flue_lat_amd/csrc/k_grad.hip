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
        store_link(grad, L.V, mu, s, G);
    }
}

template <typename FV>
__global__ void __launch_bounds__(kBS) flow_stage_kernel(Lat L, FV f, BlockOrder bo, typename FV::real* __restrict__ wout,
                                                         double* __restrict__ acc, double eps, double ca, double cb,
                                                         int reunit) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int mu = (int)(blk & 3u);
    const uint32_t s = site_block(L, bo, blk >> 2) * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
    flow_link(L, f, wout, acc, s, c, mu, eps, ca, cb, reunit);
}

}  // namespace

// brick extents (z-planes, t-slices) of the stencil variants; variant 1 is the storage order
constexpr uint32_t kBrickZ[kFlowVariants] = {0, 1, 2, 1, 2, 4};
constexpr uint32_t kBrickT[kFlowVariants] = {0, 2, 2, 4, 4, 4};

// the order of a variant, or the storage order where the bricks do not tile the lattice
// (planes of a whole number of blocks, extents divisible by the brick)
static BlockOrder block_order(const Lat& L, int variant) {
    BlockOrder bo{0, 0, 1, 1};
    if (variant < 1 || variant > kFlowVariants) variant = kFlowDefault;
    const uint32_t bz = kBrickZ[variant - 1], bt = kBrickT[variant - 1];
    const uint32_t plane = (uint32_t)L.n[1] * (uint32_t)L.n[2];
    if (bz == 0 || plane % kBS != 0 || L.n[3] % bz != 0 || L.n[0] % bt != 0) return bo;
    return BlockOrder{bz, bt, plane / kBS, (uint32_t)L.n[0] / bt};
}

// four blocks (one per direction) per 256 sites
uint32_t grad_blocks(const Lat& L) { return 4u * ((L.V + kBS - 1) / kBS); }

hipError_t launch_wilson_grad(const Lat& L, const void* U, int is_f32, double beta, void* grad, hipStream_t st,
                              int variant) {
    const uint32_t g = grad_blocks(L);
    const BlockOrder bo = block_order(L, variant);
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        wilson_grad_kernel<F><<<g, kBS, 0, st>>>(L, f, bo, beta, (typename F::real*)grad);
    });
    return hipGetLastError();
}

hipError_t launch_flow_stage(const Lat& L, const void* Win, int is_f32, void* Wout, double* acc, double eps,
                             double ca, double cb, int reunit, hipStream_t st, int variant) {
    const uint32_t g = grad_blocks(L);
    const BlockOrder bo = block_order(L, variant);
    with_view(L.V, Win, is_f32, [&](auto f) {
        using F = decltype(f);
        flow_stage_kernel<F><<<g, kBS, 0, st>>>(L, f, bo, (typename F::real*)Wout, acc, eps, ca, cb, reunit);
    });
    return hipGetLastError();
}

}  // namespace flue
