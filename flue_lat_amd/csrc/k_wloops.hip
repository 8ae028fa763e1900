// k_wloops.hip -- all rectangular Wilson loops R x T (1 <= R <= Rmax, 1 <= T <= Tmax) of
// one plane in a single pass over the lattice.
//
// The loop at x is ReTr[CA(R,T) CB(T,R)^+] (src/wloops.F90:679-791, :567-677), with
//   CA(R,T) = Lmu(x; R) Lnu(x + R mu; T)   (R steps along mu, then T along nu),
//   CB(T,R) = Lnu(x; T) Lmu(x + T nu; R)   (T steps along nu, then R along mu),
// where L(y; n) is the product of n successive links from y.  One thread per site keeps
// the Tmax chains CB(T, .) and extends them by one link per R, and runs one CA chain per
// R: about Rmax*Tmax + 2*(Rmax + Tmax) products per site for the whole table, against
// 2(R+T)-1 products for each loop walked on its own.
//
// A loop reads links up to 6 sites away along mu and nu.  In a plane containing time those
// are up to 6 time slices apart, so there the blocks are ordered with the time slice
// fastest: a block covers kBS sites of one slice, consecutive blocks take the same sites
// of consecutive slices, and the blocks that run together on an XCD share their links in L2
// (in storage order the L2 of an XCD held no more than a slice, and 9 in 10 reads missed).
// In the (z, t) and (x, z) planes a slice is also taken with z second-fastest after x, so
// the loops' z neighbours lie in the same block or the next rather than 9 blocks on.
#include "flue_common.hpp"
#include "flue_launch.hpp"

namespace flue {

namespace {

constexpr int kBS = 256;
constexpr int kWaves = kBS / 64;

// sum over the wave of a per-lane double (plain adds; the per-block partials and the
// final pass are compensated)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Grid: groups x nb logical blocks, one group per time slice for kEverySlice, else one.
// With time in the plane the slice is the fastest block index (for the whole lattice, nb
// counts the blocks of every slice, slice fastest); otherwise the block within the group is.
template <typename FV>
__global__ void __launch_bounds__(kBS) wloop_table_kernel(Lat L, FV f, int mu, int nu, int Rmax, int Tmax,
                                                          int tslice, uint32_t nb, uint32_t groups,
                                                          double* __restrict__ partials) {
    using R_ = typename FV::real;
    __shared__ double sh[kMaxLoopRT * kMaxLoopRT][kWaves];
    const bool tplane = mu == 0 || nu == 0;
    const uint32_t g = xcd_block(blockIdx.x, gridDim.x);
    uint32_t grp, blk;
    if (tslice == kEverySlice && tplane) {
        grp = g % groups;
        blk = g / groups;
    } else {
        grp = g / nb;
        blk = g - grp * nb;
    }
    // tslice >= 0: that time slice; -1: every site; kEverySlice: slice grp
    const int ts = tslice == kEverySlice ? (int)grp : tslice;
    // in a time plane the slice is taken along the plane's spatial direction; in the (x, z)
    // plane along z
    const int d = tplane ? mu + nu : (mu + nu == 4 ? 3 : 1);
    bool live;
    uint32_t s;
    if (ts >= 0) {
        const uint32_t idx = blk * kBS + threadIdx.x;
        live = idx < L.Vs;
        s = (live ? slice_site_along(L, idx, d) : 0u) + L.Vs * (uint32_t)ts;
    } else if (tplane) {
        const uint32_t NT = (uint32_t)L.n[0];
        const uint32_t chunk = blk / NT, t = blk - chunk * NT;
        const uint32_t idx = chunk * kBS + threadIdx.x;
        live = idx < L.Vs;
        s = (live ? slice_site_along(L, idx, d) : 0u) + L.Vs * t;
    } else {
        const uint32_t idx = blk * kBS + threadIdx.x;
        live = idx < L.V;
        const uint32_t t = idx / L.Vs;
        s = d == 1 ? idx : slice_site_along(L, idx - t * L.Vs, d) + L.Vs * t;
    }
    // lanes past the end work on site 0 (in bounds) and contribute nothing
    if (!live) s = 0u;
    int c[4];
    site_coords(L, s, c);
    const int cmu = c[mu], cnu = c[nu];
    const int nmu = L.ext(mu), nnu = L.ext(nu);
    const uint32_t smu = L.stride(mu), snu = L.stride(nu);
    const uint32_t base = s - (uint32_t)cmu * smu - (uint32_t)cnu * snu;
    // site of x + a mu + b nu = base + A[a] + B[b]
    uint32_t A[kMaxLoopRT + 1], B[kMaxLoopRT + 1];
#pragma unroll
    for (int k = 0; k <= kMaxLoopRT; ++k) {
        A[k] = (uint32_t)((cmu + k) % nmu) * smu;
        B[k] = (uint32_t)((cnu + k) % nnu) * snu;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;

    // CB(T, 1) for T = 1..Tmax from the nu line at x
    Mat<R_> CB[kMaxLoopRT];
    Mat<R_> Ln = load_link(f, nu, base + A[0] + B[0]);
#pragma unroll
    for (int T = 1; T <= kMaxLoopRT; ++T) {
        if (T <= Tmax) {
            if (T > 1) Ln = mul(Ln, load_link(f, nu, base + A[0] + B[T - 1]));
            CB[T - 1] = mul(Ln, load_link(f, mu, base + A[0] + B[T]));
        }
    }
    Mat<R_> Lm = load_link(f, mu, base + A[0] + B[0]);
#pragma unroll
    for (int R = 1; R <= kMaxLoopRT; ++R) {
        if (R <= Rmax) {
            if (R > 1) {
                Lm = mul(Lm, load_link(f, mu, base + A[R - 1] + B[0]));
#pragma unroll
                for (int T = 1; T <= kMaxLoopRT; ++T)
                    if (T <= Tmax) CB[T - 1] = mul(CB[T - 1], load_link(f, mu, base + A[R - 1] + B[T]));
            }
            Mat<R_> CA = mul(Lm, load_link(f, nu, base + A[R] + B[0]));
#pragma unroll
            for (int T = 1; T <= kMaxLoopRT; ++T) {
                if (T <= Tmax) {
                    if (T > 1) CA = mul(CA, load_link(f, nu, base + A[R] + B[T - 1]));
                    const double v = live ? (double)retr_nd(CA, CB[T - 1]) : 0.0;
                    const double w = wave_sum(v);
                    if (lane == 0) sh[(R - 1) * kMaxLoopRT + (T - 1)][wave] = w;
                }
            }
        }
    }
    __syncthreads();
    // thread k writes the block partial of table entry k (fixed order over the waves)
    const int k = threadIdx.x;
    if (k < kMaxLoopRT * kMaxLoopRT && (k / kMaxLoopRT) < Rmax && (k % kMaxLoopRT) < Tmax) {
        DD r = {0.0, 0.0};
#pragma unroll
        for (int w = 0; w < kWaves; ++w) r = dd_add(r, DD{sh[k][w], 0.0});
        double* p = partials + 2 * (((size_t)k * groups + grp) * nb + blk);
        p[0] = r.hi;
        p[1] = r.lo;
    }
}

// block (k, j) of the grid sums table entry (R, T) = (k / Tmax + 1, k % Tmax + 1) of slice
// group j (of gridDim.y) over its nb block partials in a fixed order and writes
// out[k * gridDim.y + j] = scale * sum
__global__ void __launch_bounds__(kBS) wloop_finalize_kernel(const double* __restrict__ partials, uint32_t nb,
                                                             int Tmax, double* __restrict__ out, double scale) {
    const int k = blockIdx.x;
    const int e = (k / Tmax) * kMaxLoopRT + (k % Tmax);
    const double* p = partials + 2 * (((size_t)e * gridDim.y + blockIdx.y) * nb);
    DD acc = {0.0, 0.0};
    for (uint32_t i = threadIdx.x; i < nb; i += kBS) acc = dd_add(acc, DD{p[2 * i], p[2 * i + 1]});
    const DD r = block_sum<kBS>(acc);
    if (threadIdx.x == 0) out[(size_t)k * gridDim.y + blockIdx.y] = (r.hi + r.lo) * scale;
}

}  // namespace

// blocks per slice group: a single slice (tslice >= 0 or kEverySlice), or the whole
// lattice (slice by slice when time is in the plane)
uint32_t wloop_blocks(const Lat& L, int tslice, int mu, int nu) {
    const uint32_t per_slice = (L.Vs + kBS - 1) / kBS;
    if (tslice != -1) return per_slice;
    return (mu == 0 || nu == 0) ? (uint32_t)L.n[0] * per_slice : (L.V + kBS - 1) / kBS;
}

// enough for every plane
size_t wloop_partial_bytes(const Lat& L, int tslice) {
    const size_t groups = tslice == kEverySlice ? (size_t)L.n[0] : 1;
    const size_t nb = std::max(wloop_blocks(L, tslice, 0, 1), wloop_blocks(L, tslice, 1, 2));
    return (size_t)kMaxLoopRT * kMaxLoopRT * groups * nb * 2 * sizeof(double);
}

hipError_t launch_wloop_table(const Lat& L, const void* U, int is_f32, int mu, int nu, int Rmax, int Tmax,
                              int tslice, double* partials, double* out, double scale, hipStream_t st) {
    const uint32_t nb = wloop_blocks(L, tslice, mu, nu);
    const uint32_t groups = tslice == kEverySlice ? (uint32_t)L.n[0] : 1u;
    with_view(L.V, U, is_f32, [&](auto f) {
        wloop_table_kernel<decltype(f)>
            <<<nb * groups, kBS, 0, st>>>(L, f, mu, nu, Rmax, Tmax, tslice, nb, groups, partials);
    });
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    wloop_finalize_kernel<<<dim3(Rmax * Tmax, groups), kBS, 0, st>>>(partials, nb, Tmax, out, scale);
    return hipGetLastError();
}

}  // namespace flue
