// k_clover.hip -- 5-loop improved field strength (hep-lat/0203008) and the magnetic B^2.
//
// Reference: Loop5MuNuCoord (src/wloops.F90:201-432) sums 28 clover leaves per site
// and plane through genericPath (208 matrix products).  Here each quadrant's seven
// rectangles a x b (a along the first direction d1 of the counter-clockwise loop, b
// along d2) are written W(a,b) = CA(a,b) CB(b,a)^+ with the corner paths
//     CA(a,b) = Line(x, d1, a) Line(x + a d1, d2, b)
//     CB(b,a) = Line(x, d2, b) Line(x + b d2, d1, a)
// which are extended one link at a time, so a quadrant costs 27 products and the
// corners stay in registers.  Only the anti-Hermitian part of the weighted sum is
// accumulated, since F = [ (Q - Q^+)/2 ]^T / 4 (src/wloops.F90:411-423).
#include "flue_launch.hpp"

namespace flue {

namespace {

constexpr double kK5 = 1.0 / 180.0;  // src/wloops.F90:219-224
constexpr double kK1 = (19.0 / 9.0) - 55.0 * kK5;
constexpr double kK2 = (1.0 / 36.0) - 16.0 * kK5;
constexpr double kK3 = 64.0 * kK5 - (32.0 / 45.0);
constexpr double kK4 = (1.0 / 15.0) - 6.0 * kK5;

// anti-Hermitian part of the weighted leaf sum: d_a = sum k Im W_aa,
// o_ab = sum k (W_ab - conj(W_ba)) for (a,b) = (0,1), (0,2), (1,2)
template <typename R>
struct AH {
    R d[3];
    R or_[3], oi[3];
};

template <typename R>
__device__ __forceinline__ void ah_zero(AH<R>& h) {
#pragma unroll
    for (int k = 0; k < 3; ++k) h.d[k] = h.or_[k] = h.oi[k] = R(0);
}

// h += w * AH-part of (CA CB^+), with W = CA CB^+, W_ab = sum_k CA_ak conj(CB_bk):
//   d_a += w Im W_aa,  o_ab += w (W_ab - conj(W_ba))
template <typename R>
__device__ __forceinline__ void acc_ah(AH<R>& h, R w, const Mat<R>& CA, const Mat<R>& CB) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        R im = CA.i[a * 3] * CB.r[a * 3];
        im = fmad(-CA.r[a * 3], CB.i[a * 3], im);
#pragma unroll
        for (int k = 1; k < 3; ++k) {
            im = fmad(CA.i[a * 3 + k], CB.r[a * 3 + k], im);
            im = fmad(-CA.r[a * 3 + k], CB.i[a * 3 + k], im);
        }
        h.d[a] = fmad(w, im, h.d[a]);
    }
    constexpr int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int a = pa[q], b = pb[q];
        // re(W_ab) - re(W_ba) and im(W_ab) + im(W_ba) accumulated in one chain each
        R xr = R(0), xi = R(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            xr = fmad(CA.r[a * 3 + k], CB.r[b * 3 + k], xr);
            xr = fmad(CA.i[a * 3 + k], CB.i[b * 3 + k], xr);
            xr = fmad(-CA.r[b * 3 + k], CB.r[a * 3 + k], xr);
            xr = fmad(-CA.i[b * 3 + k], CB.i[a * 3 + k], xr);
            xi = fmad(CA.i[a * 3 + k], CB.r[b * 3 + k], xi);
            xi = fmad(-CA.r[a * 3 + k], CB.i[b * 3 + k], xi);
            xi = fmad(CA.i[b * 3 + k], CB.r[a * 3 + k], xi);
            xi = fmad(-CA.r[b * 3 + k], CB.i[a * 3 + k], xi);
        }
        h.or_[q] = fmad(w, xr, h.or_[q]);
        h.oi[q] = fmad(w, xi, h.oi[q]);
    }
}

// One quadrant of the clover: the counter-clockwise loops leave x along d1 = S1 * e(ax1)
// and come back along -d2 = -S2 * e(ax2).  P(a1, a2) = x + a1 d1 + a2 d2 is a lattice
// point with site index sp + T1[a1] + T2[a2].  A link along d1 that leaves P is
// U_ax1(P) for S1 > 0 and U_ax1(P + d1)^+ for S1 < 0; the dagger is folded into the
// product (A U^+) instead of being materialised.
template <typename FV>
struct Quad {
    FV f;
    int ax1, ax2;
    uint32_t sp;
    uint32_t T1[4], T2[4];
};

// per-plane offsets along mu (A) and nu (B) for displacements -3..3, index k + 3
template <typename FV>
struct PlaneGeom {
    FV f;
    int mu, nu;
    uint32_t sp;
    uint32_t A[7], B[7];
};

template <typename FV>
__device__ __forceinline__ PlaneGeom<FV> plane_geom(const Lat& L, const FV& f, uint32_t s, int cmu, int cnu, int mu,
                                                    int nu) {
    PlaneGeom<FV> g;
    g.f = f;
    g.mu = mu;
    g.nu = nu;
    const uint32_t smu = L.st[mu], snu = L.st[nu];
    const int nmu = L.n[mu], nnu = L.n[nu];
    g.sp = s - (uint32_t)cmu * smu - (uint32_t)cnu * snu;
#pragma unroll
    for (int k = -3; k <= 3; ++k) {
        g.A[k + 3] = (uint32_t)wrapc(cmu + k, nmu) * smu;
        g.B[k + 3] = (uint32_t)wrapc(cnu + k, nnu) * snu;
    }
    return g;
}

// quadrant (S1 e(ax1), S2 e(ax2)) with ax1 = mu if MU_FIRST else nu
template <int S1, int S2, bool MU_FIRST, typename FV>
__device__ __forceinline__ Quad<FV> make_quad(const PlaneGeom<FV>& g) {
    Quad<FV> q;
    q.f = g.f;
    q.ax1 = MU_FIRST ? g.mu : g.nu;
    q.ax2 = MU_FIRST ? g.nu : g.mu;
    q.sp = g.sp;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        q.T1[k] = MU_FIRST ? g.A[3 + S1 * k] : g.B[3 + S1 * k];
        q.T2[k] = MU_FIRST ? g.B[3 + S2 * k] : g.A[3 + S2 * k];
    }
    return q;
}

// first link of a line along d1 (AX = 1) or d2 (AX = 2) leaving P(a1, a2), materialised
template <int S, int AX, int a1, int a2, typename FV>
__device__ __forceinline__ Mat<typename FV::real> first_link(const Quad<FV>& q) {
    const int ax = AX == 1 ? q.ax1 : q.ax2;
    if constexpr (S > 0) {
        return load_link(q.f, ax, q.sp + q.T1[a1] + q.T2[a2]);
    } else {
        constexpr int b1 = AX == 1 ? a1 + 1 : a1, b2 = AX == 2 ? a2 + 1 : a2;
        return load_link_dag(q.f, ax, q.sp + q.T1[b1] + q.T2[b2]);
    }
}

// M * (link along d1 or d2 leaving P(a1, a2))
template <int S, int AX, int a1, int a2, typename FV>
__device__ __forceinline__ Mat<typename FV::real> ext(const Quad<FV>& q, const Mat<typename FV::real>& M) {
    using R = typename FV::real;
    const int ax = AX == 1 ? q.ax1 : q.ax2;
    constexpr int b1 = (AX == 1 && S < 0) ? a1 + 1 : a1;
    constexpr int b2 = (AX == 2 && S < 0) ? a2 + 1 : a2;
    const Mat<R> U = load_link(q.f, ax, q.sp + q.T1[b1] + q.T2[b2]);
    if constexpr (S > 0)
        return mul(M, U);
    else
        return mul_nd(M, U);
}

// Scheduling fences keep each group's link loads next to the products that use them;
// without them the scheduler front-loads a quadrant's loads and runs out of registers.
#ifndef FLUE_FENCE_MODE
#define FLUE_FENCE_MODE 1
#endif
#if FLUE_FENCE_MODE == 1
#define FLUE_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define FLUE_SCHED_FENCE() ((void)0)
#endif

// the seven rectangles (1,1),(1,2),(1,3),(2,1),(2,2),(3,1),(3,3), a along d1 and b along
// d2, or only the 1x1 leaf when CLOVER is set (cloverLoopMuNuCoord, src/wloops.F90:129-199)
template <int S1, int S2, bool CLOVER, typename FV>
__device__ __forceinline__ void quadrant(const Quad<FV>& q, AH<typename FV::real>& h) {
    using R = typename FV::real;
    const R w11 = CLOVER ? R(1) : R(kK1);
    const R w22 = R(kK2), w33 = R(kK5), w12 = R(0.5 * kK3), w13 = R(0.5 * kK4);
    // row a = 1
    const Mat<R> L1 = first_link<S1, 1, 0, 0>(q);   // Line(x,d1,1)
    Mat<R> CA = ext<S2, 2, 1, 0>(q, L1);            // CA(1,1)
    Mat<R> Ld = first_link<S2, 2, 0, 0>(q);         // Line(x,d2,1)
    Mat<R> CB1 = ext<S1, 1, 0, 1>(q, Ld);           // CB(1,1)
    acc_ah(h, w11, CA, CB1);                        // W(1,1)
    if constexpr (!CLOVER) {
        FLUE_SCHED_FENCE();
        CA = ext<S2, 2, 1, 1>(q, CA);               // CA(1,2)
        Ld = ext<S2, 2, 0, 1>(q, Ld);               // Line(x,d2,2)
        Mat<R> CB2 = ext<S1, 1, 0, 2>(q, Ld);       // CB(2,1)
        acc_ah(h, w12, CA, CB2);                    // W(1,2)
        FLUE_SCHED_FENCE();
        CA = ext<S2, 2, 1, 2>(q, CA);               // CA(1,3)
        Ld = ext<S2, 2, 0, 2>(q, Ld);               // Line(x,d2,3)
        Mat<R> CB3 = ext<S1, 1, 0, 3>(q, Ld);       // CB(3,1)
        acc_ah(h, w13, CA, CB3);                    // W(1,3)
        FLUE_SCHED_FENCE();
        // row a = 2
        Mat<R> L = ext<S1, 1, 1, 0>(q, L1);         // Line(x,d1,2)
        CA = ext<S2, 2, 2, 0>(q, L);                // CA(2,1)
        CB1 = ext<S1, 1, 1, 1>(q, CB1);             // CB(1,2)
        acc_ah(h, w12, CA, CB1);                    // W(2,1)
        FLUE_SCHED_FENCE();
        CA = ext<S2, 2, 2, 1>(q, CA);               // CA(2,2)
        CB2 = ext<S1, 1, 1, 2>(q, CB2);             // CB(2,2)
        acc_ah(h, w22, CA, CB2);                    // W(2,2)
        FLUE_SCHED_FENCE();
        // row a = 3
        L = ext<S1, 1, 2, 0>(q, L);                 // Line(x,d1,3)
        CA = ext<S2, 2, 3, 0>(q, L);                // CA(3,1)
        CB1 = ext<S1, 1, 2, 1>(q, CB1);             // CB(1,3)
        acc_ah(h, w13, CA, CB1);                    // W(3,1)
        FLUE_SCHED_FENCE();
        CA = ext<S2, 2, 3, 1>(q, CA);               // CA(3,2)
        CA = ext<S2, 2, 3, 2>(q, CA);               // CA(3,3)
        CB3 = ext<S1, 1, 1, 3>(q, CB3);             // CB(3,2)
        CB3 = ext<S1, 1, 2, 3>(q, CB3);             // CB(3,3)
        acc_ah(h, w33, CA, CB3);                    // W(3,3)
    }
    FLUE_SCHED_FENCE();
}

// the four quadrants in counter-clockwise order: (+mu,+nu), (+nu,-mu), (-mu,-nu), (-nu,+mu)
template <bool CLOVER, typename FV>
__device__ __forceinline__ AH<typename FV::real> plane_ah(const PlaneGeom<FV>& g) {
    AH<typename FV::real> h;
    ah_zero(h);
    quadrant<1, 1, CLOVER>(make_quad<1, 1, true>(g), h);
    quadrant<1, -1, CLOVER>(make_quad<1, -1, false>(g), h);
    quadrant<-1, -1, CLOVER>(make_quad<-1, -1, true>(g), h);
    quadrant<-1, 1, CLOVER>(make_quad<-1, 1, false>(g), h);
    return h;
}

// sum_ab |F_ab|^2 = -ReTr(F F) for F = A^T/4 with A = AH part (all in FP64)
template <typename R>
__device__ __forceinline__ double f_norm2(const AH<R>& h) {
    double dsum = 0.0, osum = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double d = (double)h.d[k];
        const double orr = (double)h.or_[k], oi = (double)h.oi[k];
        dsum += d * d;
        osum += orr * orr + oi * oi;
    }
    return (dsum + 0.5 * osum) * (1.0 / 16.0);
}

constexpr int kBS = 256;

// B = -(1/V) sum_x sum_i ReTr(F_i F_i) over the planes (y,x), (x,z), (z,y)
// (src/wloops.F90:502-506, 530-543); this kernel writes one partial sum per block.
template <typename FV, int WPE>
__global__ void __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
magnetic_kernel(Lat L, FV f, double* __restrict__ partials) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * kBS + threadIdx.x;
    double acc = 0.0;
    if (s < L.V) {
        int c[4];
        site_coords(L, s, c);
#pragma unroll 1
        for (int pl = 0; pl < 3; ++pl) {
            // (mu, nu) = (y, x), (x, z), (z, y) in 0-based directions
            const int mu = pl == 0 ? 2 : (pl == 1 ? 1 : 3);
            const int nu = pl == 0 ? 1 : (pl == 1 ? 3 : 2);
            const int cmu = pl == 0 ? c[2] : (pl == 1 ? c[1] : c[3]);
            const int cnu = pl == 0 ? c[1] : (pl == 1 ? c[3] : c[2]);
            acc += f_norm2(plane_ah<false>(plane_geom(L, f, s, cmu, cnu, mu, nu)));
        }
    }
    DD r = block_sum<kBS>(DD{acc, 0.0});
    if (threadIdx.x == 0) {
        partials[2 * blk] = r.hi;
        partials[2 * blk + 1] = r.lo;
    }
}

struct OutStrides {
    int64_t t, x, y, z, a, b;
};

// per-site F for one plane.  layout 0: complex128 at the given strides; layout 1: SoA
// planes (element e = a*3+b, re/im) in the field precision.
template <typename FV, bool CLOVER>
__global__ void __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(sizeof(typename FV::real) == 4 ? 2 : 1)))
fmunu_kernel(Lat L, FV f, int mu, int nu, void* __restrict__ out, int layout, OutStrides os) {
    using R = typename FV::real;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
    const AH<R> h = plane_ah<CLOVER>(plane_geom(L, f, s, c[mu], c[nu], mu, nu));
    // F = A^T / 4: F_aa = i d_a / 4, F_ab = -conj(o_ab)/8, F_ba = o_ab/8 (a < b)
    R fr[9], fi[9];
    constexpr int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        fr[k * 4] = R(0);
        fi[k * 4] = h.d[k] * R(0.25);
        const int a = pa[k], b = pb[k];
        fr[a * 3 + b] = -h.or_[k] * R(0.125);
        fi[a * 3 + b] = h.oi[k] * R(0.125);
        fr[b * 3 + a] = h.or_[k] * R(0.125);
        fi[b * 3 + a] = h.oi[k] * R(0.125);
    }
    if (layout == 0) {
        double* o = (double*)out;
        const int64_t base = c[0] * os.t + c[1] * os.x + c[2] * os.y + c[3] * os.z;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int64_t off = 2 * (base + a * os.a + b * os.b);
                o[off] = (double)fr[a * 3 + b];
                o[off + 1] = (double)fi[a * 3 + b];
            }
    } else {
        R* o = (R*)out;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            o[(size_t)(2 * e) * L.V + s] = fr[e];
            o[(size_t)(2 * e + 1) * L.V + s] = fi[e];
        }
    }
}

}  // namespace

uint32_t magnetic_blocks(const Lat& L, int) { return (L.V + kBS - 1) / kBS; }

int magnetic_default_variant(int is_f32) { return is_f32 ? 2 : 1; }

// variant = occupancy target in waves per SIMD (1, 2 or 3); 0 picks the default
hipError_t launch_magnetic(const Lat& L, const void* U, int is_f32, double* partials, hipStream_t st, int variant) {
    const uint32_t g = magnetic_blocks(L, is_f32);
    if (variant == 0) variant = magnetic_default_variant(is_f32);
    if (variant < 1 || variant > 3) return hipErrorInvalidValue;
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        using R = typename F::real;
        if constexpr (!F::buffered) {
            // fields beyond the buffer-resource range: one variant only
            magnetic_kernel<F, sizeof(R) == 4 ? 2 : 1><<<g, kBS, 0, st>>>(L, f, partials);
        } else {
            if (variant == 1)
                magnetic_kernel<F, 1><<<g, kBS, 0, st>>>(L, f, partials);
            else if (variant == 2)
                magnetic_kernel<F, 2><<<g, kBS, 0, st>>>(L, f, partials);
            else
                magnetic_kernel<F, 3><<<g, kBS, 0, st>>>(L, f, partials);
        }
    });
    return hipGetLastError();
}

hipError_t launch_fmunu(const Lat& L, const void* U, int is_f32, int mu, int nu, int clover_only, void* out,
                        int out_layout, const int64_t s6[6], hipStream_t st) {
    const uint32_t g = (L.V + kBS - 1) / kBS;
    OutStrides os{0, 0, 0, 0, 0, 0};
    if (s6) os = OutStrides{s6[0], s6[1], s6[2], s6[3], s6[4], s6[5]};
    with_view(L.V, U, is_f32, [&](auto f) {
        if (clover_only)
            fmunu_kernel<decltype(f), true><<<g, kBS, 0, st>>>(L, f, mu, nu, out, out_layout, os);
        else
            fmunu_kernel<decltype(f), false><<<g, kBS, 0, st>>>(L, f, mu, nu, out, out_layout, os);
    });
    return hipGetLastError();
}

}  // namespace flue
