// flue_clover.hpp -- the 5-loop clover leaf machinery shared by the field-strength,
// magnetic and gradient kernels.
//
// Reference: Loop5MuNuCoord (src/wloops.F90:201-432) sums 28 clover leaves per site
// and plane through genericPath (208 matrix products).  Here each quadrant's seven
// rectangles a x b (a along the first direction d1 of the counter-clockwise loop, b
// along d2) are written W(a,b) = CA(a,b) CB(b,a)^+ with the corner paths
//     CA(a,b) = Line(x, d1, a) Line(x + a d1, d2, b)
//     CB(b,a) = Line(x, d2, b) Line(x + b d2, d1, a)
// which are extended one link at a time, so a quadrant costs 20 products plus 7
// anti-Hermitian-part products and the corners stay in registers.  Only the
// anti-Hermitian part of the weighted sum is accumulated, since
// F = [ (Q - Q^+)/2 ]^T / 4 (src/wloops.F90:411-423).
#pragma once
#include "flue_common.hpp"

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

// all three rows of a chain that may carry two (C2: SU(3) links, row 2 = conj(row0 x row1))
template <bool C2, typename R>
__device__ __forceinline__ Mat<R> full_rows(const Mat<R>& M) {
    Mat<R> F = M;
    if constexpr (C2) fill_row2(F);
    return F;
}

// A V for the anti-Hermitian A with parts g: A_aa = i d_a, A_ab = o_ab / 2 and
// A_ba = -conj(o_ab) / 2 for a < b (the real diagonal of A, zero, takes no multiplies)
template <typename R>
__device__ __forceinline__ Mat<R> ah_mul(const AH<R>& g, const Mat<R>& V) {
    // off-diagonal entries of A, row-major, (re, im)
    R ar[9], ai[9];
    constexpr int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int a = pa[q], b = pb[q];
        ar[a * 3 + b] = R(0.5) * g.or_[q];
        ai[a * 3 + b] = R(0.5) * g.oi[q];
        ar[b * 3 + a] = R(-0.5) * g.or_[q];
        ai[b * 3 + a] = R(0.5) * g.oi[q];
    }
    // T = A V
    Mat<R> T;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            R tr = -g.d[a] * V.i[a * 3 + j];
            R ti = g.d[a] * V.r[a * 3 + j];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (k == a) continue;
                tr = fmad(ar[a * 3 + k], V.r[k * 3 + j], tr);
                tr = fmad(-ai[a * 3 + k], V.i[k * 3 + j], tr);
                ti = fmad(ar[a * 3 + k], V.i[k * 3 + j], ti);
                ti = fmad(ai[a * 3 + k], V.r[k * 3 + j], ti);
            }
            T.r[a * 3 + j] = tr;
            T.i[a * 3 + j] = ti;
        }
    }
    return T;
}

// AH parts of V^+ A V for the anti-Hermitian A with parts g (as ah_mul).  V^+ A V is
// anti-Hermitian too, so its parts are d'_a = Im (V^+ A V)_aa and o'_ab = 2 (V^+ A V)_ab.
// V is a full (three-row) matrix.
template <typename R>
__device__ __forceinline__ AH<R> ah_conj(const AH<R>& g, const Mat<R>& V) {
    constexpr int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
    const Mat<R> T = ah_mul(g, V);
    // (V^+ T)_ab = sum_k conj(V_ka) T_kb
    AH<R> h;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        R im = V.r[a] * T.i[a];
        im = fmad(-V.i[a], T.r[a], im);
#pragma unroll
        for (int k = 1; k < 3; ++k) {
            im = fmad(V.r[k * 3 + a], T.i[k * 3 + a], im);
            im = fmad(-V.i[k * 3 + a], T.r[k * 3 + a], im);
        }
        h.d[a] = im;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int a = pa[q], b = pb[q];
        R xr = V.r[a] * T.r[b];
        R xi = V.r[a] * T.i[b];
        xr = fmad(V.i[a], T.i[b], xr);
        xi = fmad(-V.i[a], T.r[b], xi);
#pragma unroll
        for (int k = 1; k < 3; ++k) {
            xr = fmad(V.r[k * 3 + a], T.r[k * 3 + b], xr);
            xr = fmad(V.i[k * 3 + a], T.i[k * 3 + b], xr);
            xi = fmad(V.r[k * 3 + a], T.i[k * 3 + b], xi);
            xi = fmad(-V.i[k * 3 + a], T.r[k * 3 + b], xi);
        }
        h.or_[q] = R(2) * xr;
        h.oi[q] = R(2) * xi;
    }
    return h;
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
    const uint32_t smu = L.stride(mu), snu = L.stride(nu);
    const int nmu = L.ext(mu), nnu = L.ext(nu);
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

// the stored link that ext<S, AX, a1, a2> multiplies by (daggered in the product for S < 0)
template <int S, int AX, int a1, int a2, typename FV>
__device__ __forceinline__ Mat<typename FV::real> ext_link(const Quad<FV>& q) {
    const int ax = AX == 1 ? q.ax1 : q.ax2;
    constexpr int b1 = (AX == 1 && S < 0) ? a1 + 1 : a1;
    constexpr int b2 = (AX == 2 && S < 0) ? a2 + 1 : a2;
    return load_link(q.f, ax, q.sp + q.T1[b1] + q.T2[b2]);
}

// M * U (S > 0) or M * U^+ (S < 0) for a link U from ext_link
template <int S, bool C2, typename R>
__device__ __forceinline__ Mat<R> ext_by(const Mat<R>& M, const Mat<R>& U) {
    if constexpr (S > 0)
        return C2 ? mul2(M, U) : mul(M, U);
    else
        return C2 ? mul2_nd(M, U) : mul_nd(M, U);
}

// M * (link along d1 or d2 leaving P(a1, a2)); with C2 only rows 0 and 1 are formed
template <int S, int AX, int a1, int a2, bool C2 = false, typename FV>
__device__ __forceinline__ Mat<typename FV::real> ext(const Quad<FV>& q, const Mat<typename FV::real>& M) {
    return ext_by<S, C2>(M, ext_link<S, AX, a1, a2>(q));
}

// acc_ah on chain ends; with C2 the chains carry rows 0 and 1 and row 2 is rebuilt here
template <bool C2, typename R>
__device__ __forceinline__ void acc_leaf(AH<R>& h, R w, const Mat<R>& CA, const Mat<R>& CB) {
    if constexpr (C2) {
        Mat<R> A = CA, B = CB;
        fill_row2(A);
        fill_row2(B);
        acc_ah(h, w, A, B);
    } else {
        acc_ah(h, w, CA, CB);
    }
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

// Lines of length 2 and 3 from x along one direction, handed from a quadrant (where the
// direction is d2) to the next one (where it is d1); both build them with the same
// operations, so the hand-over leaves every result bit-identical.
template <typename R>
struct Lines {
    Mat<R> l2, l3;
};

// the seven rectangles (1,1),(1,2),(1,3),(2,1),(2,2),(3,1),(3,3), a along d1 and b along
// d2, or only the 1x1 leaf when CLOVER is set (cloverLoopMuNuCoord, src/wloops.F90:129-199).
// With IN the d1 lines of length 2 and 3 come from `in`; with OUT the d2 lines are saved
// to `out` for the next quadrant.
template <int S1, int S2, bool CLOVER, bool IN = false, bool OUT = false, bool C2 = false, typename FV>
__device__ __forceinline__ void quadrant(const Quad<FV>& q, AH<typename FV::real>& h,
                                         const Lines<typename FV::real>* in = nullptr,
                                         Lines<typename FV::real>* out = nullptr) {
    using R = typename FV::real;
    const R w11 = CLOVER ? R(1) : R(kK1);
    const R w22 = R(kK2), w33 = R(kK5), w12 = R(0.5 * kK3), w13 = R(0.5 * kK4);
    // row a = 1
    const Mat<R> L1 = first_link<S1, 1, 0, 0>(q);   // Line(x,d1,1)
    Mat<R> CA = ext<S2, 2, 1, 0, C2>(q, L1);            // CA(1,1)
    Mat<R> Ld = first_link<S2, 2, 0, 0>(q);         // Line(x,d2,1)
    Mat<R> CB1 = ext<S1, 1, 0, 1, C2>(q, Ld);           // CB(1,1)
    acc_leaf<C2>(h, w11, CA, CB1);                        // W(1,1)
    if constexpr (!CLOVER) {
        FLUE_SCHED_FENCE();
        CA = ext<S2, 2, 1, 1, C2>(q, CA);               // CA(1,2)
        Ld = ext<S2, 2, 0, 1, C2>(q, Ld);               // Line(x,d2,2)
        if constexpr (OUT) out->l2 = Ld;
        Mat<R> CB2 = ext<S1, 1, 0, 2, C2>(q, Ld);       // CB(2,1)
        acc_leaf<C2>(h, w12, CA, CB2);                    // W(1,2)
        FLUE_SCHED_FENCE();
        CA = ext<S2, 2, 1, 2, C2>(q, CA);               // CA(1,3)
        Ld = ext<S2, 2, 0, 2, C2>(q, Ld);               // Line(x,d2,3)
        if constexpr (OUT) out->l3 = Ld;
        Mat<R> CB3 = ext<S1, 1, 0, 3, C2>(q, Ld);       // CB(3,1)
        acc_leaf<C2>(h, w13, CA, CB3);                    // W(1,3)
        FLUE_SCHED_FENCE();
        // row a = 2
        Mat<R> L;
        if constexpr (IN)
            L = in->l2;                             // Line(x,d1,2)
        else
            L = ext<S1, 1, 1, 0, C2>(q, L1);
        CA = ext<S2, 2, 2, 0, C2>(q, L);                // CA(2,1)
        CB1 = ext<S1, 1, 1, 1, C2>(q, CB1);             // CB(1,2)
        acc_leaf<C2>(h, w12, CA, CB1);                    // W(2,1)
        FLUE_SCHED_FENCE();
        CA = ext<S2, 2, 2, 1, C2>(q, CA);               // CA(2,2)
        CB2 = ext<S1, 1, 1, 2, C2>(q, CB2);             // CB(2,2)
        acc_leaf<C2>(h, w22, CA, CB2);                    // W(2,2)
        FLUE_SCHED_FENCE();
        // row a = 3
        if constexpr (IN)
            L = in->l3;                             // Line(x,d1,3)
        else
            L = ext<S1, 1, 2, 0, C2>(q, L);
        CA = ext<S2, 2, 3, 0, C2>(q, L);                // CA(3,1)
        CB1 = ext<S1, 1, 2, 1, C2>(q, CB1);             // CB(1,3)
        acc_leaf<C2>(h, w13, CA, CB1);                    // W(3,1)
        FLUE_SCHED_FENCE();
        CA = ext<S2, 2, 3, 1, C2>(q, CA);               // CA(3,2)
        CA = ext<S2, 2, 3, 2, C2>(q, CA);               // CA(3,3)
        CB3 = ext<S1, 1, 1, 3, C2>(q, CB3);             // CB(3,2)
        CB3 = ext<S1, 1, 2, 3, C2>(q, CB3);             // CB(3,3)
        acc_leaf<C2>(h, w33, CA, CB3);                    // W(3,3)
    }
    FLUE_SCHED_FENCE();
}

// quadrant() (non-clover) with a one-link lookahead: each group of products also loads the
// link that the next group's first product needs, so that product starts on a link already
// in registers while the group's other links are loaded.  (The scheduler otherwise issues
// each link's loads just before the product that uses it.)  The products and their order are
// those of quadrant(), so the results are bit-identical.  The FP64 forms at one wave per SIMD
// use it: there no other wave covers the load latency, and the registers hold one more link.
template <int S1, int S2, bool IN, bool OUT, bool C2, typename FV>
__device__ __forceinline__ void quadrant_la(const Quad<FV>& q, AH<typename FV::real>& h,
                                            const Lines<typename FV::real>* in = nullptr,
                                            Lines<typename FV::real>* out = nullptr) {
    using R = typename FV::real;
    const R w11 = R(kK1), w22 = R(kK2), w33 = R(kK5), w12 = R(0.5 * kK3), w13 = R(0.5 * kK4);
    // row a = 1
    const Mat<R> L1 = first_link<S1, 1, 0, 0>(q);               // Line(x,d1,1)
    Mat<R> Ld = first_link<S2, 2, 0, 0>(q);                    // Line(x,d2,1)
    Mat<R> U = ext_link<S2, 2, 1, 0>(q), V = ext_link<S1, 1, 0, 1>(q);
    Mat<R> P = ext_link<S2, 2, 1, 1>(q), Pn;
    Mat<R> CA = ext_by<S2, C2>(L1, U);                               // CA(1,1)
    Mat<R> CB1 = ext_by<S1, C2>(Ld, V);                              // CB(1,1)
    acc_leaf<C2>(h, w11, CA, CB1);                                   // W(1,1)
    FLUE_SCHED_FENCE();
    U = ext_link<S2, 2, 0, 1>(q);
    V = ext_link<S1, 1, 0, 2>(q);
    Pn = ext_link<S2, 2, 1, 2>(q);
    CA = ext_by<S2, C2>(CA, P);                                      // CA(1,2)
    Ld = ext_by<S2, C2>(Ld, U);                                      // Line(x,d2,2)
    if constexpr (OUT) out->l2 = Ld;
    Mat<R> CB2 = ext_by<S1, C2>(Ld, V);                              // CB(2,1)
    acc_leaf<C2>(h, w12, CA, CB2);                                   // W(1,2)
    P = Pn;
    FLUE_SCHED_FENCE();
    U = ext_link<S2, 2, 0, 2>(q);
    V = ext_link<S1, 1, 0, 3>(q);
    if constexpr (IN)
        Pn = ext_link<S2, 2, 2, 0>(q);
    else
        Pn = ext_link<S1, 1, 1, 0>(q);
    CA = ext_by<S2, C2>(CA, P);                                      // CA(1,3)
    Ld = ext_by<S2, C2>(Ld, U);                                      // Line(x,d2,3)
    if constexpr (OUT) out->l3 = Ld;
    Mat<R> CB3 = ext_by<S1, C2>(Ld, V);                              // CB(3,1)
    acc_leaf<C2>(h, w13, CA, CB3);                                   // W(1,3)
    P = Pn;
    FLUE_SCHED_FENCE();
    // row a = 2
    if constexpr (!IN) U = ext_link<S2, 2, 2, 0>(q);
    V = ext_link<S1, 1, 1, 1>(q);
    Pn = ext_link<S2, 2, 2, 1>(q);
    Mat<R> L;
    if constexpr (IN) {
        L = in->l2;                                                  // Line(x,d1,2)
        CA = ext_by<S2, C2>(L, P);                                   // CA(2,1)
    } else {
        L = ext_by<S1, C2>(L1, P);
        CA = ext_by<S2, C2>(L, U);
    }
    CB1 = ext_by<S1, C2>(CB1, V);                                    // CB(1,2)
    acc_leaf<C2>(h, w12, CA, CB1);                                   // W(2,1)
    P = Pn;
    FLUE_SCHED_FENCE();
    V = ext_link<S1, 1, 1, 2>(q);
    if constexpr (IN)
        Pn = ext_link<S2, 2, 3, 0>(q);
    else
        Pn = ext_link<S1, 1, 2, 0>(q);
    CA = ext_by<S2, C2>(CA, P);                                      // CA(2,2)
    CB2 = ext_by<S1, C2>(CB2, V);                                    // CB(2,2)
    acc_leaf<C2>(h, w22, CA, CB2);                                   // W(2,2)
    P = Pn;
    FLUE_SCHED_FENCE();
    // row a = 3
    if constexpr (!IN) U = ext_link<S2, 2, 3, 0>(q);
    V = ext_link<S1, 1, 2, 1>(q);
    Pn = ext_link<S2, 2, 3, 1>(q);
    if constexpr (IN) {
        L = in->l3;                                                  // Line(x,d1,3)
        CA = ext_by<S2, C2>(L, P);                                   // CA(3,1)
    } else {
        L = ext_by<S1, C2>(L, P);
        CA = ext_by<S2, C2>(L, U);
    }
    CB1 = ext_by<S1, C2>(CB1, V);                                    // CB(1,3)
    acc_leaf<C2>(h, w13, CA, CB1);                                   // W(3,1)
    P = Pn;
    FLUE_SCHED_FENCE();
    U = ext_link<S2, 2, 3, 2>(q);
    V = ext_link<S1, 1, 1, 3>(q);
    Pn = ext_link<S1, 1, 2, 3>(q);
    CA = ext_by<S2, C2>(CA, P);                                      // CA(3,2)
    CA = ext_by<S2, C2>(CA, U);                                      // CA(3,3)
    CB3 = ext_by<S1, C2>(CB3, V);                                    // CB(3,2)
    CB3 = ext_by<S1, C2>(CB3, Pn);                                   // CB(3,3)
    acc_leaf<C2>(h, w33, CA, CB3);                                   // W(3,3)
    FLUE_SCHED_FENCE();
}

// Per-thread parking slots in LDS for chain matrices that wait across a row of the
// quadrant: two 3x3 complex matrices per thread, laid out [slot][18 reals][thread] so a
// wave's stores and loads are consecutive.  A thread only touches its own column, so no
// barrier is needed.
template <typename R, int BS>
struct Park {
    R* col;  // &mem[threadIdx.x] of a __shared__ R mem[2 * 18 * BS]
    __device__ __forceinline__ void put(int slot, const Mat<R>& m) const {
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            col[(slot * 18 + e) * BS] = m.r[e];
            col[(slot * 18 + 9 + e) * BS] = m.i[e];
        }
    }
    __device__ __forceinline__ Mat<R> get(int slot) const {
        Mat<R> m;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            m.r[e] = col[(slot * 18 + e) * BS];
            m.i[e] = col[(slot * 18 + 9 + e) * BS];
        }
        return m;
    }
};

// The seven rectangles of quadrant() with the chains that wait for a later row parked in
// LDS: CB(1, .) from its first leaf on and CB(2, .) between rows 1 and 2; CB(3, .) is built
// again in row 3 (three products more than quadrant() without line hand-over).  At most
// about five matrices are live at once, so the FP64 form fits two waves per SIMD.
template <int S1, int S2, bool C2, typename FV, typename PK>
__device__ __forceinline__ void quadrant_parked(const Quad<FV>& q, AH<typename FV::real>& h, const PK& pk) {
    using R = typename FV::real;
    const R w11 = R(kK1), w22 = R(kK2), w33 = R(kK5), w12 = R(0.5 * kK3), w13 = R(0.5 * kK4);
    // row a = 1
    Mat<R> CA = ext<S2, 2, 1, 0, C2>(q, first_link<S1, 1, 0, 0>(q));  // CA(1,1)
    Mat<R> Ld = first_link<S2, 2, 0, 0>(q);                            // Line(x,d2,1)
    {
        const Mat<R> CB = ext<S1, 1, 0, 1, C2>(q, Ld);                 // CB(1,1)
        acc_leaf<C2>(h, w11, CA, CB);
        pk.put(0, CB);
    }
    FLUE_SCHED_FENCE();
    CA = ext<S2, 2, 1, 1, C2>(q, CA);                                  // CA(1,2)
    Ld = ext<S2, 2, 0, 1, C2>(q, Ld);                                  // Line(x,d2,2)
    {
        const Mat<R> CB = ext<S1, 1, 0, 2, C2>(q, Ld);                 // CB(2,1)
        acc_leaf<C2>(h, w12, CA, CB);
        pk.put(1, CB);
    }
    FLUE_SCHED_FENCE();
    CA = ext<S2, 2, 1, 2, C2>(q, CA);                                  // CA(1,3)
    Ld = ext<S2, 2, 0, 2, C2>(q, Ld);                                  // Line(x,d2,3)
    acc_leaf<C2>(h, w13, CA, ext<S1, 1, 0, 3, C2>(q, Ld));             // CB(3,1)
    FLUE_SCHED_FENCE();
    // row a = 2
    Mat<R> L = ext<S1, 1, 1, 0, C2>(q, first_link<S1, 1, 0, 0>(q));   // Line(x,d1,2)
    CA = ext<S2, 2, 2, 0, C2>(q, L);                                   // CA(2,1)
    {
        const Mat<R> CB = ext<S1, 1, 1, 1, C2>(q, pk.get(0));          // CB(1,2)
        acc_leaf<C2>(h, w12, CA, CB);
        pk.put(0, CB);
    }
    FLUE_SCHED_FENCE();
    CA = ext<S2, 2, 2, 1, C2>(q, CA);                                  // CA(2,2)
    acc_leaf<C2>(h, w22, CA, ext<S1, 1, 1, 2, C2>(q, pk.get(1)));      // CB(2,2)
    FLUE_SCHED_FENCE();
    // row a = 3
    L = ext<S1, 1, 2, 0, C2>(q, L);                                    // Line(x,d1,3)
    CA = ext<S2, 2, 3, 0, C2>(q, L);                                   // CA(3,1)
    acc_leaf<C2>(h, w13, CA, ext<S1, 1, 2, 1, C2>(q, pk.get(0)));      // CB(1,3)
    FLUE_SCHED_FENCE();
    CA = ext<S2, 2, 3, 1, C2>(q, CA);                                  // CA(3,2)
    CA = ext<S2, 2, 3, 2, C2>(q, CA);                                  // CA(3,3)
    FLUE_SCHED_FENCE();
    {
        Mat<R> CB = ext<S2, 2, 0, 1, C2>(q, first_link<S2, 2, 0, 0>(q));  // Line(x,d2,2)
        CB = ext<S2, 2, 0, 2, C2>(q, CB);                                  // Line(x,d2,3)
        CB = ext<S1, 1, 0, 3, C2>(q, CB);                                  // CB(3,1)
        CB = ext<S1, 1, 1, 3, C2>(q, CB);                                  // CB(3,2)
        CB = ext<S1, 1, 2, 3, C2>(q, CB);                                  // CB(3,3)
        acc_leaf<C2>(h, w33, CA, CB);
    }
    FLUE_SCHED_FENCE();
}

// plane_ah with every quadrant in its parked form
template <bool C2, typename FV, typename PK>
__device__ __forceinline__ AH<typename FV::real> plane_ah_parked(const PlaneGeom<FV>& g, const PK& pk) {
    AH<typename FV::real> h;
    ah_zero(h);
    quadrant_parked<1, 1, C2>(make_quad<1, 1, true>(g), h, pk);
    quadrant_parked<1, -1, C2>(make_quad<1, -1, false>(g), h, pk);
    quadrant_parked<-1, -1, C2>(make_quad<-1, -1, true>(g), h, pk);
    quadrant_parked<-1, 1, C2>(make_quad<-1, 1, false>(g), h, pk);
    return h;
}

// the four quadrants in counter-clockwise order: (+mu,+nu), (+nu,-mu), (-mu,-nu), (-nu,+mu).
// Each quadrant's d2 is the next one's d1; with SHARE the lines along it are computed once.
// With C2 the link chains carry two rows (SU(3) links only; see mul2).  LA (with SHARE) takes
// the quadrants in their lookahead form, quadrant_la.
template <bool CLOVER, bool SHARE = false, bool C2 = false, bool LA = false, typename FV>
__device__ __forceinline__ AH<typename FV::real> plane_ah(const PlaneGeom<FV>& g) {
    using R = typename FV::real;
    AH<R> h;
    ah_zero(h);
    if constexpr (LA && SHARE && !CLOVER) {
        Lines<R> a, b;
        quadrant_la<1, 1, false, true, C2>(make_quad<1, 1, true>(g), h, nullptr, &a);
        quadrant_la<1, -1, true, true, C2>(make_quad<1, -1, false>(g), h, &a, &b);
        quadrant_la<-1, -1, true, true, C2>(make_quad<-1, -1, true>(g), h, &b, &a);
        quadrant_la<-1, 1, true, false, C2>(make_quad<-1, 1, false>(g), h, &a, nullptr);
    } else if constexpr (SHARE && !CLOVER) {
        Lines<R> a, b;
        quadrant<1, 1, false, false, true, C2>(make_quad<1, 1, true>(g), h, nullptr, &a);
        quadrant<1, -1, false, true, true, C2>(make_quad<1, -1, false>(g), h, &a, &b);
        quadrant<-1, -1, false, true, true, C2>(make_quad<-1, -1, true>(g), h, &b, &a);
        quadrant<-1, 1, false, true, false, C2>(make_quad<-1, 1, false>(g), h, &a, nullptr);
    } else {
        quadrant<1, 1, CLOVER, false, false, C2>(make_quad<1, 1, true>(g), h);
        quadrant<1, -1, CLOVER, false, false, C2>(make_quad<1, -1, false>(g), h);
        quadrant<-1, -1, CLOVER, false, false, C2>(make_quad<-1, -1, true>(g), h);
        quadrant<-1, 1, CLOVER, false, false, C2>(make_quad<-1, 1, false>(g), h);
    }
    return h;
}

// sum_ab |F_ab|^2 = -ReTr(F F) for F = A^T/4 with A the accumulated anti-Hermitian part
// (FP64): |F_aa|^2 = d_a^2/16 and |F_ab|^2 + |F_ba|^2 = |o_ab|^2/32
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

// -ReTr F_i^2 summed over the three magnetic planes (y,x), (x,z), (z,y) at site s
// (src/wloops.F90:502-506, 530-533)
template <typename FV>
__device__ __forceinline__ double magnetic_site(const Lat& L, const FV& f, uint32_t s) {
    // FP32: two-row chains and shared lines, as in the single-field kernel
    constexpr bool kC2 = sizeof(typename FV::real) == 4;
    int c[4];
    site_coords(L, s, c);
    double acc = 0.0;
#pragma unroll 1
    for (int pl = 0; pl < 3; ++pl) {
        const int mu = pl == 0 ? 2 : (pl == 1 ? 1 : 3);
        const int nu = pl == 0 ? 1 : (pl == 1 ? 3 : 2);
        const int cmu = pl == 0 ? c[2] : (pl == 1 ? c[1] : c[3]);
        const int cnu = pl == 0 ? c[1] : (pl == 1 ? c[3] : c[2]);
        acc += f_norm2(plane_ah<false, kC2, kC2>(plane_geom(L, f, s, cmu, cnu, mu, nu)));
    }
    return acc;
}


}  // namespace
}  // namespace flue
