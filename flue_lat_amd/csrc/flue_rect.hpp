// flue_rect.hpp -- rectangle-first clover sums for the magnetic B.
//
// The clover sum at a site adds, for each of the seven rectangle shapes and each rectangle
// with the site as a corner, the loop around it read from that corner (the 28 leaves of
// Loop5MuNuCoord, src/wloops.F90:234-290).  Here every rectangle is formed once, from its
// lower-left corner x in the (d1, d2) orientation (a links along d1, b along d2), and its
// leaves at the other three corners follow from the one at x:
//   W(x + a d1 + b d2) = CB^+ CA                       CA, CB the corner chains of quadrant()
//   W(x + a d1)        = V1^+ W(x) V1,   V1 = Line(x, d1, a)
//   W(x + b d2)        = V3^+ W(x) V3,   V3 = Line(x, d2, b)
// Conjugation commutes with the anti-Hermitian projection, so only AH parts are moved.  The
// rectangles that share the length a (or b) share the corner x + a d1 (x + b d2) and the line
// to it: their weighted AH parts are summed first and moved once.  Per base site that is 20
// chain products, 14 AH closings and 6 conjugations, against 74 products and 28 closings for
// the four quadrants of quadrant() with shared lines.  (Reversing the orientation of a plane
// negates every F, so B, a sum of |F|^2, may use either.)
#pragma once
#include "flue_clover.hpp"

namespace flue {
namespace {

// h += w * AH-part of (CB^+ CA): with W_ab = sum_k conj(CB_ka) CA_kb,
//   d_a += w Im W_aa,  o_ab += w (W_ab - conj(W_ba))
template <typename R>
__device__ __forceinline__ void acc_ah_dag(AH<R>& h, R w, const Mat<R>& CA, const Mat<R>& CB) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        R im = CB.r[a] * CA.i[a];
        im = fmad(-CB.i[a], CA.r[a], im);
#pragma unroll
        for (int k = 1; k < 3; ++k) {
            im = fmad(CB.r[k * 3 + a], CA.i[k * 3 + a], im);
            im = fmad(-CB.i[k * 3 + a], CA.r[k * 3 + a], im);
        }
        h.d[a] = fmad(w, im, h.d[a]);
    }
    constexpr int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const int a = pa[q], b = pb[q];
        // conj(CB_ka) CA_kb - CB_kb conj(CA_ka), summed over k in one chain per component
        R xr = R(0), xi = R(0);
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int ka = k * 3 + a, kb = k * 3 + b;
            xr = fmad(CB.r[ka], CA.r[kb], xr);
            xr = fmad(CB.i[ka], CA.i[kb], xr);
            xr = fmad(-CB.r[kb], CA.r[ka], xr);
            xr = fmad(-CB.i[kb], CA.i[ka], xr);
            xi = fmad(CB.r[ka], CA.i[kb], xi);
            xi = fmad(-CB.i[ka], CA.r[kb], xi);
            xi = fmad(-CB.i[kb], CA.r[ka], xi);
            xi = fmad(CB.r[kb], CA.i[ka], xi);
        }
        h.or_[q] = fmad(w, xr, h.or_[q]);
        h.oi[q] = fmad(w, xi, h.oi[q]);
    }
}

template <typename R>
__device__ __forceinline__ void ah_add(AH<R>& h, const AH<R>& g) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        h.d[k] += g.d[k];
        h.or_[k] += g.or_[k];
        h.oi[k] += g.oi[k];
    }
}

// One rectangle closed at both ends of its diagonal: the weighted AH part of the leaf at x
// (W = CA CB^+) goes to h0 and to the two groups it belongs to; the weighted leaf at the
// opposite corner (CB^+ CA) is returned.
template <bool C2, typename R>
__device__ __forceinline__ AH<R> rect_close(R w, const Mat<R>& CA, const Mat<R>& CB, AH<R>& h0, AH<R>& ga,
                                            AH<R>& gb) {
    const Mat<R> A = full_rows<C2>(CA), B = full_rows<C2>(CB);
    AH<R> t;
    ah_zero(t);
    acc_ah(t, w, A, B);
    ah_add(h0, t);
    ah_add(ga, t);
    ah_add(gb, t);
    AH<R> c;
    ah_zero(c);
    acc_ah_dag(c, w, A, B);
    return c;
}

// The seven rectangles with lower-left corner x in quadrant q = (+d1, +d2).  emit(da, db, p)
// adds the AH parts p to the site x + da d1 + db d2; the calls come in a fixed order.  With
// LA each group of products also loads the link of the next group's first product (P), so
// that product starts on a link in registers (for forms at one wave per SIMD).
template <bool C2, bool LA, typename FV, typename Emit>
__device__ __forceinline__ void rect_base(const Quad<FV>& q, Emit&& emit) {
    using R = typename FV::real;
    const R w11 = R(kK1), w22 = R(kK2), w33 = R(kK5), w12 = R(0.5 * kK3), w13 = R(0.5 * kK4);
    AH<R> h0, ga, gb1, gb2, gb3;  // leaves at x; current row a; columns b = 1, 2, 3
    ah_zero(h0);
    ah_zero(ga);
    ah_zero(gb1);
    ah_zero(gb2);
    ah_zero(gb3);
    Mat<R> P;
    // row a = 1
    const Mat<R> L1 = first_link<1, 1, 0, 0>(q);   // Line(x,d1,1)
    Mat<R> CA = ext<1, 2, 1, 0, C2>(q, L1);            // CA(1,1)
    const Mat<R> Ld1 = first_link<1, 2, 0, 0>(q);  // Line(x,d2,1)
    Mat<R> CB1 = ext<1, 1, 0, 1, C2>(q, Ld1);          // CB(1,1)
    if constexpr (LA) P = ext_link<1, 2, 1, 1>(q);
    emit(1, 1, rect_close<C2>(w11, CA, CB1, h0, ga, gb1));
    FLUE_SCHED_FENCE();
    CA = LA ? ext_by<1, C2>(CA, P) : ext<1, 2, 1, 1, C2>(q, CA);  // CA(1,2)
    const Mat<R> Ld2 = ext<1, 2, 0, 1, C2>(q, Ld1);    // Line(x,d2,2)
    Mat<R> CB2 = ext<1, 1, 0, 2, C2>(q, Ld2);          // CB(2,1)
    if constexpr (LA) P = ext_link<1, 2, 1, 2>(q);
    emit(1, 2, rect_close<C2>(w12, CA, CB2, h0, ga, gb2));
    FLUE_SCHED_FENCE();
    CA = LA ? ext_by<1, C2>(CA, P) : ext<1, 2, 1, 2, C2>(q, CA);  // CA(1,3)
    const Mat<R> Ld3 = ext<1, 2, 0, 2, C2>(q, Ld2);    // Line(x,d2,3)
    Mat<R> CB3 = ext<1, 1, 0, 3, C2>(q, Ld3);          // CB(3,1)
    if constexpr (LA) P = ext_link<1, 1, 1, 0>(q);
    emit(1, 3, rect_close<C2>(w13, CA, CB3, h0, ga, gb3));
    emit(1, 0, ah_conj(ga, L1));
    ah_zero(ga);
    FLUE_SCHED_FENCE();
    // row a = 2
    Mat<R> L = LA ? ext_by<1, C2>(L1, P) : ext<1, 1, 1, 0, C2>(q, L1);  // Line(x,d1,2)
    CA = ext<1, 2, 2, 0, C2>(q, L);                    // CA(2,1)
    CB1 = ext<1, 1, 1, 1, C2>(q, CB1);                 // CB(1,2)
    if constexpr (LA) P = ext_link<1, 2, 2, 1>(q);
    emit(2, 1, rect_close<C2>(w12, CA, CB1, h0, ga, gb1));
    FLUE_SCHED_FENCE();
    CA = LA ? ext_by<1, C2>(CA, P) : ext<1, 2, 2, 1, C2>(q, CA);  // CA(2,2)
    CB2 = ext<1, 1, 1, 2, C2>(q, CB2);                 // CB(2,2)
    if constexpr (LA) P = ext_link<1, 1, 2, 0>(q);
    emit(2, 2, rect_close<C2>(w22, CA, CB2, h0, ga, gb2));
    emit(2, 0, ah_conj(ga, full_rows<C2>(L)));
    ah_zero(ga);
    emit(0, 2, ah_conj(gb2, full_rows<C2>(Ld2)));  // b = 2: (1,2), (2,2)
    FLUE_SCHED_FENCE();
    // row a = 3
    L = LA ? ext_by<1, C2>(L, P) : ext<1, 1, 2, 0, C2>(q, L);  // Line(x,d1,3)
    CA = ext<1, 2, 3, 0, C2>(q, L);                    // CA(3,1)
    CB1 = ext<1, 1, 2, 1, C2>(q, CB1);                 // CB(1,3)
    if constexpr (LA) P = ext_link<1, 2, 3, 1>(q);
    emit(3, 1, rect_close<C2>(w13, CA, CB1, h0, ga, gb1));
    emit(0, 1, ah_conj(gb1, Ld1));                 // b = 1: (1,1), (2,1), (3,1)
    FLUE_SCHED_FENCE();
    CA = LA ? ext_by<1, C2>(CA, P) : ext<1, 2, 3, 1, C2>(q, CA);  // CA(3,2)
    CA = ext<1, 2, 3, 2, C2>(q, CA);                   // CA(3,3)
    CB3 = ext<1, 1, 1, 3, C2>(q, CB3);                 // CB(3,2)
    CB3 = ext<1, 1, 2, 3, C2>(q, CB3);                 // CB(3,3)
    emit(3, 3, rect_close<C2>(w33, CA, CB3, h0, ga, gb3));
    emit(3, 0, ah_conj(ga, full_rows<C2>(L)));     // a = 3: (3,1), (3,3)
    emit(0, 3, ah_conj(gb3, full_rows<C2>(Ld3)));  // b = 3: (1,3), (3,3)
    emit(0, 0, h0);
    FLUE_SCHED_FENCE();
}

}  // namespace
}  // namespace flue
