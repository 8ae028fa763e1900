// k_b2grad.hip -- adjoint of the magnetic B^2 (gradient with respect to every link).
//
// B = -(1/V) sum_x sum_i ReTr F_i(x)^2 with F_i = [A_i]^T / 4 and A_i = (Q_i - Q_i^+)/2
// the anti-Hermitian part of the weighted leaf sum of plane i (src/wloops.F90:411-423,
// 502-543).  dB = -(1/(8V)) sum_{x,i} ReTr(A_i(x) dQ_i(x)), so for a link U every leaf
// W = P U S (or P U^+ S) of every site x and magnetic plane through U contributes
// +k U S A(x) P (or -k S A(x) P U^+), and the gradient is (1/(16V)) P_TA of the sum.
//
// The leaves through U are enumerated as rectangles: each a x b rectangle of the
// 5-loop set that contains U is a leaf of each of its four corners (one per quadrant,
// all with the weight k_ab).  The clover leaves are recomputed per site in a first pass
// that stores only A_i(x).  Two forms then build the gradient:
//
// gather (variant 1): every link walks each rectangle through it from its far end with the
// running product X (from the walk start to the current vertex), which gives all four
// corner insertions at once: sum_c X_c A(c) X_c^+ F with F the walk's full product (links
// are unitary, so the product from a corner to the walk's end is X_c^+ F).  A rectangle
// of perimeter n is walked by each of its n links: n(n + 9) products per rectangle.
//
// reverse (variant 2, the default): every rectangle is walked once, from its lower-left
// corner x0, and hands each of its n links its term; see b2_rect_kernel.  About 4n + 7
// products per rectangle; the corner insertions and the per-link terms are anti-Hermitian,
// so they are carried as their nine anti-Hermitian parts and moved as V^+ A V (ah_conj,
// cheaper than two full products).  The terms go to fixed slots of a buffer, one per
// (plane, rectangle shape, step of the loop) and base site, and a second pass adds the 52
// slots of each link in a fixed order, so the result is deterministic without atomics.
// Rectangles lie in spatial planes, so time slices are independent: the slot buffer covers a
// slab of whole slices and the two passes run slab by slab.
#include "flue_clover.hpp"
#include "flue_flow.hpp"
#include "flue_launch.hpp"
#include "flue_su3x.hpp"

namespace flue {

namespace {

constexpr int kBS = 256;
// magnetic planes (mu, nu) in 0-based directions: (y,x), (x,z), (z,y)  (src/wloops.F90:502-506)
__constant__ int kPlaneMu[3] = {2, 1, 3};
__constant__ int kPlaneNu[3] = {1, 3, 2};

template <int A_, int B_>
struct AB {
    static constexpr int a = A_, b = B_;
};

// c[d] for a spatial direction d known only at run time (selects, so c stays in registers)
__device__ __forceinline__ int coord_of(const int c[4], int d) { return d == 1 ? c[1] : (d == 2 ? c[2] : c[3]); }

// A_i(x) stored as 9 planes per magnetic plane: d0 d1 d2, o01r o02r o12r, o01i o02i o12i
template <typename FV>
__global__ void __launch_bounds__(kBS) magnetic_ah_kernel(Lat L, FV f, typename FV::real* __restrict__ Abuf) {
    using R = typename FV::real;
    const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
#pragma unroll 1
    for (int pl = 0; pl < 3; ++pl) {
        const int mu = pl == 0 ? 2 : (pl == 1 ? 1 : 3);
        const int nu = pl == 0 ? 1 : (pl == 1 ? 3 : 2);
        const int cmu = pl == 0 ? c[2] : (pl == 1 ? c[1] : c[3]);
        const int cnu = pl == 0 ? c[1] : (pl == 1 ? c[3] : c[2]);
        const AH<R> h = plane_ah<false, sizeof(R) == 4, sizeof(R) == 4>(plane_geom(L, f, s, cmu, cnu, mu, nu));
        R* o = Abuf + (size_t)pl * 9 * L.V + s;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            o[(size_t)k * L.V] = h.d[k];
            o[(size_t)(3 + k) * L.V] = h.or_[k];
            o[(size_t)(6 + k) * L.V] = h.oi[k];
        }
    }
}

// A = (Q - Q^+)/2 from the stored anti-Hermitian data: A_aa = i d_a, A_ab = o_ab/2,
// A_ba = -conj(o_ab)/2 (a < b)
template <typename R>
__device__ __forceinline__ Mat<R> load_ah(const R* __restrict__ Abuf, uint32_t V, int pl, uint32_t s) {
    const R* p = Abuf + (size_t)pl * 9 * V + s;
    Mat<R> A;
    constexpr int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        A.r[k * 4] = R(0);
        A.i[k * 4] = p[(size_t)k * V];
        const R orr = p[(size_t)(3 + k) * V] * R(0.5), oi = p[(size_t)(6 + k) * V] * R(0.5);
        A.r[pa[k] * 3 + pb[k]] = orr;
        A.i[pa[k] * 3 + pb[k]] = oi;
        A.r[pb[k] * 3 + pa[k]] = -orr;
        A.i[pb[k] * 3 + pa[k]] = oi;
    }
    return A;
}

// Y += X A X^+
template <typename R>
__device__ __forceinline__ void add_conj(Mat<R>& Y, const Mat<R>& X, const Mat<R>& A) {
    const Mat<R> T = mul_nd(mul(X, A), X);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        Y.r[e] += T.r[e];
        Y.i[e] += T.i[e];
    }
}

// a walker over one plane (directions ax_mu, ax_nu) that tracks the site index
struct PlaneWalker {
    uint32_t s;
    int ci, cj;  // coordinates along ax_mu and ax_nu
};

__device__ __forceinline__ void pw_step(const Lat& L, PlaneWalker& w, int ax_mu, int ax_nu, int di, int dj) {
    if (di > 0) {
        if (w.ci == L.ext(ax_mu) - 1) {
            w.ci = 0;
            w.s -= (uint32_t)(L.ext(ax_mu) - 1) * L.stride(ax_mu);
        } else {
            ++w.ci;
            w.s += L.stride(ax_mu);
        }
    } else if (di < 0) {
        if (w.ci == 0) {
            w.ci = L.ext(ax_mu) - 1;
            w.s += (uint32_t)(L.ext(ax_mu) - 1) * L.stride(ax_mu);
        } else {
            --w.ci;
            w.s -= L.stride(ax_mu);
        }
    }
    if (dj > 0) {
        if (w.cj == L.ext(ax_nu) - 1) {
            w.cj = 0;
            w.s -= (uint32_t)(L.ext(ax_nu) - 1) * L.stride(ax_nu);
        } else {
            ++w.cj;
            w.s += L.stride(ax_nu);
        }
    } else if (dj < 0) {
        if (w.cj == 0) {
            w.cj = L.ext(ax_nu) - 1;
            w.s += (uint32_t)(L.ext(ax_nu) - 1) * L.stride(ax_nu);
        } else {
            --w.cj;
            w.s -= L.stride(ax_nu);
        }
    }
}

// direction (di, dj) of step k of the counter-clockwise a x b loop from its lower-left corner
__device__ __forceinline__ void loop_dir(int k, int a, int b, int& di, int& dj) {
    di = dj = 0;
    if (k < a)
        di = 1;
    else if (k < a + b)
        dj = 1;
    else if (k < 2 * a + b)
        di = -1;
    else
        dj = -1;
}

__device__ __forceinline__ bool is_corner(int k, int a, int b) {
    return k == 0 || k == a || k == a + b || k == 2 * a + b;
}

// gradient of B for the link U_mu(x), mu a spatial direction
template <typename FV>
__global__ void __launch_bounds__(kBS) b2_grad_kernel(Lat L, FV f, const typename FV::real* __restrict__ Abuf,
                                                      double scale, typename FV::real* __restrict__ grad) {
    using R = typename FV::real;
    const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
    // temporal links do not enter B
    {
        Mat<R> Z;
#pragma unroll
        for (int e = 0; e < 9; ++e) Z.r[e] = Z.i[e] = R(0);
        store_link(grad, L.V, 0, s, Z);
    }
    constexpr double k5 = 1.0 / 180.0, k1 = (19.0 / 9.0) - 55.0 * k5, k2 = (1.0 / 36.0) - 16.0 * k5,
                     k3 = 64.0 * k5 - (32.0 / 45.0), k4 = (1.0 / 15.0) - 6.0 * k5;
    const int shape_a[7] = {1, 2, 3, 1, 2, 1, 3};
    const int shape_b[7] = {1, 2, 3, 2, 1, 3, 1};
    const R shape_k[7] = {R(k1), R(k2), R(k5), R(0.5 * k3), R(0.5 * k3), R(0.5 * k4), R(0.5 * k4)};
#pragma unroll 1
    for (int mu = 1; mu < 4; ++mu) {
        const Mat<R> U = load_link(f, mu, s);
        Mat<R> G;
#pragma unroll
        for (int e = 0; e < 9; ++e) G.r[e] = G.i[e] = R(0);
#pragma unroll 1
        for (int pl = 0; pl < 3; ++pl) {
            const int pmu = kPlaneMu[pl], pnu = kPlaneNu[pl];
            if (pmu != mu && pnu != mu) continue;
            const bool along_mu = (pmu == mu);
#pragma unroll 1
            for (int sh = 0; sh < 7; ++sh) {
                const int a = shape_a[sh], b = shape_b[sh], n = 2 * (a + b);
                const int len = along_mu ? a : b;
#pragma unroll 1
                for (int side = 0; side < 2; ++side) {
#pragma unroll 1
                    for (int pos = 0; pos < len; ++pos) {
                        // step index jl of the link in the loop and whether it is traversed forward
                        int jl;
                        bool fwd;
                        if (along_mu) {
                            fwd = side == 0;                       // bottom (+mu) or top (-mu)
                            jl = fwd ? pos : 2 * a + b - 1 - pos;
                        } else {
                            fwd = side == 0;                       // right (+nu) or left (-nu)
                            jl = fwd ? a + pos : 2 * a + 2 * b - 1 - pos;
                        }
                        // the walk starts at the vertex after the link: its far end when
                        // traversed forward, x itself when traversed backward
                        PlaneWalker w{s, c[pmu], c[pnu]};
                        if (fwd) pw_step(L, w, pmu, pnu, along_mu ? 1 : 0, along_mu ? 0 : 1);
                        Mat<R> X = identity<R>(), Y;
#pragma unroll
                        for (int e = 0; e < 9; ++e) Y.r[e] = Y.i[e] = R(0);
                        int k = jl + 1 == n ? 0 : jl + 1;
                        for (int t = 0; t < n; ++t) {
                            if (is_corner(k, a, b)) add_conj(Y, X, load_ah(Abuf, L.V, pl, w.s));
                            if (t < n - 1) {
                                int di, dj;
                                loop_dir(k, a, b, di, dj);
                                const int ax = di != 0 ? pmu : pnu;
                                if (di + dj > 0) {
                                    X = mul(X, load_link(f, ax, w.s));
                                    pw_step(L, w, pmu, pnu, di, dj);
                                } else {
                                    pw_step(L, w, pmu, pnu, di, dj);
                                    X = mul_nd(X, load_link(f, ax, w.s));
                                }
                            }
                            k = k + 1 == n ? 0 : k + 1;
                        }
                        // forward: U Y F; backward: -Y F U^+
                        const Mat<R> YF = mul(Y, X);
                        const Mat<R> M = fwd ? mul(U, YF) : mul_nd(YF, U);
                        const R w8 = fwd ? shape_k[sh] : -shape_k[sh];
#pragma unroll
                        for (int e = 0; e < 9; ++e) {
                            G.r[e] = fmad(w8, M.r[e], G.r[e]);
                            G.i[e] = fmad(w8, M.i[e], G.i[e]);
                        }
                    }
                }
            }
        }
        Mat<R> P = proj_ta(G);
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            P.r[e] *= (R)scale;
            P.i[e] *= (R)scale;
        }
        store_link(grad, L.V, mu, s, P);
    }
}

// ---------------------------------------------------------------- reverse mode
//
// Rectangle a x b in plane (mu, nu) with lower-left corner x0: its counter-clockwise loop
// has vertices v_k = P(i, j) = x0 + i mu + j nu, links L_k (k = 0 .. n-1, n = 2(a + b):
// a steps +mu, b steps +nu, a steps -mu, b steps -nu), prefixes X_k = L_0 ... L_{k-1} and
// the loop W0 = X_n.  Its corners sit at steps 0, a, a+b, 2a+b; with the corner sums
// transported to x0, Ac~ = X_c A(v_c) X_c^+, the leaves of the rectangle at its four
// corners give the link of step k
//     forward step  (L_k = U):    M = X_k^+ Z X_k
//     backward step (L_k = U^+):  M = -X_{k+1}^+ Z X_{k+1}
// with Z = (sum of Ac~ over the corners after step k) W0 + W0 (sum over the corners at or
// before it): one matrix per side of the rectangle.  (The leaf at corner c through link k
// is P(v_{k+1} -> c) A(c) P(c -> v_k) after the link; writing the paths with prefixes and
// using unitarity gives the forms above.)  The prefixes are the corner chains of the
// forward kernel: Line(mu, p) and CA(a, q) = Line(mu, a) Line(v_a, nu, q) on the bottom and
// right sides, and for the top and left sides X_{k+1} = W0 CB(b, m) or W0 Line(nu, m) with
// CB(b, m) = Line(nu, b) Line(v_{2a+b}, mu, m), so with Z' = W0^+ Z W0 they read
// M = -Y^+ Z' Y for the CB-side chain Y.  Per side:
//     bottom  Z  = (Aa~ + Aab~ + A2ab~) W0 + W0 A0
//     right   Z  = (Aab~ + A2ab~) W0 + W0 (A0 + Aa~)
//     top     Z' = (Ahat + A0 + Aa~ + Aab~) W0,       Ahat = Line(nu, b) A(v_{2a+b}) Line(nu, b)^+
//     left    Z' = (A0 + Aa~ + Aab~ + A2ab~) W0,      A2ab~ = W0 Ahat W0^+

constexpr int kRectSlots = 52;  // loop steps of the seven shapes: 4+8+12+6+6+8+8

// shapes in slot order, with the slot of their step 0 and their weight
// (k1, k2, k5 for 1x1, 2x2, 3x3; k3/2 for 1x2, 2x1; k4/2 for 1x3, 3x1; src/wloops.F90:219-224)
template <int A, int B>
struct RectShape;
template <> struct RectShape<1, 1> { static constexpr int off = 0;  static constexpr double w = kK1; };
template <> struct RectShape<2, 2> { static constexpr int off = 4;  static constexpr double w = kK2; };
template <> struct RectShape<3, 3> { static constexpr int off = 12; static constexpr double w = kK5; };
template <> struct RectShape<1, 2> { static constexpr int off = 24; static constexpr double w = 0.5 * kK3; };
template <> struct RectShape<2, 1> { static constexpr int off = 30; static constexpr double w = 0.5 * kK3; };
template <> struct RectShape<1, 3> { static constexpr int off = 36; static constexpr double w = 0.5 * kK4; };
template <> struct RectShape<3, 1> { static constexpr int off = 44; static constexpr double w = 0.5 * kK4; };

template <typename R>
__device__ __forceinline__ Mat<R> madd(const Mat<R>& a, const Mat<R>& b) {
    Mat<R> c;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        c.r[e] = a.r[e] + b.r[e];
        c.i[e] = a.i[e] + b.i[e];
    }
    return c;
}

// chain step M * U (two rows with C2)
template <bool C2, typename R>
__device__ __forceinline__ Mat<R> chain(const Mat<R>& M, const Mat<R>& U) {
    if constexpr (C2)
        return mul2(M, U);
    else
        return mul(M, U);
}

// slot buffer layout: a slot holds 8 coordinates for each of `pitch` bases (the slab's
// sites rounded up to 64); the 8 coordinates of 64 consecutive bases are 64 adjacent reals
// each, so one wavefront's coordinates of a slot fill one contiguous 8 x 64 run
__device__ __forceinline__ uint32_t slot_pitch(uint32_t nsl) { return (nsl + 63u) & ~63u; }
__device__ __forceinline__ size_t slot_at(int c, uint32_t o) {
    return ((size_t)(o >> 6) * 8 + (size_t)c) * 64 + (o & 63u);
}

// anti-Hermitian parts of M: d_a = Im M_aa, o_ab = M_ab - conj(M_ba)
template <typename R>
__device__ __forceinline__ AH<R> ah_of(const Mat<R>& M) {
    AH<R> h;
    constexpr int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        h.d[k] = M.i[k * 4];
        h.or_[k] = M.r[pa[k] * 3 + pb[k]] - M.r[pb[k] * 3 + pa[k]];
        h.oi[k] = M.i[pa[k] * 3 + pb[k]] + M.i[pb[k] * 3 + pa[k]];
    }
    return h;
}

// the stored parts of A_pl(s), as load_ah reads them
template <typename R>
__device__ __forceinline__ AH<R> load_ah_parts(const R* __restrict__ Abuf, uint32_t V, int pl, uint32_t s) {
    const R* p = Abuf + (size_t)pl * 9 * V + s;
    AH<R> h;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        h.d[k] = p[(size_t)k * V];
        h.or_[k] = p[(size_t)(3 + k) * V];
        h.oi[k] = p[(size_t)(6 + k) * V];
    }
    return h;
}

// slot k of this base: w * (su(3) coordinates of the traceless part of the anti-Hermitian
// matrix with parts h, coordinates as su3_to8)
template <typename R>
__device__ __forceinline__ void emit_slot(R* __restrict__ slots, uint32_t nsl, uint32_t o, int slot, R w,
                                          const AH<R>& h) {
    const R tr = (h.d[0] + h.d[1] + h.d[2]) * R(1.0 / 3.0);
    R v[8];
    v[0] = h.d[0] - tr;
    v[1] = h.d[1] - tr;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v[2 + 2 * k] = R(0.5) * h.or_[k];
        v[3 + 2 * k] = R(0.5) * h.oi[k];
    }
    const uint32_t pitch = slot_pitch(nsl);
    R* p = slots + (size_t)slot * 8 * pitch;
#pragma unroll
    for (int c = 0; c < 8; ++c) st_stream(p + slot_at(c, o), w * v[c]);
}

// parts of F A F^+ for the stored A at s (anti-Hermitian), F a full matrix
template <typename R>
__device__ __forceinline__ AH<R> corner(const Mat<R>& F, const R* __restrict__ Abuf, uint32_t V, int pl, uint32_t s) {
    return ah_conj(load_ah_parts(Abuf, V, pl, s), dag(F));
}

template <typename R>
__device__ __forceinline__ AH<R> ah_plus(const AH<R>& a, const AH<R>& b) {
    AH<R> c;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        c.d[k] = a.d[k] + b.d[k];
        c.or_[k] = a.or_[k] + b.or_[k];
        c.oi[k] = a.oi[k] + b.oi[k];
    }
    return c;
}

// W A for the anti-Hermitian A with parts h, as -(A W^+)^+
template <typename R>
__device__ __forceinline__ Mat<R> mul_ah(const Mat<R>& W, const AH<R>& h) {
    const Mat<R> T = ah_mul(h, dag(W));
    Mat<R> M;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            M.r[a * 3 + b] = -T.r[b * 3 + a];
            M.i[a * 3 + b] = T.i[b * 3 + a];
        }
    }
    return M;
}

// One rectangle shape A x B for every base site of the slab [s0, s0 + nsl) and the
// magnetic plane blockIdx.y: writes its n = 2(A + B) slots (plane, shape, step).
template <int A, int B, typename FV>
__global__ void __launch_bounds__(kBS) b2_rect_kernel(Lat L, FV f, const typename FV::real* __restrict__ Abuf,
                                                      uint32_t s0, uint32_t nsl, typename FV::real* __restrict__ slots) {
    using R = typename FV::real;
    constexpr bool C2 = sizeof(R) == 4;  // FP32: the chains carry two rows (SU(3) links)
    constexpr int OFF = RectShape<A, B>::off;
    const R w = R(RectShape<A, B>::w);
    const int pl = (int)blockIdx.y;
    const uint32_t o = xcd_block(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
    if (o >= nsl) return;
    const int mu = pl == 0 ? 2 : (pl == 1 ? 1 : 3);
    const int nu = pl == 0 ? 1 : (pl == 1 ? 3 : 2);
    const uint32_t x0 = s0 + o;
    int c[4];
    site_coords(L, x0, c);
    const uint32_t smu = L.stride(mu), snu = L.stride(nu);
    const int cmu = mu == 1 ? c[1] : (mu == 2 ? c[2] : c[3]);
    const int cnu = nu == 1 ? c[1] : (nu == 2 ? c[2] : c[3]);
    const uint32_t sp = x0 - (uint32_t)cmu * smu - (uint32_t)cnu * snu;
    uint32_t Tm[4], Tn[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        Tm[k] = (uint32_t)wrapc(cmu + k, L.ext(mu)) * smu;
        Tn[k] = (uint32_t)wrapc(cnu + k, L.ext(nu)) * snu;
    }
#define P_(i, j) (sp + Tm[i] + Tn[j])
    R* __restrict__ out = slots + (size_t)(pl * kRectSlots + OFF) * 8 * slot_pitch(nsl);

    // ---- the loop and its corner sums
    AH<R> Aa, Aab, Ahat;  // corner sums moved to x0 (anti-Hermitian parts)
    Mat<R> W0;
    {
        Mat<R> X = load_link(f, mu, P_(0, 0));  // Line(mu, 1)
#pragma unroll
        for (int p = 1; p < A; ++p) X = chain<C2>(X, load_link(f, mu, P_(p, 0)));
        Aa = corner(full_rows<C2>(X), Abuf, L.V, pl, P_(A, 0));
#pragma unroll
        for (int q = 0; q < B; ++q) X = chain<C2>(X, load_link(f, nu, P_(A, q)));  // CA(A, B)
        const Mat<R> CA = full_rows<C2>(X);
        Aab = corner(CA, Abuf, L.V, pl, P_(A, B));
        Mat<R> Y = load_link(f, nu, P_(0, 0));  // Line(nu, 1)
#pragma unroll
        for (int q = 1; q < B; ++q) Y = chain<C2>(Y, load_link(f, nu, P_(0, q)));
        Ahat = corner(full_rows<C2>(Y), Abuf, L.V, pl, P_(0, B));
#pragma unroll
        for (int p = 0; p < A; ++p) Y = chain<C2>(Y, load_link(f, mu, P_(p, B)));  // CB(B, A)
        W0 = mul_nd(CA, full_rows<C2>(Y));
    }
    const AH<R> A0 = load_ah_parts(Abuf, L.V, pl, P_(0, 0));
    const AH<R> A2ab = ah_conj(Ahat, dag(W0));
    const AH<R> T3 = ah_plus(ah_plus(A0, Aa), Aab);
    // the side sums Z enter the slots as X^+ Z X projected on its traceless anti-Hermitian
    // part; for unitary X that is X^+ AH(Z) X, so only the AH parts of Z are carried
    const AH<R> Zt = ah_of(ah_mul(ah_plus(Ahat, T3), W0));                        // top, W0-conjugated
    const AH<R> Zl = ah_of(ah_mul(ah_plus(T3, A2ab), W0));                        // left, W0-conjugated
    const AH<R> Sab = ah_plus(Aab, A2ab);
    const AH<R> Zr = ah_of(madd(ah_mul(Sab, W0), mul_ah(W0, ah_plus(A0, Aa))));   // right
    const AH<R> Zb = ah_of(madd(ah_mul(ah_plus(Aa, Sab), W0), mul_ah(W0, A0)));   // bottom

    // ---- bottom (steps 0 .. A-1, X = Line(mu, p)) and right (steps A .. A+B-1, X = CA(A, q))
    emit_slot(out, nsl, o, 0, w, Zb);
    {
        Mat<R> X = load_link(f, mu, P_(0, 0));
#pragma unroll
        for (int p = 1; p < A; ++p) {
            emit_slot(out, nsl, o, p, w, ah_conj(Zb, full_rows<C2>(X)));
            X = chain<C2>(X, load_link(f, mu, P_(p, 0)));
        }
#pragma unroll
        for (int q = 0; q < B; ++q) {
            emit_slot(out, nsl, o, A + q, w, ah_conj(Zr, full_rows<C2>(X)));
            if (q + 1 < B) X = chain<C2>(X, load_link(f, nu, P_(A, q)));
        }
    }
    // ---- left (step 2A+B+j ends at P(0, B-1-j): Y = Line(nu, B-1-j)) and top (step A+B+j
    // ends at P(A-1-j, B): Y = CB(B, A-1-j)), both backward, walked with the chain ascending
    emit_slot(out, nsl, o, 2 * A + 2 * B - 1, -w, Zl);
    {
        Mat<R> Y = load_link(f, nu, P_(0, 0));
#pragma unroll
        for (int m = 1; m < B; ++m) {
            emit_slot(out, nsl, o, 2 * A + 2 * B - 1 - m, -w, ah_conj(Zl, full_rows<C2>(Y)));
            Y = chain<C2>(Y, load_link(f, nu, P_(0, m)));
        }
#pragma unroll
        for (int m = 0; m < A; ++m) {
            emit_slot(out, nsl, o, 2 * A + B - 1 - m, -w, ah_conj(Zt, full_rows<C2>(Y)));
            if (m + 1 < A) Y = chain<C2>(Y, load_link(f, mu, P_(m, B)));
        }
    }
#undef P_
}

// Second pass: every link of the slab adds its 52 slots (26 in each of the two magnetic
// planes through its direction) in a fixed order; temporal links get 0.
// blockIdx.y is the direction.
template <typename R>
__global__ void __launch_bounds__(kBS) b2_slot_sum_kernel(Lat L, const R* __restrict__ slots, uint32_t s0,
                                                          uint32_t nsl, double scale, R* __restrict__ grad) {
    const int d = (int)blockIdx.y;
    const uint32_t o = xcd_block(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
    if (o >= nsl) return;
    const uint32_t x = s0 + o;
    Mat<R> G;
    if (d == 0) {
#pragma unroll
        for (int e = 0; e < 9; ++e) G.r[e] = G.i[e] = R(0);
        store_link_stream(grad, L.V, 0, x, G);
        return;
    }
    int c[4];
    site_coords(L, x, c);
    double acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.0;
    // slab-relative base site of the rectangle whose step `slot` is this link
    auto add = [&](int pl, int slot, uint32_t base) {
        const uint32_t pitch = slot_pitch(nsl);
        const R* p = slots + ((size_t)(pl * kRectSlots + slot) * 8) * pitch;
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] += (double)ld_stream(p + slot_at(k, base - s0));
    };
    // plane where d is the first direction (d = mu): bottom and top sides
    {
        const int pl = d == 2 ? 0 : (d == 1 ? 1 : 2);
        const int e = pl == 0 ? 1 : (pl == 1 ? 3 : 2);  // that plane's nu
        const uint32_t sd = L.stride(d), se = L.stride(e);
        const int cd = coord_of(c, d), ce = coord_of(c, e);
        const uint32_t sp = x - (uint32_t)cd * sd - (uint32_t)ce * se;
        uint32_t Td[4], Te[4];  // sites at -k along d and e
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            Td[k] = (uint32_t)wrapc(cd - k, L.ext(d)) * sd;
            Te[k] = (uint32_t)wrapc(ce - k, L.ext(e)) * se;
        }
        auto shape = [&](auto ab) {
            constexpr int A = decltype(ab)::a, B = decltype(ab)::b, OFF = RectShape<A, B>::off;
#pragma unroll
            for (int p = 0; p < A; ++p) add(pl, OFF + p, sp + Td[p] + Te[0]);              // bottom, base x - p d
#pragma unroll
            for (int j = 0; j < A; ++j) add(pl, OFF + A + B + j, sp + Td[A - 1 - j] + Te[B]);  // top
        };
        shape(AB<1, 1>()); shape(AB<2, 2>()); shape(AB<3, 3>()); shape(AB<1, 2>());
        shape(AB<2, 1>()); shape(AB<1, 3>()); shape(AB<3, 1>());
    }
    // plane where d is the second direction (d = nu): right and left sides
    {
        const int pl = d == 1 ? 0 : (d == 3 ? 1 : 2);
        const int e = pl == 0 ? 2 : (pl == 1 ? 1 : 3);  // that plane's mu
        const uint32_t sd = L.stride(d), se = L.stride(e);
        const int cd = coord_of(c, d), ce = coord_of(c, e);
        const uint32_t sp = x - (uint32_t)cd * sd - (uint32_t)ce * se;
        uint32_t Td[4], Te[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            Td[k] = (uint32_t)wrapc(cd - k, L.ext(d)) * sd;
            Te[k] = (uint32_t)wrapc(ce - k, L.ext(e)) * se;
        }
        auto shape = [&](auto ab) {
            constexpr int A = decltype(ab)::a, B = decltype(ab)::b, OFF = RectShape<A, B>::off;
#pragma unroll
            for (int q = 0; q < B; ++q) add(pl, OFF + A + q, sp + Te[A] + Td[q]);                  // right
#pragma unroll
            for (int j = 0; j < B; ++j) add(pl, OFF + 2 * A + B + j, sp + Te[0] + Td[B - 1 - j]);  // left
        };
        shape(AB<1, 1>()); shape(AB<2, 2>()); shape(AB<3, 3>()); shape(AB<1, 2>());
        shape(AB<2, 1>()); shape(AB<1, 3>()); shape(AB<3, 1>());
    }
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = acc[k] * scale;
    const Mat<double> Gd = su3_from8(v);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        G.r[e] = (R)Gd.r[e];
        G.i[e] = (R)Gd.i[e];
    }
    store_link_stream(grad, L.V, d, x, G);
}

}  // namespace

hipError_t launch_magnetic_ah(const Lat& L, const void* U, int is_f32, void* Abuf, hipStream_t st) {
    const uint32_t g = (L.V + kBS - 1) / kBS;
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        magnetic_ah_kernel<F><<<g, kBS, 0, st>>>(L, f, (typename F::real*)Abuf);
    });
    return hipGetLastError();
}

uint32_t b2_slab_slices(const Lat& L, int is_f32) {
    const uint64_t per_slice = (uint64_t)L.Vs * 3 * kRectSlots * 8;  // slot reals per time slice
    uint64_t S = kB2SlotReals / per_slice;
    if (S < 1) S = 1;
    if (S > (uint64_t)L.n[0]) S = (uint64_t)L.n[0];
    return (uint32_t)S;
}

size_t b2_slot_bytes(const Lat& L, int is_f32) {
    const uint64_t nsl = (uint64_t)b2_slab_slices(L, is_f32) * L.Vs;
    return (size_t)((nsl + 63) & ~(uint64_t)63) * 3 * kRectSlots * 8 * (is_f32 ? 4 : 8);
}

hipError_t launch_b2_grad_rev(const Lat& L, const void* U, int is_f32, const void* Abuf, void* slots, void* grad,
                              hipStream_t st) {
    const double scale = 1.0 / (16.0 * (double)L.V);
    const uint32_t S = b2_slab_slices(L, is_f32);
    hipError_t e = hipSuccess;
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        using R = typename F::real;
        const R* A = (const R*)Abuf;
        R* sl = (R*)slots;
        for (uint32_t t0 = 0; t0 < (uint32_t)L.n[0] && e == hipSuccess; t0 += S) {
            const uint32_t ns = std::min<uint32_t>(S, (uint32_t)L.n[0] - t0);
            const uint32_t s0 = t0 * L.Vs, nsl = ns * L.Vs;
            const dim3 g1((nsl + kBS - 1) / kBS, 3), g2((nsl + kBS - 1) / kBS, 4);
            b2_rect_kernel<1, 1, F><<<g1, kBS, 0, st>>>(L, f, A, s0, nsl, sl);
            b2_rect_kernel<2, 2, F><<<g1, kBS, 0, st>>>(L, f, A, s0, nsl, sl);
            b2_rect_kernel<3, 3, F><<<g1, kBS, 0, st>>>(L, f, A, s0, nsl, sl);
            b2_rect_kernel<1, 2, F><<<g1, kBS, 0, st>>>(L, f, A, s0, nsl, sl);
            b2_rect_kernel<2, 1, F><<<g1, kBS, 0, st>>>(L, f, A, s0, nsl, sl);
            b2_rect_kernel<1, 3, F><<<g1, kBS, 0, st>>>(L, f, A, s0, nsl, sl);
            b2_rect_kernel<3, 1, F><<<g1, kBS, 0, st>>>(L, f, A, s0, nsl, sl);
            b2_slot_sum_kernel<R><<<g2, kBS, 0, st>>>(L, sl, s0, nsl, scale, (R*)grad);
            e = hipGetLastError();
        }
    });
    return e;
}

hipError_t launch_b2_grad(const Lat& L, const void* U, int is_f32, const void* Abuf, void* grad, hipStream_t st) {
    const uint32_t g = (L.V + kBS - 1) / kBS;
    const double scale = 1.0 / (16.0 * (double)L.V);
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        using R = typename F::real;
        b2_grad_kernel<F><<<g, kBS, 0, st>>>(L, f, (const R*)Abuf, scale, (R*)grad);
    });
    return hipGetLastError();
}

}  // namespace flue
