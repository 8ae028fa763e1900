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
// all with the weight k_ab).  Walking the rectangle once from U's far end with the
// running product X (from the walk start to the current vertex) gives all four corner
// insertions at once: sum_c X_c A(c) X_c^+ F with F the walk's full product (links are
// unitary, so the product from a corner to the walk's end is X_c^+ F).  The clover
// leaves are recomputed per site in a first pass that stores only A_i(x).
#include "flue_clover.hpp"
#include "flue_launch.hpp"
#include "flue_su3x.hpp"

namespace flue {

namespace {

constexpr int kBS = 256;
// magnetic planes (mu, nu) in 0-based directions: (y,x), (x,z), (z,y)  (src/wloops.F90:502-506)
__constant__ int kPlaneMu[3] = {2, 1, 3};
__constant__ int kPlaneNu[3] = {1, 3, 2};

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

}  // namespace

hipError_t launch_magnetic_ah(const Lat& L, const void* U, int is_f32, void* Abuf, hipStream_t st) {
    const uint32_t g = (L.V + kBS - 1) / kBS;
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        magnetic_ah_kernel<F><<<g, kBS, 0, st>>>(L, f, (typename F::real*)Abuf);
    });
    return hipGetLastError();
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
