// k_grad.hip -- gradients of the Wilson plaquette action and of the magnetic B^2 with
// respect to every link, and the Wilson-flow (RK3) stage built on the first.
//
// Convention (no reference counterpart): for a link U = U_mu(x) the Lie derivative is
// D^a f = d/ds f(exp(s T^a) U)|_0 with T^a = -i lambda_a / 2, and the gradient stored
// per link is G = sum_a T^a D^a f, traceless anti-Hermitian; D^a f = -2 Tr(T^a G).
//   Wilson action S = beta sum_p (1 - ReTr P/3):   G = (beta/6) P_TA(U Sigma)
//   B = -(1/V) sum_x sum_i ReTr F_i(x)^2:          G = (1/(16V)) P_TA(sum_leaves +-k M)
// where Sigma is the staple sum and, for every 5-loop leaf through U, M is the loop
// product starting at U with A = (Q - Q^+)/2 of the leaf's base site inserted at that
// site (Q the untransposed weighted leaf sum of that site and plane).
#include "flue_launch.hpp"
#include "flue_su3x.hpp"

namespace flue {

namespace {

constexpr int kBS = 256;

// staple sum Sigma_mu(x): U_mu(x) Sigma is the sum of the six plaquettes through the link
template <typename FV>
__device__ __forceinline__ Mat<typename FV::real> staple_sum(const Lat& L, const FV& f, uint32_t s, const int c[4],
                                                             int mu) {
    using R = typename FV::real;
    Mat<R> St;
#pragma unroll
    for (int e = 0; e < 9; ++e) St.r[e] = St.i[e] = R(0);
    int cm[4] = {c[0], c[1], c[2], c[3]};
    const uint32_t sm = step_fwd(L, s, cm, mu);  // x + mu
#pragma unroll 1
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        // up: U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
        int cn[4] = {c[0], c[1], c[2], c[3]};
        const uint32_t sn = step_fwd(L, s, cn, nu);
        const Mat<R> a = mul_nd(load_link(f, nu, sm), load_link(f, mu, sn));
        const Mat<R> up = mul_nd(a, load_link(f, nu, s));
        // down: U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
        int cmn[4] = {cm[0], cm[1], cm[2], cm[3]};
        const uint32_t smn = step_bwd(L, sm, cmn, nu);
        int cd[4] = {c[0], c[1], c[2], c[3]};
        const uint32_t sd = step_bwd(L, s, cd, nu);
        // = (U_mu(x-nu) U_nu(x+mu-nu))^+ U_nu(x-nu)
        const Mat<R> dn = mul_dn(mul(load_link(f, mu, sd), load_link(f, nu, smn)), load_link(f, nu, sd));
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            St.r[e] += up.r[e] + dn.r[e];
            St.i[e] += up.i[e] + dn.i[e];
        }
    }
    return St;
}

template <typename R>
__device__ __forceinline__ Mat<double> to_double(const Mat<R>& m) {
    Mat<double> d;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        d.r[e] = (double)m.r[e];
        d.i[e] = (double)m.i[e];
    }
    return d;
}

template <typename R>
__device__ __forceinline__ Mat<R> from_double(const Mat<double>& m) {
    Mat<R> d;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        d.r[e] = (R)m.r[e];
        d.i[e] = (R)m.i[e];
    }
    return d;
}

// Wilson action gradient for the four links of a site
template <typename FV>
__global__ void __launch_bounds__(kBS) wilson_grad_kernel(Lat L, FV f, double beta,
                                                          typename FV::real* __restrict__ grad) {
    using R = typename FV::real;
    const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
#pragma unroll 1
    for (int mu = 0; mu < 4; ++mu) {
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

// su(3) element X <-> 8 real coordinates (h0, h1, a, b, c, d, e, f):
//   X_00 = i h0, X_11 = i h1, X_22 = -i (h0 + h1), X_01 = a + i b, X_02 = c + i d, X_12 = e + i f
__device__ __forceinline__ void su3_to8(const Mat<double>& X, double v[8]) {
    v[0] = X.i[0];
    v[1] = X.i[4];
    v[2] = X.r[1];
    v[3] = X.i[1];
    v[4] = X.r[2];
    v[5] = X.i[2];
    v[6] = X.r[5];
    v[7] = X.i[5];
}
__device__ __forceinline__ Mat<double> su3_from8(const double v[8]) {
    Mat<double> X;
    X.r[0] = X.r[4] = X.r[8] = 0.0;
    X.i[0] = v[0];
    X.i[4] = v[1];
    X.i[8] = -(v[0] + v[1]);
    X.r[1] = v[2];
    X.i[1] = v[3];
    X.r[3] = -v[2];
    X.i[3] = v[3];
    X.r[2] = v[4];
    X.i[2] = v[5];
    X.r[6] = -v[4];
    X.i[6] = v[5];
    X.r[5] = v[6];
    X.i[5] = v[7];
    X.r[7] = -v[6];
    X.i[7] = v[7];
    return X;
}

// One stage of Luescher's RK3 Wilson flow (JHEP 08 (2010) 071, appendix C):
//   Z = eps * (-P_TA(W Sigma)),  acc <- ca * acc + cb * Z,  W_out = exp(acc) W_in
// acc is kept in FP64 as 8 su(3) coordinates per link; the exponential, the product and
// (after the last stage) the reunitarisation are done in FP64.
template <typename FV>
__global__ void __launch_bounds__(kBS) flow_stage_kernel(Lat L, FV f, typename FV::real* __restrict__ wout,
                                                         double* __restrict__ acc, double eps, double ca, double cb,
                                                         int reunit) {
    using R = typename FV::real;
    const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
#pragma unroll 1
    for (int mu = 0; mu < 4; ++mu) {
        const Mat<R> U = load_link(f, mu, s);
        const Mat<double> Z = to_double(proj_ta(mul(U, staple_sum(L, f, s, c, mu))));
        double z[8], a[8];
        su3_to8(Z, z);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            double* p = acc + (size_t)(mu * 8 + k) * L.V + s;
            const double prev = ca != 0.0 ? *p : 0.0;
            a[k] = ca * prev - cb * eps * z[k];
            *p = a[k];
        }
        Mat<double> W = mul(su3_exp(su3_from8(a)), to_double(U));
        if (reunit) fix_su3(W);
        store_link(wout, L.V, mu, s, from_double<R>(W));
    }
}

}  // namespace

uint32_t grad_blocks(const Lat& L) { return (L.V + kBS - 1) / kBS; }

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
