// flue_flow.hpp -- per-site pieces of the Wilson-action gradient and the RK3 Wilson flow,
// shared by the single-field and the ensemble kernels.
#pragma once
#include "flue_common.hpp"
#include "flue_su3x.hpp"

namespace flue {
namespace {

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
        // down: U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) = (U_mu(x-nu) U_nu(x+mu-nu))^+ U_nu(x-nu)
        int cmn[4] = {cm[0], cm[1], cm[2], cm[3]};
        const uint32_t smn = step_bwd(L, sm, cmn, nu);
        int cd[4] = {c[0], c[1], c[2], c[3]};
        const uint32_t sd = step_bwd(L, s, cd, nu);
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

// One RK3 stage for the link U_mu(s), c the coordinates of s (Luescher, JHEP 08 (2010) 071,
// appendix C):
//   Z = -P_TA(W Sigma),  acc <- ca * acc + cb * eps * Z,  W_out = exp(acc) W_in
// acc holds 8 su(3) coordinates per link in FP64 (plane (mu*8+k) of V doubles); the
// exponential, the product and the reunitarisation (reunit != 0) are FP64.  ca == 0
// starts a fresh accumulator.  Links are independent within a stage: W_in is only read.
template <typename FV>
__device__ __forceinline__ void flow_link(const Lat& L, const FV& f, typename FV::real* __restrict__ wout,
                                          double* __restrict__ acc, uint32_t s, const int c[4], int mu, double eps,
                                          double ca, double cb, int reunit) {
    using R = typename FV::real;
    {
        const Mat<R> U = load_link(f, mu, s);
        const Mat<double> Z = to_double(proj_ta(mul(U, staple_sum(L, f, s, c, mu))));
        double z[8], a[8];
        su3_to8(Z, z);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            double* p = acc + (size_t)(mu * 8 + k) * L.V + s;
            const double prev = ca != 0.0 ? ca * *p : 0.0;
            a[k] = prev - cb * eps * z[k];
            *p = a[k];
        }
        Mat<double> W = mul(su3_exp(su3_from8(a)), to_double(U));
        if (reunit) fix_su3(W);
        store_link(wout, L.V, mu, s, from_double<R>(W));
    }
}

// sum over the 6 planes of ReTr of the plaquette at site s (src/wloops.F90:947-1006)
template <typename FV>
__device__ __forceinline__ double plaquette_site(const Lat& L, const FV& f, uint32_t s, int mu0, int mu1, int nu1) {
    using R = typename FV::real;
    int c[4];
    site_coords(L, s, c);
    // forward neighbour of s along each direction
    uint32_t sf[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) sf[d] = c[d] == L.n[d] - 1 ? s - (uint32_t)(L.n[d] - 1) * L.st[d] : s + L.st[d];
    // the six planes unrolled (the enabled set is uniform), so the link loads of several
    // planes are in flight together
    double acc = 0.0;
#pragma unroll
    for (int mu = 0; mu < 3; ++mu) {
        if (mu < mu0 || mu > mu1) continue;
        const Mat<R> Umu = load_link(f, mu, s);
#pragma unroll
        for (int nu = mu + 1; nu < 4; ++nu) {
            if (nu > nu1) continue;
            const Mat<R> A = mul(Umu, load_link(f, nu, sf[mu]));
            const Mat<R> B = mul(load_link(f, nu, s), load_link(f, mu, sf[nu]));
            acc += (double)retr_nd(A, B);
        }
    }
    return acc;
}

}  // namespace
}  // namespace flue
