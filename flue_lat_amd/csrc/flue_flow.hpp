// flue_flow.hpp -- per-site pieces of the Wilson-action gradient and the RK3 Wilson flow,
// shared by the single-field and the ensemble kernels.
#pragma once
#include "flue_clover.hpp"
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
// acc holds 8 su(3) coordinates per link in the field's precision (plane (mu*8+k) of V
// reals); the exponential, the product and the reunitarisation (reunit != 0, the step's last
// stage, which leaves acc unwritten since no later stage reads it) are FP64.  ca == 0 starts
// a fresh accumulator.  Links are independent within a stage: W_in is only read.
template <typename FV>
__device__ __forceinline__ void flow_link(const Lat& L, const FV& f, typename FV::real* __restrict__ wout,
                                          typename FV::real* __restrict__ acc, uint32_t s, const int c[4], int mu,
                                          double eps, double ca, double cb, int reunit) {
    using R = typename FV::real;
    {
        const Mat<R> U = load_link(f, mu, s);
        const Mat<double> Z = to_double(proj_ta(mul(U, staple_sum(L, f, s, c, mu))));
        double z[8], a[8];
        su3_to8(Z, z);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            R* p = acc + (size_t)(mu * 8 + k) * L.V + s;
            const double prev = ca != 0.0 ? ca * (double)*p : 0.0;
            a[k] = prev - cb * eps * z[k];
            if (!reunit) *p = (R)a[k];
        }
        Mat<double> W = mul(su3_exp(su3_from8(a)), to_double(U));
        if (reunit) fix_su3(W);
        store_link(wout, L.V, mu, s, from_double<R>(W));
    }
}

// ---------------------------------------------------------------- plaquette-first stencil
//
// The gradient of a link is the sum of the traceless anti-Hermitian parts of the six
// plaquettes that start with it: G_mu(x) ~ P_TA(U_mu(x) Sigma) = sum over nu of P_TA of the
// "up" loop and of the "down" loop.  With P_ab(y) = U_a(y) U_b(y+a) U_a(y+b)^+ U_b(y)^+ for
// a < b and T_ab(y) = P_TA(P_ab(y)), for unitary links
//     up   loop of U_mu(x) in (mu, nu):  s T_ab(x)
//     down loop of U_mu(x) in (mu, nu):  -s V^+ T_ab(x - nu) V,   V = U_nu(x - nu)
// with (a, b) = (min, max)(mu, nu) and s = +1 for mu < nu, -1 otherwise (the loop is P or P^+,
// and the down loop is V^+ P(x - nu)^(+-1) V).  A first pass forms each plaquette once per
// site and stores T (8 su(3) coordinates x 6 planes); the second pass reads 3 links instead
// of 18 per link.

// plane index of (a, b), a < b: (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
__device__ __forceinline__ int plaq_index(int a, int b) { return a == 0 ? b - 1 : (a == 1 ? b + 1 : 5); }

// The plaquette parts of the six planes written for one site: A B^+ with A = U_a(x) U_b(x+a),
// B = U_b(x) U_a(x+b)
template <typename R>
__device__ __forceinline__ void store_plaq_ta(R* __restrict__ T, uint32_t V, int p, uint32_t s, const Mat<R>& A,
                                              const Mat<R>& B) {
    AH<R> h;
    ah_zero(h);
    acc_ah(h, R(1), A, B);
    const R tr = (h.d[0] + h.d[1] + h.d[2]) * R(1.0 / 3.0);
    R* o = T + (size_t)p * 8 * V + s;
    st_stream(o, h.d[0] - tr);
    st_stream(o + (size_t)V, h.d[1] - tr);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        st_stream(o + (size_t)(2 + 2 * k) * V, R(0.5) * h.or_[k]);
        st_stream(o + (size_t)(3 + 2 * k) * V, R(0.5) * h.oi[k]);
    }
}

// T_ab(s) of the six planes of site s into T[(p * 8 + k) * V + s]
template <typename FV>
__device__ __forceinline__ void plaquette_ta_site(const Lat& L, const FV& f, uint32_t s, typename FV::real* __restrict__ T) {
    using R = typename FV::real;
    int c[4];
    site_coords(L, s, c);
    uint32_t sf[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) sf[d] = c[d] == L.n[d] - 1 ? s - (uint32_t)(L.n[d] - 1) * L.st[d] : s + L.st[d];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const Mat<R> Ua = load_link(f, a, s);
#pragma unroll
        for (int b = a + 1; b < 4; ++b) {
            const Mat<R> A = mul(Ua, load_link(f, b, sf[a]));
            const Mat<R> B = mul(load_link(f, b, s), load_link(f, a, sf[b]));
            store_plaq_ta(T, L.V, plaq_index(a, b), s, A, B);
        }
    }
}

template <typename R>
__device__ __forceinline__ Mat<R> load_ta(const R* __restrict__ T, uint32_t V, int p, uint32_t s) {
    const R* q = T + (size_t)p * 8 * V + s;
    R v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = q[(size_t)k * V];
    Mat<R> X;
    X.r[0] = X.r[4] = X.r[8] = R(0);
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

// P_TA(U_mu(x) Sigma_mu(x)) from the stored plaquette parts (links unitary)
template <typename FV>
__device__ __forceinline__ Mat<typename FV::real> ta_force(const Lat& L, const FV& f,
                                                           const typename FV::real* __restrict__ T, uint32_t s,
                                                           const int c[4], int mu) {
    using R = typename FV::real;
    Mat<R> Z;
#pragma unroll
    for (int e = 0; e < 9; ++e) Z.r[e] = Z.i[e] = R(0);
#pragma unroll 1
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        const int p = mu < nu ? plaq_index(mu, nu) : plaq_index(nu, mu);
        const R sg = mu < nu ? R(1) : R(-1);
        int cd[4] = {c[0], c[1], c[2], c[3]};
        const uint32_t sd = step_bwd(L, s, cd, nu);
        const Mat<R> V = load_link(f, nu, sd);
        const Mat<R> D = mul_dn(V, mul(load_ta(T, L.V, p, sd), V));  // V^+ T(x - nu) V
        const Mat<R> Tx = load_ta(T, L.V, p, s);
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            Z.r[e] = fmad(sg, Tx.r[e] - D.r[e], Z.r[e]);
            Z.i[e] = fmad(sg, Tx.i[e] - D.i[e], Z.i[e]);
        }
    }
    return proj_ta(Z);
}

// Force of link (x, mu) at one slice of a walk along t: the in-slice terms as in ta_force,
// the t term from the previous slice's T_{0 mu}(x) and U_0(x), carried by the caller in
// Tprev and U0prev (mu spatial); both are replaced by this slice's values for the next.
template <typename FV>
__device__ __forceinline__ Mat<typename FV::real> ta_force_walk(const Lat& L, const FV& f,
                                                                const typename FV::real* __restrict__ T, uint32_t s,
                                                                const int c[4], int mu, Mat<typename FV::real>& Tprev,
                                                                Mat<typename FV::real>& U0prev) {
    using R = typename FV::real;
    Mat<R> Z;
#pragma unroll
    for (int e = 0; e < 9; ++e) Z.r[e] = Z.i[e] = R(0);
#pragma unroll 1
    for (int nu = 0; nu < 4; ++nu) {
        if (nu == mu) continue;
        const int p = mu < nu ? plaq_index(mu, nu) : plaq_index(nu, mu);
        const R sg = mu < nu ? R(1) : R(-1);
        const Mat<R> Tx = load_ta(T, L.V, p, s);
        Mat<R> D;
        if (nu == 0) {
            D = mul_dn(U0prev, mul(Tprev, U0prev));  // V^+ T(x - t) V from the previous slice
            Tprev = Tx;
            U0prev = load_link(f, 0, s);
        } else {
            int cd[4] = {c[0], c[1], c[2], c[3]};
            const uint32_t sd = step_bwd(L, s, cd, nu);
            const Mat<R> V = load_link(f, nu, sd);
            D = mul_dn(V, mul(load_ta(T, L.V, p, sd), V));
        }
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            Z.r[e] = fmad(sg, Tx.r[e] - D.r[e], Z.r[e]);
            Z.i[e] = fmad(sg, Tx.i[e] - D.i[e], Z.i[e]);
        }
    }
    return proj_ta(Z);
}

// the RK3 update of link (s, mu) given its force Zr = P_TA(U Sigma).  reunit marks the step's
// last stage: the new link is reunitarised, and acc is left unwritten (the next step's first
// stage does not read it)
template <typename FV>
__device__ __forceinline__ void flow_update(const Lat& L, const FV& f, typename FV::real* __restrict__ wout,
                                            typename FV::real* __restrict__ acc, uint32_t s, int mu, double eps, double ca,
                                            double cb, int reunit, const Mat<typename FV::real>& Zr) {
    using R = typename FV::real;
    const Mat<double> Z = to_double(Zr);
    double z[8], a[8];
    su3_to8(Z, z);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        R* p = acc + (size_t)(mu * 8 + k) * L.V + s;
        const double prev = ca != 0.0 ? ca * (double)ld_stream(p) : 0.0;
        a[k] = prev - cb * eps * z[k];
        if (!reunit) st_stream(p, (R)a[k]);
    }
    Mat<double> W = mul(su3_exp(su3_from8(a)), to_double(load_link(f, mu, s)));
    if (reunit) fix_su3(W);
    store_link_stream(wout, L.V, mu, s, from_double<R>(W));
}

// flow_link with the force from the stored plaquette parts
template <typename FV>
__device__ __forceinline__ void flow_link_pf(const Lat& L, const FV& f, const typename FV::real* __restrict__ T,
                                             typename FV::real* __restrict__ wout, typename FV::real* __restrict__ acc,
                                             uint32_t s, const int c[4], int mu, double eps, double ca, double cb,
                                             int reunit) {
    flow_update(L, f, wout, acc, s, mu, eps, ca, cb, reunit, ta_force(L, f, T, s, c, mu));
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
