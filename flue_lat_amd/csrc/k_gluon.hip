// k_gluon.hip -- the scalar gluon propagator of FLUE's gluonCalc module.
//
// mom_scalar_kernel: calc_mom_space_scalarD (src/gluonCalc/gluonProp.f90:112-250) on a
// Fourier-transformed field Uq held like a gauge field (36 complex planes over the momenta).
// One thread per output momentum q reads U(q) and U(-q) of each direction and forms
//   A+ = (i/2) e^{-i th/2} TCS(U(q), U(-q)),  A- = (i/2) e^{+i th/2} TCS(U(-q), U(q)),
//   D(q) = prefac * sum_mu sum_a Re(c_a(A+) c_a(A-)),   (times xi^2 for mu = 0)
// with TCS(L, R) = L - R^+ - Tr(L - R^+)/3 on all nine entries (src/SU3MatrixOps.f90:94-105)
// and c_a the Gell-Mann components of src/SU3MatrixOps.f90:107-122.
//
// scalar_gluon_kernel: scalarGluonProp (src/gluonCalc/gluonProp.f90:17-110), one thread per
// entry of D(NS,NS,NS,NS).
#include "flue_common.hpp"
#include "flue_launch.hpp"

namespace flue {

namespace {

constexpr int kBS = 256;

struct C2d {
    double re, im;
};

__device__ __forceinline__ C2d cm(C2d a, C2d b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// 3x3 complex in double, row-major (a*3+b)
struct M3d {
    C2d e[9];
};

template <typename R>
__device__ __forceinline__ M3d to_d(const Mat<R>& m) {
    M3d o;
#pragma unroll
    for (int i = 0; i < 9; ++i) o.e[i] = {(double)m.r[i], (double)m.i[i]};
    return o;
}

// L - R^+ minus a third of its trace on every entry
__device__ __forceinline__ M3d tcs(const M3d& L, const M3d& R) {
    M3d o;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const C2d l = L.e[a * 3 + b], r = R.e[b * 3 + a];
            o.e[a * 3 + b] = {l.re - r.re, l.im + r.im};
        }
    const double tr = (o.e[0].re + o.e[4].re) + o.e[8].re, ti = (o.e[0].im + o.e[4].im) + o.e[8].im;
    const double t3r = tr / 3.0, t3i = ti / 3.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) o.e[i] = {o.e[i].re - t3r, o.e[i].im - t3i};
    return o;
}

// Gell-Mann components: c[2] is complex, the others real
struct Colour {
    double c[8];
    double c3i;
};

__device__ __forceinline__ Colour colour_decomp(const M3d& A) {
    Colour k;
    k.c[0] = A.e[1].re + A.e[3].re;
    k.c[1] = -A.e[1].im + A.e[3].im;
    k.c[2] = A.e[0].re - A.e[4].re;
    k.c3i = A.e[0].im - A.e[4].im;
    k.c[3] = A.e[2].re + A.e[6].re;
    k.c[4] = -A.e[2].im + A.e[6].im;
    k.c[5] = A.e[5].re + A.e[7].re;
    k.c[6] = -A.e[5].im + A.e[7].im;
    k.c[7] = A.e[0].re + A.e[4].re - (2.0 / sqrt(3.0)) * A.e[8].re;
    return k;
}

__device__ __forceinline__ M3d scale(const M3d& A, C2d p) {
    M3d o;
#pragma unroll
    for (int i = 0; i < 9; ++i) o.e[i] = cm(p, A.e[i]);
    return o;
}

// Lq: lattice of the transformed field (its extents E); output momenta 0 <= q < dext,
// written as complex128 at D + sum q_i * dstr_i (complex elements, order t, x, y, z)
template <typename FV>
__global__ void __launch_bounds__(kBS) mom_scalar_kernel(Lat Lq, FV f, int mu_start, double xi, int4 dext,
                                                         longlong4 dstr, uint32_t nD, double* __restrict__ D) {
    const uint32_t o = blockIdx.x * kBS + threadIdx.x;
    if (o >= nD) return;
    int q[4];
    uint32_t r = o;
    q[0] = (int)(r % (uint32_t)dext.x);
    r /= (uint32_t)dext.x;
    q[1] = (int)(r % (uint32_t)dext.y);
    r /= (uint32_t)dext.y;
    q[2] = (int)(r % (uint32_t)dext.z);
    q[3] = (int)(r / (uint32_t)dext.z);
    double qhat[4];
    uint32_t sp = 0, sn = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int E = Lq.ext(i);
        qhat[i] = (2.0 * M_PI / 1.0) * (double)q[i] / (double)E;
        const int qn = q[i] == 0 ? 0 : E - q[i];
        sp += (uint32_t)q[i] * Lq.stride(i);
        sn += (uint32_t)qn * Lq.stride(i);
    }
    double acc = 0.0;
    for (int mu = mu_start; mu <= 3; ++mu) {
        const M3d Up = to_d(load_link(f, mu, sp)), Un = to_d(load_link(f, mu, sn));
        // the phase of direction mu uses qhat(mu) of a 1-based qhat: the component of
        // direction mu - 1 (the t component for mu = 0)
        const double th = qhat[mu >= 1 ? mu - 1 : 0];
        double sh, ch;
        sincos(0.5 * th, &sh, &ch);
        // (i/2) e^{-i th/2} and (i/2) e^{+i th/2}
        const C2d pp = {0.5 * sh, 0.5 * ch}, pn = {-0.5 * sh, 0.5 * ch};
        const Colour cp = colour_decomp(scale(tcs(Up, Un), pp));
        const Colour cn = colour_decomp(scale(tcs(Un, Up), pn));
        const double w = mu == 0 ? xi * xi : 1.0;
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < 8; ++a)
            s += a == 2 ? (w * cp.c[a]) * cn.c[a] - (w * cp.c3i) * cn.c3i : (w * cp.c[a]) * cn.c[a];
        acc += s;
    }
    const double Vd = (double)Lq.ext(0) * (double)Lq.ext(1) * (double)Lq.ext(2) * (double)Lq.ext(3);
    // src/gluonCalc/gluonProp.f90:216,225 (qx counted twice, qz not at all for mu_start = 1)
    const bool four = (mu_start == 1 && q[1] + q[2] + q[1] < 1) ||
                      (mu_start == 0 && q[0] + q[1] + q[2] + q[3] + q[1] < 1);
    const double prefac = 2.0 / (8.0 * Vd * (four ? 4.0 : 3.0));
    const int64_t off = q[0] * dstr.x + q[1] * dstr.y + q[2] * dstr.z + q[3] * dstr.w;
    D[2 * off] = acc * prefac;
    D[2 * off + 1] = 0.0;
}

// D(t,x,y,z) (Fortran order over NS^4) of scalarGluonProp; entries with t >= NT are 0.
// The reference sums V identical terms exp(-i c q) U one by one (:69-73); this takes V
// times the term.
template <typename FV>
__global__ void __launch_bounds__(kBS) scalar_gluon_kernel(Lat L, FV f, int NS, uint32_t nD,
                                                           double* __restrict__ D) {
    const uint32_t o = blockIdx.x * kBS + threadIdx.x;
    if (o >= nD) return;
    int c[4];
    uint32_t r = o;
    c[0] = (int)(r % (uint32_t)NS);
    r /= (uint32_t)NS;
    c[1] = (int)(r % (uint32_t)NS);
    r /= (uint32_t)NS;
    c[2] = (int)(r % (uint32_t)NS);
    c[3] = (int)(r / (uint32_t)NS);
    if (c[0] >= L.n[0]) {
        D[o] = 0.0;
        return;
    }
    const uint32_t s = (uint32_t)c[1] * L.st[1] + (uint32_t)c[2] * L.st[2] + (uint32_t)c[3] * L.st[3] +
                       (uint32_t)c[0] * L.st[0];
    const double Vd = (double)L.V;
    double acc = 0.0;
    for (int mu = 0; mu < 4; ++mu) {
        // 1-based coordinate and qhat = 2 pi coord / N of direction mu
        const double cm1 = (double)(c[mu] + 1);
        const double qh = (2.0 * M_PI / 1.0) * cm1 / (double)L.ext(mu);
        double sn, cs;
        sincos(cm1 * qh, &sn, &cs);
        const M3d U = to_d(load_link(f, mu, s));
        M3d T1, T2;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const C2d a = cm(C2d{cs, -sn}, U.e[i]), b = cm(C2d{cs, sn}, U.e[i]);
            T1.e[i] = {Vd * a.re, Vd * a.im};
            T2.e[i] = {Vd * b.re, Vd * b.im};
        }
        double hs, hc;
        sincos(qh * 0.5, &hs, &hc);
        const M3d A = scale(scale(tcs(T2, T1), C2d{0.5 * hs, 0.5 * hc}), C2d{-0.5 * hs, 0.5 * hc});
        const Colour k = colour_decomp(A);
        double t = 0.0;
#pragma unroll
        for (int a = 0; a < 8; ++a) t += a == 2 ? k.c[a] * k.c[a] - k.c3i * k.c3i : k.c[a] * k.c[a];
        acc += t;
    }
    D[o] = ((1.0 / 3.0) * (1.0 / 8.0) * Vd) * acc;
}

}  // namespace

hipError_t launch_mom_scalar(const Lat& Lq, const void* Uq, int is_f32, int mu_start, double xi,
                             const int32_t dext[4], const int64_t dstr[4], double* D, hipStream_t st) {
    const uint64_t nD64 = (uint64_t)dext[0] * dext[1] * dext[2] * dext[3];
    if (nD64 > 0x7FFFFFFFull) return hipErrorInvalidValue;
    const uint32_t nD = (uint32_t)nD64;
    const uint32_t g = (nD + kBS - 1) / kBS;
    const int4 de = make_int4(dext[0], dext[1], dext[2], dext[3]);
    const longlong4 ds = make_longlong4(dstr[0], dstr[1], dstr[2], dstr[3]);
    with_view(Lq.V, Uq, is_f32, [&](auto f) {
        mom_scalar_kernel<decltype(f)><<<g, kBS, 0, st>>>(Lq, f, mu_start, xi, de, ds, nD, D);
    });
    return hipGetLastError();
}

hipError_t launch_scalar_gluon(const Lat& L, const void* U, int is_f32, int NS, double* D, hipStream_t st) {
    // D holds NS^4 entries indexed in 32 bits: refuse anything the grid cannot cover
    const uint64_t nD64 = (uint64_t)NS * NS * NS * NS;
    if (NS < 1 || nD64 > 0x7FFFFFFFull) return hipErrorInvalidValue;
    const uint32_t nD = (uint32_t)nD64;
    const uint32_t g = (nD + kBS - 1) / kBS;
    with_view(L.V, U, is_f32, [&](auto f) {
        scalar_gluon_kernel<decltype(f)><<<g, kBS, 0, st>>>(L, f, NS, nD, D);
    });
    return hipGetLastError();
}

}  // namespace flue
