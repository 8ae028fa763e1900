// flue_su3x.hpp -- FP64 su(3)/SU(3) utilities: exponential map, reunitarisation,
// traceless anti-Hermitian projection and the counter-based link generator.
#pragma once
#include "flue_common.hpp"

namespace flue {

// exp(X) by scaling and squaring a Taylor polynomial of degree <= 16 (Horner form).
__device__ __forceinline__ Mat<double> su3_exp(const Mat<double>& X) {
    double nrm = 0.0;
#pragma unroll
    for (int e = 0; e < 9; ++e) nrm += X.r[e] * X.r[e] + X.i[e] * X.i[e];
    nrm = sqrt(nrm);
    int s = 0;
    while (nrm > 0.125 && s < 60) {
        nrm *= 0.5;
        ++s;
    }
    const double scale = ldexp(1.0, -s);
    Mat<double> Y;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        Y.r[e] = X.r[e] * scale;
        Y.i[e] = X.i[e] * scale;
    }
    // Taylor degree: the smallest K <= 16 whose first dropped term, |Y|^(K+1)/(K+1)!, is
    // below 1e-17 (|Y| <= 0.125 after the scaling, so K <= 11; small flow steps need ~8)
    int K = 1;
    double term = 0.5 * nrm * nrm;  // |Y|^2 / 2!, the first term a degree-1 polynomial drops
    while (K < 16 && term > 1e-17) {
        ++K;
        term *= nrm / (double)(K + 1);
    }
    Mat<double> E = identity<double>();
    for (int k = K; k >= 1; --k) {
        Mat<double> T = mul(Y, E);
        const double inv = 1.0 / (double)k;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            E.r[e] = T.r[e] * inv + ((e % 4 == 0) ? 1.0 : 0.0);
            E.i[e] = T.i[e] * inv;
        }
    }
    for (int k = 0; k < s; ++k) E = mul(E, E);
    return E;
}

// Rows 1, 2 by Gram-Schmidt, row 3 = conj(row1 x row2) -- the FixSU3Matrix scheme
// (src/gaugeIO/ILDG_bin.f90:138-151).
__device__ __forceinline__ void fix_su3(Mat<double>& U) {
    double n1 = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) n1 += U.r[j] * U.r[j] + U.i[j] * U.i[j];
    n1 = 1.0 / sqrt(n1);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        U.r[j] *= n1;
        U.i[j] *= n1;
    }
    // <v1, v2> = sum conj(v1) v2
    double dr = 0.0, di = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        dr += U.r[j] * U.r[3 + j] + U.i[j] * U.i[3 + j];
        di += U.r[j] * U.i[3 + j] - U.i[j] * U.r[3 + j];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double vr = U.r[j] * dr - U.i[j] * di;
        double vi = U.r[j] * di + U.i[j] * dr;
        U.r[3 + j] -= vr;
        U.i[3 + j] -= vi;
    }
    double n2 = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) n2 += U.r[3 + j] * U.r[3 + j] + U.i[3 + j] * U.i[3 + j];
    n2 = 1.0 / sqrt(n2);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        U.r[3 + j] *= n2;
        U.i[3 + j] *= n2;
    }
    double n3 = 0.0;
#pragma unroll
    for (int ic = 0; ic < 3; ++ic) {
        const int jc = (ic + 1) % 3, kc = (jc + 1) % 3;
        // conj(v1_j v2_k - v1_k v2_j)
        double pr = U.r[jc] * U.r[3 + kc] - U.i[jc] * U.i[3 + kc] - (U.r[kc] * U.r[3 + jc] - U.i[kc] * U.i[3 + jc]);
        double pi = U.r[jc] * U.i[3 + kc] + U.i[jc] * U.r[3 + kc] - (U.r[kc] * U.i[3 + jc] + U.i[kc] * U.r[3 + jc]);
        U.r[6 + ic] = pr;
        U.i[6 + ic] = -pi;
        n3 += pr * pr + pi * pi;
    }
    n3 = 1.0 / sqrt(n3);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        U.r[6 + j] *= n3;
        U.i[6 + j] *= n3;
    }
}

// P_TA(M) = (M - M^+)/2 - Tr(M - M^+)/6
template <typename R>
__device__ __forceinline__ Mat<R> proj_ta(const Mat<R>& M) {
    Mat<R> T;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            T.r[a * 3 + b] = R(0.5) * (M.r[a * 3 + b] - M.r[b * 3 + a]);
            T.i[a * 3 + b] = R(0.5) * (M.i[a * 3 + b] + M.i[b * 3 + a]);
        }
    const R tr = (T.i[0] + T.i[4] + T.i[8]) / R(3);
    T.i[0] -= tr;
    T.i[4] -= tr;
    T.i[8] -= tr;
    return T;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Link generator: 8 N(0,1) coefficients keyed by (seed, link id) through Philox and
// Box-Muller, H = sum_a c_a lambda_a / 2, U = fix(exp(i eps H)).  The link id is the
// reference-order offset t + NT(x + NX(y + NY z)) + V mu, so a field does not depend
// on the order the links are generated in.
__device__ __forceinline__ Mat<double> random_link(uint64_t seed, uint64_t link, double eps) {
    double g[8];
#pragma unroll
    for (int blk = 0; blk < 2; ++blk) {
        uint32_t c[4] = {(uint32_t)link, (uint32_t)(link >> 32), (uint32_t)blk, 0x464C5545u};
        philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const double u1 = ((double)c[2 * p] + 0.5) * 0x1p-32;
            const double u2 = ((double)c[2 * p + 1] + 0.5) * 0x1p-32;
            const double rad = sqrt(-2.0 * log(u1));
            const double ang = 6.283185307179586476925286766559 * u2;
            g[4 * blk + 2 * p] = rad * cos(ang);
            g[4 * blk + 2 * p + 1] = rad * sin(ang);
        }
    }
    // X = i eps/2 sum_a g_a lambda_a, written out from the Gell-Mann matrices
    const double h = 0.5 * eps;
    const double r3 = 0.57735026918962576450914878050196;  // 1/sqrt(3)
    Mat<double> X;
    // i * lambda_3, lambda_8 diagonal parts
    X.r[0] = 0.0;
    X.i[0] = h * (g[2] + r3 * g[7]);
    X.r[4] = 0.0;
    X.i[4] = h * (-g[2] + r3 * g[7]);
    X.r[8] = 0.0;
    X.i[8] = h * (-2.0 * r3 * g[7]);
    // (0,1): lambda_1 -> 1, lambda_2 -> -i ; times i: i*g1 + g2
    X.r[1] = h * g[1];
    X.i[1] = h * g[0];
    X.r[3] = -h * g[1];
    X.i[3] = h * g[0];
    // (0,2): lambda_4, lambda_5
    X.r[2] = h * g[4];
    X.i[2] = h * g[3];
    X.r[6] = -h * g[4];
    X.i[6] = h * g[3];
    // (1,2): lambda_6, lambda_7
    X.r[5] = h * g[6];
    X.i[5] = h * g[5];
    X.r[7] = -h * g[6];
    X.i[7] = h * g[5];
    Mat<double> U = su3_exp(X);
    fix_su3(U);
    return U;
}

}  // namespace flue
