// flue_common.hpp -- device-side building blocks shared by the FLUE HIP kernels.
//
// Link storage: 36 complex planes, one per direction and colour element, each a
// contiguous array over the sites of (re, im) pairs:
//     real index = (((mu * 3 + a) * 3 + b) * V + s) * 2 + ri
// with the site index s = x + NX * (y + NY * (z + NZ * t)), so x is fastest and every
// time slice is one contiguous block of Vs = NX*NY*NZ sites of each plane.  Directions
// use the reference's numbering shifted to 0: mu = 0 is time, 1..3 are x, y, z
// (src/wloops.F90:21 with mu=1 time).
//
// All matrices live in registers as 9 real + 9 imaginary parts, row-major (a*3+b).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flue {

constexpr int kPlanes = 72;  // 4 directions x 9 elements x {re, im}

struct Lat {
    int n[4];         // extent per direction mu (0 = t, 1 = x, 2 = y, 3 = z)
    uint32_t st[4];   // site-index stride per direction
    uint32_t V;       // number of sites
    uint32_t Vs;      // sites per time slice

    // extent and stride for a run-time direction mu
    __host__ __device__ __forceinline__ int ext(int mu) const {
        return mu == 0 ? n[0] : (mu == 1 ? n[1] : (mu == 2 ? n[2] : n[3]));
    }
    __host__ __device__ __forceinline__ uint32_t stride(int mu) const {
        return mu == 0 ? st[0] : (mu == 1 ? st[1] : (mu == 2 ? st[2] : st[3]));
    }
};

inline Lat make_lat(const int32_t dims_tx_yz[4]) {
    // dims are given in the reference order (NT, NX, NY, NZ)
    Lat L;
    for (int i = 0; i < 4; ++i) L.n[i] = dims_tx_yz[i];
    L.st[1] = 1;
    L.st[2] = (uint32_t)L.n[1];
    L.st[3] = (uint32_t)L.n[1] * L.n[2];
    L.st[0] = (uint32_t)L.n[1] * L.n[2] * L.n[3];
    L.Vs = L.st[0];
    L.V = L.Vs * (uint32_t)L.n[0];
    return L;
}

template <typename R>
struct Mat {
    R r[9], i[9];
};

// ---------------------------------------------------------------- field access

// Read-only view of a field: 36 complex planes (mu, a, b), each V sites of (re, im)
// pairs, real index ((mu*9 + a*3 + b) * V + s) * 2 + ri.  A wave reads one plane of 64
// consecutive sites with one 64-bit (FP32) or 128-bit (FP64) load per lane.  BUF = true
// reads through one buffer resource per direction (9 planes), so every load of a link
// shares one 32-bit vector offset and takes its plane offset from a scalar register; it
// needs 18 * V * sizeof(R) < 2^32 bytes.  BUF = false is the plain global-memory form
// for larger fields.
template <typename R, bool BUF>
struct FieldView {
    using real = R;
    static constexpr bool buffered = BUF;
    const R* __restrict__ p;
    uint32_t V;

    // complex plane (mu, e) with e = a*3 + b
    __device__ __forceinline__ const R* cplane(int mu, int e) const { return p + (size_t)(mu * 9 + e) * 2 * V; }
    __device__ __forceinline__ __amdgpu_buffer_rsrc_t dir_rsrc(int mu) const {
        const char* b = (const char*)p + (size_t)mu * 18u * V * sizeof(R);
        // num_records is the unsigned 32-bit byte count of the 9 planes (fits_buffer keeps it
        // below 2^32); the builtin takes it as an int, so its bit pattern is passed unchanged.
        // dword3 0x00020000: 32-bit data format, no swizzle -- the raw-buffer setting for gfx9 parts
        const uint32_t nrec = 18u * V * (uint32_t)sizeof(R);
        return __builtin_amdgcn_make_buffer_rsrc((void*)b, (short)0, __builtin_bit_cast(int, nrec), 0x00020000);
    }
    // one (re, im) pair: a 64-bit (FP32) or 128-bit (FP64) load, reinterpreted as a whole
    // (taking dwords out of the loaded vector one by one is not used: with bit casts of
    // single elements the compiler narrowed such loads to one dword)
    __device__ __forceinline__ void bld2(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, R& re, R& im) const {
        typedef R r2 __attribute__((ext_vector_type(2)));
        r2 v;
        if constexpr (sizeof(R) == 4)
            v = __builtin_bit_cast(r2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
        else
            v = __builtin_bit_cast(r2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
        re = v.x;
        im = v.y;
    }
};

template <typename R>
struct Pair2;
template <>
struct Pair2<float> {
    using type = float2;
};
template <>
struct Pair2<double> {
    using type = double2;
};

template <typename R, bool BUF>
__device__ __forceinline__ Mat<R> load_link(const FieldView<R, BUF>& f, int mu, uint32_t s) {
    Mat<R> m;
    if constexpr (BUF) {
        const __amdgpu_buffer_rsrc_t r = f.dir_rsrc(mu);
        const uint32_t vo = s * 2u * (uint32_t)sizeof(R), pb = f.V * 2u * (uint32_t)sizeof(R);
#pragma unroll
        for (int e = 0; e < 9; ++e) f.bld2(r, vo, e * pb, m.r[e], m.i[e]);
    } else {
        using P = typename Pair2<R>::type;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            const P v = reinterpret_cast<const P*>(f.cplane(mu, e))[s];
            m.r[e] = v.x;
            m.i[e] = v.y;
        }
    }
    return m;
}

// U^dagger loaded directly: element (a,b) of U^+ is conj(U_ba)
template <typename R, bool BUF>
__device__ __forceinline__ Mat<R> load_link_dag(const FieldView<R, BUF>& f, int mu, uint32_t s) {
    const Mat<R> u = load_link(f, mu, s);
    Mat<R> m;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            m.r[a * 3 + b] = u.r[b * 3 + a];
            m.i[a * 3 + b] = -u.i[b * 3 + a];
        }
    return m;
}

// host side: can the field be read through per-direction buffer resources?
inline bool fits_buffer(uint32_t V, int is_f32) {
    return (uint64_t)18 * V * (is_f32 ? 4 : 8) < (1ull << 32);
}

// host side: call fn with the FieldView type matching the precision and size
template <typename Fn>
inline void with_view(uint32_t V, const void* U, int is_f32, Fn&& fn) {
    if (is_f32) {
        if (fits_buffer(V, 1))
            fn(FieldView<float, true>{(const float*)U, V});
        else
            fn(FieldView<float, false>{(const float*)U, V});
    } else {
        if (fits_buffer(V, 0))
            fn(FieldView<double, true>{(const double*)U, V});
        else
            fn(FieldView<double, false>{(const double*)U, V});
    }
}

template <typename R>
__device__ __forceinline__ void store_link(R* __restrict__ p, uint32_t V, int mu, uint32_t s, const Mat<R>& m) {
    using P = typename Pair2<R>::type;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        P v;
        v.x = m.r[e];
        v.y = m.i[e];
        reinterpret_cast<P*>(p + (size_t)(mu * 9 + e) * 2 * V)[s] = v;
    }
}

// Streaming accesses: data read or written once per kernel (accumulators, output links,
// stored plaquette parts) take the non-temporal cache policy, so they do not push out of L2
// the links and plaquette parts that neighbouring sites read again.
#ifndef FLUE_NT
#define FLUE_NT 1
#endif
template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
#if FLUE_NT
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
#if FLUE_NT
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

// store_link with streaming stores
template <typename R>
__device__ __forceinline__ void store_link_stream(R* __restrict__ p, uint32_t V, int mu, uint32_t s, const Mat<R>& m) {
    typedef R r2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        r2 v;
        v.x = m.r[e];
        v.y = m.i[e];
        st_stream(reinterpret_cast<r2*>(p + (size_t)(mu * 9 + e) * 2 * V) + s, v);
    }
}

// ---------------------------------------------------------------- SU(3) algebra
//
// Complex multiply-accumulates are written as explicit FMA chains: an expression such as
// acc += a*b - c*d would otherwise become a multiply, an FMA and an add.

template <typename R>
__device__ __forceinline__ R fmad(R a, R b, R c) {
    if constexpr (sizeof(R) == 4)
        return __builtin_fmaf(a, b, c);
    else
        return __builtin_fma(a, b, c);
}

// (cr, ci) += (ar + i ai) (br + i bi)
template <typename R>
__device__ __forceinline__ void cmac(R& cr, R& ci, R ar, R ai, R br, R bi) {
    cr = fmad(ar, br, cr);
    cr = fmad(-ai, bi, cr);
    ci = fmad(ar, bi, ci);
    ci = fmad(ai, br, ci);
}

// (cr, ci) += (ar + i ai) conj(br + i bi)
template <typename R>
__device__ __forceinline__ void cmac_c(R& cr, R& ci, R ar, R ai, R br, R bi) {
    cr = fmad(ar, br, cr);
    cr = fmad(ai, bi, cr);
    ci = fmad(ai, br, ci);
    ci = fmad(-ar, bi, ci);
}

template <typename R>
__device__ __forceinline__ Mat<R> mul(const Mat<R>& A, const Mat<R>& B) {
    Mat<R> C;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            R cr = A.r[a * 3] * B.r[b];
            R ci = A.r[a * 3] * B.i[b];
            cr = fmad(-A.i[a * 3], B.i[b], cr);
            ci = fmad(A.i[a * 3], B.r[b], ci);
#pragma unroll
            for (int k = 1; k < 3; ++k) cmac(cr, ci, A.r[a * 3 + k], A.i[a * 3 + k], B.r[k * 3 + b], B.i[k * 3 + b]);
            C.r[a * 3 + b] = cr;
            C.i[a * 3 + b] = ci;
        }
    return C;
}

// C = A * B^dagger
template <typename R>
__device__ __forceinline__ Mat<R> mul_nd(const Mat<R>& A, const Mat<R>& B) {
    Mat<R> C;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            R cr = A.r[a * 3] * B.r[b * 3];
            R ci = A.i[a * 3] * B.r[b * 3];
            cr = fmad(A.i[a * 3], B.i[b * 3], cr);
            ci = fmad(-A.r[a * 3], B.i[b * 3], ci);
#pragma unroll
            for (int k = 1; k < 3; ++k)
                cmac_c(cr, ci, A.r[a * 3 + k], A.i[a * 3 + k], B.r[b * 3 + k], B.i[b * 3 + k]);
            C.r[a * 3 + b] = cr;
            C.i[a * 3 + b] = ci;
        }
    return C;
}

// Rows 0 and 1 of A * B and of A * B^dagger (row 2 of the result is left zero).  For
// SU(3) matrices row 2 follows from the other two (fill_row2), so a product chain of
// SU(3) links can carry two rows: 72 instead of 108 multiply-adds per step.
template <typename R>
__device__ __forceinline__ Mat<R> mul2(const Mat<R>& A, const Mat<R>& B) {
    Mat<R> C;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            R cr = A.r[a * 3] * B.r[b];
            R ci = A.r[a * 3] * B.i[b];
            cr = fmad(-A.i[a * 3], B.i[b], cr);
            ci = fmad(A.i[a * 3], B.r[b], ci);
#pragma unroll
            for (int k = 1; k < 3; ++k) cmac(cr, ci, A.r[a * 3 + k], A.i[a * 3 + k], B.r[k * 3 + b], B.i[k * 3 + b]);
            C.r[a * 3 + b] = cr;
            C.i[a * 3 + b] = ci;
        }
#pragma unroll
    for (int b = 6; b < 9; ++b) C.r[b] = C.i[b] = R(0);
    return C;
}

template <typename R>
__device__ __forceinline__ Mat<R> mul2_nd(const Mat<R>& A, const Mat<R>& B) {
    Mat<R> C;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            R cr = A.r[a * 3] * B.r[b * 3];
            R ci = A.i[a * 3] * B.r[b * 3];
            cr = fmad(A.i[a * 3], B.i[b * 3], cr);
            ci = fmad(-A.r[a * 3], B.i[b * 3], ci);
#pragma unroll
            for (int k = 1; k < 3; ++k)
                cmac_c(cr, ci, A.r[a * 3 + k], A.i[a * 3 + k], B.r[b * 3 + k], B.i[b * 3 + k]);
            C.r[a * 3 + b] = cr;
            C.i[a * 3 + b] = ci;
        }
#pragma unroll
    for (int b = 6; b < 9; ++b) C.r[b] = C.i[b] = R(0);
    return C;
}

// row 2 = conj(row 0 x row 1), exact for SU(3) (unitary, det 1)
template <typename R>
__device__ __forceinline__ void fill_row2(Mat<R>& M) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int j = (c + 1) % 3, k = (c + 2) % 3;
        // x = r0_j r1_k - r0_k r1_j; row2_c = conj(x)
        R xr = M.r[j] * M.r[3 + k];
        R xi = M.r[j] * M.i[3 + k];
        xr = fmad(-M.i[j], M.i[3 + k], xr);
        xi = fmad(M.i[j], M.r[3 + k], xi);
        xr = fmad(-M.r[k], M.r[3 + j], xr);
        xr = fmad(M.i[k], M.i[3 + j], xr);
        xi = fmad(-M.r[k], M.i[3 + j], xi);
        xi = fmad(-M.i[k], M.r[3 + j], xi);
        M.r[6 + c] = xr;
        M.i[6 + c] = -xi;
    }
}

// C = A^dagger * B
template <typename R>
__device__ __forceinline__ Mat<R> mul_dn(const Mat<R>& A, const Mat<R>& B) {
    Mat<R> C;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            // sum_k conj(A_ka) B_kb
            R cr = A.r[a] * B.r[b];
            R ci = A.r[a] * B.i[b];
            cr = fmad(A.i[a], B.i[b], cr);
            ci = fmad(-A.i[a], B.r[b], ci);
#pragma unroll
            for (int k = 1; k < 3; ++k)
                cmac_c(cr, ci, B.r[k * 3 + b], B.i[k * 3 + b], A.r[k * 3 + a], A.i[k * 3 + a]);
            C.r[a * 3 + b] = cr;
            C.i[a * 3 + b] = ci;
        }
    return C;
}

template <typename R>
__device__ __forceinline__ Mat<R> dag(const Mat<R>& A) {
    Mat<R> C;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            C.r[a * 3 + b] = A.r[b * 3 + a];
            C.i[a * 3 + b] = -A.i[b * 3 + a];
        }
    return C;
}

template <typename R>
__device__ __forceinline__ Mat<R> identity() {
    Mat<R> C;
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        C.r[e] = (e % 4 == 0) ? R(1) : R(0);
        C.i[e] = R(0);
    }
    return C;
}

// Re Tr(A B^dagger) = sum_ab Re(A_ab conj(B_ab))
template <typename R>
__device__ __forceinline__ R retr_nd(const Mat<R>& A, const Mat<R>& B) {
    R s = A.r[0] * B.r[0];
    s = fmad(A.i[0], B.i[0], s);
#pragma unroll
    for (int e = 1; e < 9; ++e) {
        s = fmad(A.r[e], B.r[e], s);
        s = fmad(A.i[e], B.i[e], s);
    }
    return s;
}

// Re Tr(A B)
template <typename R>
__device__ __forceinline__ R retr_mul(const Mat<R>& A, const Mat<R>& B) {
    R s = R(0);
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            s = fmad(A.r[a * 3 + k], B.r[k * 3 + a], s);
            s = fmad(-A.i[a * 3 + k], B.i[k * 3 + a], s);
        }
    return s;
}

template <typename R>
__device__ __forceinline__ R retr(const Mat<R>& A) {
    return A.r[0] + A.r[4] + A.r[8];
}

// ---------------------------------------------------------------- periodic coordinates

// (c + k) mod n for |k| <= 3 and c in [0, n), branch-free (selects only) so that
// callers stay a single basic block
__device__ __forceinline__ int wrapc(int v, int n) {
#pragma unroll
    for (int i = 0; i < 4; ++i) v = (v < 0) ? v + n : v;
#pragma unroll
    for (int i = 0; i < 4; ++i) v = (v >= n) ? v - n : v;
    return v;
}

__device__ __forceinline__ void site_coords(const Lat& L, uint32_t s, int c[4]) {
    uint32_t q = s;
    c[1] = (int)(q % (uint32_t)L.n[1]);
    q /= (uint32_t)L.n[1];
    c[2] = (int)(q % (uint32_t)L.n[2]);
    q /= (uint32_t)L.n[2];
    c[3] = (int)(q % (uint32_t)L.n[3]);
    c[0] = (int)(q / (uint32_t)L.n[3]);
}

// site reached from s (coords c) by one step along +mu or -mu
__device__ __forceinline__ uint32_t step_fwd(const Lat& L, uint32_t s, int* c, int mu) {
    if (c[mu] == L.ext(mu) - 1) {
        c[mu] = 0;
        return s - (uint32_t)(L.ext(mu) - 1) * L.stride(mu);
    }
    c[mu] += 1;
    return s + L.stride(mu);
}
__device__ __forceinline__ uint32_t step_bwd(const Lat& L, uint32_t s, int* c, int mu) {
    if (c[mu] == 0) {
        c[mu] = L.ext(mu) - 1;
        return s + (uint32_t)(L.ext(mu) - 1) * L.stride(mu);
    }
    c[mu] -= 1;
    return s - L.stride(mu);
}

// Offset within a time slice of its idx-th site when the slice is taken x fastest, then
// direction d, then the remaining spatial direction.  For d = 1 or 2 that is storage order;
// for d = 3 (z) consecutive sites then span a few rows along z, so kernels whose sites reach
// several steps along z find those sites in the same block or the next one.
__device__ __forceinline__ uint32_t slice_site_along(const Lat& L, uint32_t idx, int d) {
    if (d != 3) return idx;
    const uint32_t NX = (uint32_t)L.n[1], NZ = (uint32_t)L.n[3];
    const uint32_t x = idx % NX, r = idx / NX;
    const uint32_t z = r % NZ, y = r / NZ;
    return x + NX * (y + (uint32_t)L.n[2] * z);
}

// Hardware hands consecutive workgroups to the 8 XCDs in turn; give each XCD a
// contiguous range of logical blocks so that its private L2 sees neighbouring sites.
__device__ __forceinline__ uint32_t xcd_block(uint32_t bid, uint32_t nblocks) {
    const uint32_t X = 8;
    uint32_t q = nblocks / X, r = nblocks % X;
    uint32_t x = bid % X, k = bid / X;
    return x * q + (x < r ? x : r) + k;
}

// ---------------------------------------------------------------- compensated sums

struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

__device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    double lo = s.lo + a.lo + b.lo;
    double hi = s.hi + lo;
    return {hi, lo - (hi - s.hi)};
}

__device__ __forceinline__ DD dd_shfl_xor(DD v, int m) {
    return {__shfl_xor(v.hi, m), __shfl_xor(v.lo, m)};
}

// Deterministic block sum of a compensated pair: fixed butterfly inside each
// wavefront, then the waves in index order.  Result valid in thread 0.
template <int BS>
__device__ __forceinline__ DD block_sum(DD v) {
    __shared__ double sh_hi[BS / 64], sh_lo[BS / 64];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = dd_add(v, dd_shfl_xor(v, m));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh_hi[w] = v.hi;
        sh_lo[w] = v.lo;
    }
    __syncthreads();
    DD r = {0.0, 0.0};
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < BS / 64; ++k) r = dd_add(r, DD{sh_hi[k], sh_lo[k]});
    }
    return r;
}

}  // namespace flue
