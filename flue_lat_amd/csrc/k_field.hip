// k_field.hip -- moving gauge fields between the reference's complex layout and the
// SoA planes, precision conversion, and the on-device link generator.
#include "flue_launch.hpp"
#include "flue_su3x.hpp"

namespace flue {

namespace {

template <typename Rin, typename Rout>
__global__ void __launch_bounds__(256) import_kernel(Lat L, const Rin* __restrict__ src, int64_t st_t, int64_t st_x,
                                                     int64_t st_y, int64_t st_z, int64_t st_mu, int64_t st_a,
                                                     int64_t st_b, Rout* __restrict__ dst) {
    // one thread per (site, mu); threads run along the reference's t-fastest order so
    // that reads of a Fortran-ordered source are contiguous
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (uint64_t)L.V * 4) return;
    const uint32_t mu = (uint32_t)(gid / L.V);
    uint32_t q = (uint32_t)(gid - (uint64_t)mu * L.V);
    const int t = (int)(q % (uint32_t)L.n[0]);
    q /= (uint32_t)L.n[0];
    const int x = (int)(q % (uint32_t)L.n[1]);
    q /= (uint32_t)L.n[1];
    const int y = (int)(q % (uint32_t)L.n[2]);
    const int z = (int)(q / (uint32_t)L.n[2]);
    const int64_t base = t * st_t + x * st_x + y * st_y + z * st_z + (int64_t)mu * st_mu;
    const uint32_t s = x + L.st[2] * y + L.st[3] * z + L.st[0] * t;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int64_t off = 2 * (base + a * st_a + b * st_b);
            Rout* d = dst + ((size_t)(mu * 9 + a * 3 + b) * L.V + s) * 2;
            d[0] = (Rout)src[off];
            d[1] = (Rout)src[off + 1];
        }
}

template <typename Rin, typename Rout>
__global__ void __launch_bounds__(256) export_kernel(Lat L, const Rin* __restrict__ src, Rout* __restrict__ dst,
                                                     int64_t st_t, int64_t st_x, int64_t st_y, int64_t st_z,
                                                     int64_t st_mu, int64_t st_a, int64_t st_b) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (uint64_t)L.V * 4) return;
    const uint32_t mu = (uint32_t)(gid / L.V);
    uint32_t q = (uint32_t)(gid - (uint64_t)mu * L.V);
    const int t = (int)(q % (uint32_t)L.n[0]);
    q /= (uint32_t)L.n[0];
    const int x = (int)(q % (uint32_t)L.n[1]);
    q /= (uint32_t)L.n[1];
    const int y = (int)(q % (uint32_t)L.n[2]);
    const int z = (int)(q / (uint32_t)L.n[2]);
    const int64_t base = t * st_t + x * st_x + y * st_y + z * st_z + (int64_t)mu * st_mu;
    const uint32_t s = x + L.st[2] * y + L.st[3] * z + L.st[0] * t;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int64_t off = 2 * (base + a * st_a + b * st_b);
            const Rin* q = src + ((size_t)(mu * 9 + a * 3 + b) * L.V + s) * 2;
            dst[off] = (Rout)q[0];
            dst[off + 1] = (Rout)q[1];
        }
}

template <typename R>
__global__ void __launch_bounds__(256) random_kernel(Lat L, uint64_t seed, double eps, R* __restrict__ dst) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (uint64_t)L.V * 4) return;
    const uint32_t mu = (uint32_t)(gid / L.V);
    const uint32_t s = (uint32_t)(gid - (uint64_t)mu * L.V);
    int c[4];
    site_coords(L, s, c);
    const uint64_t sref = (uint64_t)c[0] + (uint64_t)L.n[0] * (c[1] + (uint64_t)L.n[1] * (c[2] + (uint64_t)L.n[2] * c[3]));
    const Mat<double> U = random_link(seed, sref + (uint64_t)L.V * mu, eps);
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        R* d = dst + ((size_t)(mu * 9 + e) * L.V + s) * 2;
        d[0] = (R)U.r[e];
        d[1] = (R)U.i[e];
    }
}

template <typename Rin, typename Rout>
__global__ void __launch_bounds__(256) convert_kernel(size_t n, const Rin* __restrict__ src, Rout* __restrict__ dst) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < n) dst[gid] = (Rout)src[gid];
}

inline uint32_t grid_for(uint64_t n) { return (uint32_t)((n + 255) / 256); }

// ---------------------------------------------------------------- SU(3) check
//
// Deviation of every link from SU(3), in FP64 whatever the storage precision.  One thread
// per site checks its four links; each block writes three values: the number of links
// with a non-finite entry, max |(U U^+ - 1)_ab| and max |det U - 1| over its finite links.
// Counts and maxima do not depend on the order of combination, so the result is the
// same on every run.

struct Dev3 {
    double bad, du, dd;
};

__device__ __forceinline__ Dev3 dev3_join(Dev3 a, Dev3 b) {
    return {a.bad + b.bad, fmax(a.du, b.du), fmax(a.dd, b.dd)};
}

// block-wide combination of Dev3 (256 threads); valid in thread 0
__device__ __forceinline__ Dev3 dev3_block(Dev3 v) {
    __shared__ double sh[3][4];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
        v = dev3_join(v, Dev3{__shfl_xor(v.bad, m), __shfl_xor(v.du, m), __shfl_xor(v.dd, m)});
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][w] = v.bad;
        sh[1][w] = v.du;
        sh[2][w] = v.dd;
    }
    __syncthreads();
    Dev3 r = {0.0, 0.0, 0.0};
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) r = dev3_join(r, Dev3{sh[0][k], sh[1][k], sh[2][k]});
    return r;
}

template <typename FV>
__global__ void __launch_bounds__(256) check_kernel(Lat L, FV f, double* __restrict__ part) {
    const uint32_t s = blockIdx.x * 256u + threadIdx.x;
    Dev3 v = {0.0, 0.0, 0.0};
    if (s < L.V) {
#pragma unroll 1
        for (int mu = 0; mu < 4; ++mu) {
            const auto u = load_link(f, mu, s);
            double ur[9], ui[9];
            bool fin = true;
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                ur[e] = (double)u.r[e];
                ui[e] = (double)u.i[e];
                fin = fin && isfinite(ur[e]) && isfinite(ui[e]);
            }
            if (!fin) {
                v.bad += 1.0;
                continue;
            }
            // (U U^+)_ab = sum_k U_ak conj(U_bk)
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    double mr = a == b ? -1.0 : 0.0, mi = 0.0;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        mr += ur[a * 3 + k] * ur[b * 3 + k] + ui[a * 3 + k] * ui[b * 3 + k];
                        mi += ui[a * 3 + k] * ur[b * 3 + k] - ur[a * 3 + k] * ui[b * 3 + k];
                    }
                    v.du = fmax(v.du, sqrt(mr * mr + mi * mi));
                }
            // det U by the first row's cofactors
            double dr = 0.0, di = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int j = (c + 1) % 3, k = (c + 2) % 3;
                // cofactor C_0c = U_1j U_2k - U_1k U_2j
                const double cr = ur[3 + j] * ur[6 + k] - ui[3 + j] * ui[6 + k] - ur[3 + k] * ur[6 + j] +
                                  ui[3 + k] * ui[6 + j];
                const double ci = ur[3 + j] * ui[6 + k] + ui[3 + j] * ur[6 + k] - ur[3 + k] * ui[6 + j] -
                                  ui[3 + k] * ur[6 + j];
                dr += ur[c] * cr - ui[c] * ci;
                di += ur[c] * ci + ui[c] * cr;
            }
            v.dd = fmax(v.dd, sqrt((dr - 1.0) * (dr - 1.0) + di * di));
        }
    }
    const Dev3 r = dev3_block(v);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = r.bad;
        part[3 * blockIdx.x + 1] = r.du;
        part[3 * blockIdx.x + 2] = r.dd;
    }
}

__global__ void __launch_bounds__(256) check_final_kernel(const double* __restrict__ part, uint32_t n,
                                                          double* __restrict__ out) {
    Dev3 v = {0.0, 0.0, 0.0};
    for (uint32_t i = threadIdx.x; i < n; i += 256u) v = dev3_join(v, Dev3{part[3 * i], part[3 * i + 1], part[3 * i + 2]});
    const Dev3 r = dev3_block(v);
    if (threadIdx.x == 0) {
        out[0] = r.bad;
        out[1] = r.du;
        out[2] = r.dd;
    }
}

}  // namespace

uint32_t check_blocks(const Lat& L) { return (L.V + 255u) / 256u; }

hipError_t launch_check_field(const Lat& L, const void* U, int is_f32, double* part, double* out3, hipStream_t st) {
    const uint32_t nb = check_blocks(L);
    with_view(L.V, U, is_f32, [&](auto f) { check_kernel<decltype(f)><<<nb, 256, 0, st>>>(L, f, part); });
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    check_final_kernel<<<1, 256, 0, st>>>(part, nb, out3);
    return hipGetLastError();
}

hipError_t launch_import(const Lat& L, const void* src, int src_is_f32, const int64_t s[7], void* dst,
                         int dst_is_f32, hipStream_t st) {
    const uint32_t g = grid_for((uint64_t)L.V * 4);
    if (src_is_f32 && dst_is_f32)
        import_kernel<float, float><<<g, 256, 0, st>>>(L, (const float*)src, s[0], s[1], s[2], s[3], s[4], s[5], s[6],
                                                       (float*)dst);
    else if (src_is_f32)
        import_kernel<float, double><<<g, 256, 0, st>>>(L, (const float*)src, s[0], s[1], s[2], s[3], s[4], s[5],
                                                        s[6], (double*)dst);
    else if (dst_is_f32)
        import_kernel<double, float><<<g, 256, 0, st>>>(L, (const double*)src, s[0], s[1], s[2], s[3], s[4], s[5],
                                                        s[6], (float*)dst);
    else
        import_kernel<double, double><<<g, 256, 0, st>>>(L, (const double*)src, s[0], s[1], s[2], s[3], s[4], s[5],
                                                         s[6], (double*)dst);
    return hipGetLastError();
}

hipError_t launch_export(const Lat& L, const void* src, int src_is_f32, void* dst, int dst_is_f32,
                         const int64_t s[7], hipStream_t st) {
    const uint32_t g = grid_for((uint64_t)L.V * 4);
    if (src_is_f32 && dst_is_f32)
        export_kernel<float, float><<<g, 256, 0, st>>>(L, (const float*)src, (float*)dst, s[0], s[1], s[2], s[3], s[4],
                                                       s[5], s[6]);
    else if (src_is_f32)
        export_kernel<float, double><<<g, 256, 0, st>>>(L, (const float*)src, (double*)dst, s[0], s[1], s[2], s[3],
                                                        s[4], s[5], s[6]);
    else if (dst_is_f32)
        export_kernel<double, float><<<g, 256, 0, st>>>(L, (const double*)src, (float*)dst, s[0], s[1], s[2], s[3],
                                                        s[4], s[5], s[6]);
    else
        export_kernel<double, double><<<g, 256, 0, st>>>(L, (const double*)src, (double*)dst, s[0], s[1], s[2], s[3],
                                                         s[4], s[5], s[6]);
    return hipGetLastError();
}

hipError_t launch_random_field(const Lat& L, uint64_t seed, double eps, void* dst, int is_f32, hipStream_t st) {
    const uint32_t g = grid_for((uint64_t)L.V * 4);
    if (is_f32)
        random_kernel<float><<<g, 256, 0, st>>>(L, seed, eps, (float*)dst);
    else
        random_kernel<double><<<g, 256, 0, st>>>(L, seed, eps, (double*)dst);
    return hipGetLastError();
}

hipError_t launch_convert(const Lat& L, const void* src, int src_is_f32, void* dst, int dst_is_f32, hipStream_t st) {
    const size_t n = (size_t)L.V * kPlanes;
    const uint32_t g = grid_for(n);
    if (src_is_f32 == dst_is_f32)
        return hipMemcpyAsync(dst, src, n * (src_is_f32 ? 4 : 8), hipMemcpyDeviceToDevice, st);
    if (src_is_f32)
        convert_kernel<float, double><<<g, 256, 0, st>>>(n, (const float*)src, (double*)dst);
    else
        convert_kernel<double, float><<<g, 256, 0, st>>>(n, (const double*)src, (float*)dst);
    return hipGetLastError();
}

}  // namespace flue
