// k_clover.hip -- 5-loop improved field strength (hep-lat/0203008) and the magnetic B^2.
//
// The leaf machinery (corner chains, anti-Hermitian accumulation) is in flue_clover.hpp.
#include "flue_clover.hpp"
#include "flue_launch.hpp"

namespace flue {

namespace {

constexpr int kBS = 256;

// B = -(1/V) sum_x sum_i ReTr(F_i F_i) over the planes (y,x), (x,z), (z,y)
// (src/wloops.F90:502-506, 530-543); this kernel writes one partial sum per block.
template <typename FV, int WPE>
__global__ void __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
magnetic_kernel(Lat L, FV f, double* __restrict__ partials) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * kBS + threadIdx.x;
    double acc = 0.0;
    if (s < L.V) {
        int c[4];
        site_coords(L, s, c);
#pragma unroll 1
        for (int pl = 0; pl < 3; ++pl) {
            // (mu, nu) = (y, x), (x, z), (z, y) in 0-based directions
            const int mu = pl == 0 ? 2 : (pl == 1 ? 1 : 3);
            const int nu = pl == 0 ? 1 : (pl == 1 ? 3 : 2);
            const int cmu = pl == 0 ? c[2] : (pl == 1 ? c[1] : c[3]);
            const int cnu = pl == 0 ? c[1] : (pl == 1 ? c[3] : c[2]);
            acc += f_norm2(plane_ah<false>(plane_geom(L, f, s, cmu, cnu, mu, nu)));
        }
    }
    DD r = block_sum<kBS>(DD{acc, 0.0});
    if (threadIdx.x == 0) {
        partials[2 * blk] = r.hi;
        partials[2 * blk + 1] = r.lo;
    }
}

struct OutStrides {
    int64_t t, x, y, z, a, b;
};

// per-site F for one plane.  layout 0: complex128 at the given strides; layout 1: SoA
// planes (element e = a*3+b, re/im) in the field precision.
template <typename FV, bool CLOVER>
__global__ void __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(sizeof(typename FV::real) == 4 ? 2 : 1)))
fmunu_kernel(Lat L, FV f, int mu, int nu, void* __restrict__ out, int layout, OutStrides os) {
    using R = typename FV::real;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * kBS + threadIdx.x;
    if (s >= L.V) return;
    int c[4];
    site_coords(L, s, c);
    const AH<R> h = plane_ah<CLOVER>(plane_geom(L, f, s, c[mu], c[nu], mu, nu));
    // F = A^T / 4: F_aa = i d_a / 4, F_ab = -conj(o_ab)/8, F_ba = o_ab/8 (a < b)
    R fr[9], fi[9];
    constexpr int pa[3] = {0, 0, 1}, pb[3] = {1, 2, 2};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        fr[k * 4] = R(0);
        fi[k * 4] = h.d[k] * R(0.25);
        const int a = pa[k], b = pb[k];
        fr[a * 3 + b] = -h.or_[k] * R(0.125);
        fi[a * 3 + b] = h.oi[k] * R(0.125);
        fr[b * 3 + a] = h.or_[k] * R(0.125);
        fi[b * 3 + a] = h.oi[k] * R(0.125);
    }
    if (layout == 0) {
        double* o = (double*)out;
        const int64_t base = c[0] * os.t + c[1] * os.x + c[2] * os.y + c[3] * os.z;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int64_t off = 2 * (base + a * os.a + b * os.b);
                o[off] = (double)fr[a * 3 + b];
                o[off + 1] = (double)fi[a * 3 + b];
            }
    } else {
        R* o = (R*)out;
#pragma unroll
        for (int e = 0; e < 9; ++e) {
            o[(size_t)(2 * e) * L.V + s] = fr[e];
            o[(size_t)(2 * e + 1) * L.V + s] = fi[e];
        }
    }
}

}  // namespace

uint32_t magnetic_blocks(const Lat& L, int) { return (L.V + kBS - 1) / kBS; }

int magnetic_default_variant(int is_f32) { return is_f32 ? 2 : 1; }

// variant = occupancy target in waves per SIMD (1, 2 or 3); 0 picks the default
hipError_t launch_magnetic(const Lat& L, const void* U, int is_f32, double* partials, hipStream_t st, int variant) {
    const uint32_t g = magnetic_blocks(L, is_f32);
    if (variant == 0) variant = magnetic_default_variant(is_f32);
    if (variant < 1 || variant > 3) return hipErrorInvalidValue;
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        using R = typename F::real;
        if constexpr (!F::buffered) {
            // fields beyond the buffer-resource range: one variant only
            magnetic_kernel<F, sizeof(R) == 4 ? 2 : 1><<<g, kBS, 0, st>>>(L, f, partials);
        } else {
            if (variant == 1)
                magnetic_kernel<F, 1><<<g, kBS, 0, st>>>(L, f, partials);
            else if (variant == 2)
                magnetic_kernel<F, 2><<<g, kBS, 0, st>>>(L, f, partials);
            else
                magnetic_kernel<F, 3><<<g, kBS, 0, st>>>(L, f, partials);
        }
    });
    return hipGetLastError();
}

hipError_t launch_fmunu(const Lat& L, const void* U, int is_f32, int mu, int nu, int clover_only, void* out,
                        int out_layout, const int64_t s6[6], hipStream_t st) {
    const uint32_t g = (L.V + kBS - 1) / kBS;
    OutStrides os{0, 0, 0, 0, 0, 0};
    if (s6) os = OutStrides{s6[0], s6[1], s6[2], s6[3], s6[4], s6[5]};
    with_view(L.V, U, is_f32, [&](auto f) {
        if (clover_only)
            fmunu_kernel<decltype(f), true><<<g, kBS, 0, st>>>(L, f, mu, nu, out, out_layout, os);
        else
            fmunu_kernel<decltype(f), false><<<g, kBS, 0, st>>>(L, f, mu, nu, out, out_layout, os);
    });
    return hipGetLastError();
}

}  // namespace flue
