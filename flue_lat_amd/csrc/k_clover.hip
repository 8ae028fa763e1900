// k_clover.hip -- 5-loop improved field strength (hep-lat/0203008) and the magnetic B^2.
//
// The leaf machinery (corner chains, anti-Hermitian accumulation) is in flue_clover.hpp.
#include "flue_clover.hpp"
#include "flue_launch.hpp"

namespace flue {

namespace {

constexpr int kBS = 256;

// Two forms of the magnetic kernel, picked per precision (or by the autotuner):
//
// site-major: one thread per site does the three planes in turn, sites in storage order.
//
// plane-major: the grid is (nb, 3) and blockIdx.y is the plane, so the dispatcher runs
// the planes one after another.  The blocks of one XCD run through a contiguous range
// of a per-plane visit order; each order makes the sites in flight whole layers of
// the two plane directions, stacked along the third spatial direction (rho), so the
// +-3 reach of the leaves stays inside layers that are already in that XCD's L2:
//   plane 0 (y,x), rho = z:  x, y, z, t            (the storage order)
//   plane 1 (x,z), rho = y:  x, z, y, t
//   plane 2 (z,y), rho = x:  x mod xw, y, z, x / xw, t   (x in chunks of xw sites, at
//                            least 128 bytes of one element plane, so loads coalesce)
// This cuts the traffic behind L2 several-fold; it pays where occupancy is low (FP64).
struct PlaneOrder {
    uint32_t xw;   // x chunk width for plane 2 (divides NX)
};

__device__ __forceinline__ uint32_t plane_site(const Lat& L, const PlaneOrder& po, int pl, uint32_t o, int c[4]) {
    const uint32_t nx = (uint32_t)L.n[1], ny = (uint32_t)L.n[2], nz = (uint32_t)L.n[3];
    uint32_t q;
    if (pl == 0) {
        site_coords(L, o, c);
    } else if (pl == 1) {
        c[1] = (int)(o % nx);
        q = o / nx;
        c[3] = (int)(q % nz);
        q /= nz;
        c[2] = (int)(q % ny);
        c[0] = (int)(q / ny);
    } else {
        const uint32_t xi = o % po.xw;
        q = o / po.xw;
        c[2] = (int)(q % ny);
        q /= ny;
        c[3] = (int)(q % nz);
        q /= nz;
        const uint32_t nxc = nx / po.xw;
        c[1] = (int)((q % nxc) * po.xw + xi);
        c[0] = (int)(q / nxc);
    }
    return (uint32_t)c[1] + nx * ((uint32_t)c[2] + ny * ((uint32_t)c[3] + nz * (uint32_t)c[0]));
}

// (mu, nu) of magnetic plane pl: (y, x), (x, z), (z, y) in 0-based directions
__device__ __forceinline__ int plane_mu(int pl) { return pl == 0 ? 2 : (pl == 1 ? 1 : 3); }
__device__ __forceinline__ int plane_nu(int pl) { return pl == 0 ? 1 : (pl == 1 ? 3 : 2); }

// B = -(1/V) sum_x sum_i ReTr(F_i F_i) over the planes (y,x), (x,z), (z,y)
// (src/wloops.F90:502-506, 530-543); each block writes one compensated partial sum.
template <typename FV, int WPE, bool PLANE_MAJOR, bool C2 = sizeof(typename FV::real) == 4>
__global__ void __launch_bounds__(kBS) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
magnetic_kernel(Lat L, FV f, PlaneOrder po, uint32_t nb, double* __restrict__ partials) {
    // C2 (FP32 always; FP64 for links declared SU(3)): link chains carry two rows and row 2
    // is rebuilt where a leaf closes.  Plain FP64 keeps full products, so it makes no
    // assumption on the links.  Consecutive quadrants share their lines where the registers
    // allow it: always with C2, at one wave per SIMD otherwise.
    constexpr bool kC2 = C2;
    // FP64 at two waves per SIMD: the waiting chains are parked in LDS (quadrant_parked)
    constexpr bool kPark = sizeof(typename FV::real) == 8 && WPE == 2;
    constexpr bool kShare = !kPark && (kC2 || WPE == 1);
    // FP64 at one wave per SIMD: each group of products loads the next group's first link
    constexpr bool kLA = sizeof(typename FV::real) == 8 && WPE == 1;
    // the plane comes from blockIdx.y so that it, and every buffer descriptor derived
    // from it, stays in scalar registers (a division of blockIdx.x would not)
    const int pl = (int)blockIdx.y;
    const uint32_t blk = xcd_block(blockIdx.x, nb);
    const uint32_t o = blk * kBS + threadIdx.x;
    double acc = 0.0;
    if (o < L.V) {
        int c[4];
        if constexpr (kPark) {
            __shared__ typename FV::real park_mem[2 * 18 * kBS];
            const Park<typename FV::real, kBS> pk{park_mem + threadIdx.x};
            if constexpr (PLANE_MAJOR) {
                const uint32_t s = plane_site(L, po, pl, o, c);
                const int mu = plane_mu(pl), nu = plane_nu(pl);
                acc = f_norm2(plane_ah_parked<kC2>(plane_geom(L, f, s, c[mu], c[nu], mu, nu), pk));
            } else {
                site_coords(L, o, c);
#pragma unroll 1
                for (int p = 0; p < 3; ++p) {
                    const int mu = plane_mu(p), nu = plane_nu(p);
                    acc += f_norm2(plane_ah_parked<kC2>(plane_geom(L, f, o, c[mu], c[nu], mu, nu), pk));
                }
            }
        } else if constexpr (PLANE_MAJOR) {
            const uint32_t s = plane_site(L, po, pl, o, c);
            const int mu = plane_mu(pl), nu = plane_nu(pl);
            acc = f_norm2(plane_ah<false, kShare, kC2, kLA>(plane_geom(L, f, s, c[mu], c[nu], mu, nu)));
        } else {
            site_coords(L, o, c);
#pragma unroll 1
            for (int p = 0; p < 3; ++p) {
                const int mu = plane_mu(p), nu = plane_nu(p);
                acc += f_norm2(plane_ah<false, kShare, kC2, kLA>(plane_geom(L, f, o, c[mu], c[nu], mu, nu)));
            }
        }
    }
    DD r = block_sum<kBS>(DD{acc, 0.0});
    if (threadIdx.x == 0) {
        const uint32_t i = (uint32_t)pl * nb + blk;
        partials[2 * i] = r.hi;
        partials[2 * i + 1] = r.lo;
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
    uint32_t s;
    if (mu == 0 || nu == 0) {
        // the leaves reach 3 slices either way: blocks take the same sites of consecutive
        // slices one after another, so the blocks running together share those slices in L2
        const uint32_t NT = (uint32_t)L.n[0];
        const uint32_t chunk = blk / NT, t = blk - chunk * NT;
        const uint32_t idx = chunk * kBS + threadIdx.x;
        if (idx >= L.Vs) return;
        s = slice_site_along(L, idx, mu + nu) + L.Vs * t;  // mu + nu: the spatial direction
    } else {
        s = blk * kBS + threadIdx.x;
        if (s >= L.V) return;
        if (mu + nu == 4) {  // plane (x, z): each slice taken along z
            const uint32_t t = s / L.Vs;
            s = slice_site_along(L, s - t * L.Vs, 3) + L.Vs * t;
        }
    }
    int c[4];
    site_coords(L, s, c);
    const AH<R> h = plane_ah<CLOVER, sizeof(R) == 4, sizeof(R) == 4>(plane_geom(L, f, s, c[mu], c[nu], mu, nu));
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

uint32_t magnetic_blocks(const Lat& L, int) { return 3 * ((L.V + kBS - 1) / kBS); }

// x chunk width of plane 2: the smallest divisor of NX spanning >= 128 bytes, else NX
static PlaneOrder plane_order(const Lat& L, size_t real_bytes) {
    PlaneOrder po{(uint32_t)L.n[1]};
    for (int w = 1; w < L.n[1]; ++w)
        if (L.n[1] % w == 0 && (size_t)w * real_bytes >= 128) {
            po.xw = (uint32_t)w;
            break;
        }
    return po;
}

int magnetic_default_variant(int is_f32) { return is_f32 ? 5 : 4; }

// variant 1..3: site-major with an occupancy target of 1..3 waves per SIMD;
// variant 4..6: plane-major with 1..3 waves per SIMD; 7: rectangle-first (k_rect.hip, needs
// fpbuf and rect_fits); 0 picks the precision's default.
// su3 (FP64 only) selects two-row chains for the plane-major forms at 1 and 2 waves per
// SIMD; the other forms, and fields beyond the buffer-resource range, keep full products.
// *npart receives the number of partial sums written (nb or 3 nb).
hipError_t launch_magnetic(const Lat& L, const void* U, int is_f32, int su3, double* partials, hipStream_t st,
                           int variant, uint32_t* npart, void* fpbuf) {
    const uint32_t nb = magnetic_blocks(L, is_f32) / 3;
    if (variant == 0) variant = magnetic_default_variant(is_f32);
    if (variant < 1 || variant > kMagneticVariants) return hipErrorInvalidValue;
    if (variant == kMagneticRect) {
        if (!fpbuf) return hipErrorInvalidValue;
        return launch_magnetic_rect(L, U, is_f32, su3, partials, fpbuf, st, npart);
    }
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        using R = typename F::real;
        const PlaneOrder po = plane_order(L, sizeof(R));
        const dim3 g1(nb, 1), g3(nb, 3);
        if constexpr (!F::buffered) {
            // fields beyond the buffer-resource range: one form only
            magnetic_kernel<F, sizeof(R) == 4 ? 2 : 1, false><<<g1, kBS, 0, st>>>(L, f, po, nb, partials);
            variant = 1;
        } else if (sizeof(R) == 8 && su3 && (variant == 4 || variant == 5)) {
            if (variant == 4)
                magnetic_kernel<F, 1, true, true><<<g3, kBS, 0, st>>>(L, f, po, nb, partials);
            else
                magnetic_kernel<F, 2, true, true><<<g3, kBS, 0, st>>>(L, f, po, nb, partials);
        } else {
            switch (variant) {
                case 1: magnetic_kernel<F, 1, false><<<g1, kBS, 0, st>>>(L, f, po, nb, partials); break;
                case 2: magnetic_kernel<F, 2, false><<<g1, kBS, 0, st>>>(L, f, po, nb, partials); break;
                case 3: magnetic_kernel<F, 3, false><<<g1, kBS, 0, st>>>(L, f, po, nb, partials); break;
                case 4: magnetic_kernel<F, 1, true><<<g3, kBS, 0, st>>>(L, f, po, nb, partials); break;
                case 5: magnetic_kernel<F, 2, true><<<g3, kBS, 0, st>>>(L, f, po, nb, partials); break;
                default: magnetic_kernel<F, 3, true><<<g3, kBS, 0, st>>>(L, f, po, nb, partials); break;
            }
        }
    });
    if (npart) *npart = variant >= 4 ? 3 * nb : nb;
    return hipGetLastError();
}

hipError_t launch_fmunu(const Lat& L, const void* U, int is_f32, int mu, int nu, int clover_only, void* out,
                        int out_layout, const int64_t s6[6], hipStream_t st) {
    // planes containing time: every slice's blocks, slice fastest
    const uint32_t g = (mu == 0 || nu == 0) ? (uint32_t)L.n[0] * ((L.Vs + kBS - 1) / kBS) : (L.V + kBS - 1) / kBS;
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
