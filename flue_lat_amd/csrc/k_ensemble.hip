// k_ensemble.hip -- ensembles of gauge fields with different extents processed in one
// launch per operation: the RK3 flow stage, the plaquette sum and the magnetic B.
//
// The fields lie back to back, each in the SoA layout of its own lattice.  Every
// configuration owns a contiguous range of logical blocks (256 sites each); a persistent
// grid walks the logical blocks in contiguous chunks per workgroup (chunks of one XCD
// adjacent), finds the configuration of a logical block by binary search in a table
// kept in device memory, and runs the same per-site (flow: per-link) code as the
// single-field kernels.
// Sums are reduced per logical block and then per configuration in a fixed order.
#include "flue_clover.hpp"
#include "flue_flow.hpp"
#include "flue_launch.hpp"

namespace flue {

namespace {

constexpr int kBS = 256;

// The table is read through the constant address space and its entries are used in
// place: uniform loads from it become scalar loads, so the lattice geometry and every
// buffer descriptor built from it stay in scalar registers.  (A local copy of an entry
// is indexed at run time by direction, lives in memory, and turns every link load into
// a loop over the distinct descriptors of the wave.)
using CEnsEntry = const __attribute__((address_space(4))) EnsEntry;

__device__ __forceinline__ const EnsEntry& entry(const EnsEntry* tab, int c) {
    return *(const EnsEntry*)((CEnsEntry*)tab + c);
}

__device__ __forceinline__ int find_cfg(const EnsEntry* __restrict__ tab, int n, uint32_t lb) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].blk0 <= lb)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// logical blocks [lo, hi) of this workgroup: chunks of `chunk` blocks (chunk = total /
// gridDim rounded up, from the host).  No division here: a division would be lowered to
// vector instructions, and every value derived from it -- the configuration entry and
// each buffer descriptor -- would then leave scalar registers.
__device__ __forceinline__ void my_range(uint32_t total, uint32_t chunk, uint32_t& lo, uint32_t& hi) {
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const uint64_t l = (uint64_t)b * chunk;
    lo = l < total ? (uint32_t)l : total;
    hi = total - lo > chunk ? lo + chunk : total;
}

template <typename R, bool BUF>
__global__ void __launch_bounds__(kBS) ens_flow_kernel(const EnsEntry* __restrict__ tab, int n, uint32_t total,
                                                       uint32_t chunk, const R* __restrict__ win, R* __restrict__ wout,
                                                       R* __restrict__ acc, double eps, double ca, double cb,
                                                       int reunit) {
    // units = (logical block, direction), direction fastest: one thread per link, and the
    // four units of a block, which read the same neighbour links, run close together
    uint32_t lo, hi;
    my_range(4u * total, chunk, lo, hi);
    for (uint32_t lu = lo; lu < hi; ++lu) {
        const uint32_t lb = lu >> 2;
        const int mu = (int)(lu & 3u);
        const EnsEntry& e = entry(tab, find_cfg(tab, n, lb));
        const uint32_t s = (lb - e.blk0) * kBS + threadIdx.x;
        if (s < e.L.V) {
            const FieldView<R, BUF> f{win + e.off, e.L.V};
            int c[4];
            site_coords(e.L, s, c);
            flow_link(e.L, f, wout + e.off, acc + e.acc_off, s, c, mu, eps, ca, cb, reunit);
        }
    }
}

// Plaquette-first stencil over the ensemble: the six plaquette parts of every site of every
// field (48 reals per site, config c's block at 2/3 of its field offset), then the links.
template <typename R, bool BUF>
__global__ void __launch_bounds__(kBS) ens_plaq_ta_kernel(const EnsEntry* __restrict__ tab, int n, uint32_t total,
                                                          uint32_t chunk, const R* __restrict__ win, R* __restrict__ T) {
    uint32_t lo, hi;
    my_range(total, chunk, lo, hi);
    for (uint32_t lb = lo; lb < hi; ++lb) {
        const EnsEntry& e = entry(tab, find_cfg(tab, n, lb));
        const uint32_t s = (lb - e.blk0) * kBS + threadIdx.x;
        if (s < e.L.V) {
            const FieldView<R, BUF> f{win + e.off, e.L.V};
            plaquette_ta_site(e.L, f, s, T + e.off / 72 * 48);
        }
    }
}

template <typename R, bool BUF>
__global__ void __launch_bounds__(kBS) ens_flow_pf_kernel(const EnsEntry* __restrict__ tab, int n, uint32_t total,
                                                          uint32_t chunk, const R* __restrict__ win,
                                                          const R* __restrict__ T, R* __restrict__ wout,
                                                          R* __restrict__ acc, double eps, double ca, double cb,
                                                          int reunit) {
    uint32_t lo, hi;
    my_range(4u * total, chunk, lo, hi);
    for (uint32_t lu = lo; lu < hi; ++lu) {
        const uint32_t lb = lu >> 2;
        const int mu = (int)(lu & 3u);
        const EnsEntry& e = entry(tab, find_cfg(tab, n, lb));
        const uint32_t s = (lb - e.blk0) * kBS + threadIdx.x;
        if (s < e.L.V) {
            const FieldView<R, BUF> f{win + e.off, e.L.V};
            int c[4];
            site_coords(e.L, s, c);
            flow_link_pf(e.L, f, T + e.off / 72 * 48, wout + e.off, acc + e.acc_off, s, c, mu, eps, ca, cb, reunit);
        }
    }
}

// OBS = 0: sum of ReTr over all 6V plaquettes; OBS = 1: sum of -ReTr F_i^2 (magnetic)
template <typename R, bool BUF, int OBS>
__global__ void __launch_bounds__(kBS) ens_sum_kernel(const EnsEntry* __restrict__ tab, int n, uint32_t total,
                                                      uint32_t chunk, const R* __restrict__ fld, double* __restrict__ partials) {
    uint32_t lo, hi;
    my_range(total, chunk, lo, hi);
    for (uint32_t lb = lo; lb < hi; ++lb) {
        const EnsEntry& e = entry(tab, find_cfg(tab, n, lb));
        const uint32_t s = (lb - e.blk0) * kBS + threadIdx.x;
        double v = 0.0;
        if (s < e.L.V) {
            const FieldView<R, BUF> f{fld + e.off, e.L.V};
            v = OBS == 0 ? plaquette_site(e.L, f, s, 0, 3, 3) : magnetic_site(e.L, f, s);
        }
        const DD r = block_sum<kBS>(DD{v, 0.0});
        if (threadIdx.x == 0) {
            partials[2 * lb] = r.hi;
            partials[2 * lb + 1] = r.lo;
        }
        __syncthreads();  // block_sum's shared slots are reused by the next logical block
    }
}

// one workgroup per configuration: out[c * stride] = scale_c * sum of its block partials
__global__ void __launch_bounds__(kBS) ens_finalize_kernel(const EnsEntry* __restrict__ tab,
                                                           const double* __restrict__ partials,
                                                           double* __restrict__ out, int stride, int per_volume) {
    const EnsEntry e = tab[blockIdx.x];
    const uint32_t lo = e.blk0 + (uint32_t)(((uint64_t)e.nblk * threadIdx.x) / kBS);
    const uint32_t hi = e.blk0 + (uint32_t)(((uint64_t)e.nblk * (threadIdx.x + 1)) / kBS);
    DD acc = {0.0, 0.0};
    for (uint32_t i = lo; i < hi; ++i) acc = dd_add(acc, DD{partials[2 * i], partials[2 * i + 1]});
    const DD r = block_sum<kBS>(acc);
    if (threadIdx.x == 0) {
        const double v = r.hi + r.lo;
        out[(size_t)blockIdx.x * stride] = per_volume ? v / (double)e.L.V : v;
    }
}

template <typename R>
__global__ void __launch_bounds__(kBS) copy_kernel(size_t n, const R* __restrict__ src, R* __restrict__ dst) {
    for (size_t i = (size_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (size_t)gridDim.x * kBS) dst[i] = src[i];
}

template <typename Fn>
void with_real(int is_f32, bool buf, Fn&& fn) {
    if (is_f32) {
        if (buf)
            fn(float(0), std::true_type());
        else
            fn(float(0), std::false_type());
    } else {
        if (buf)
            fn(double(0), std::true_type());
        else
            fn(double(0), std::false_type());
    }
}

}  // namespace

hipError_t launch_ens_flow_stage(const EnsEntry* tab, int n, uint32_t total, uint32_t maxV, const void* win,
                                 int is_f32, void* wout, void* acc, double eps, double ca, double cb, int reunit,
                                 uint32_t grid, void* T, hipStream_t st) {
    with_real(is_f32, fits_buffer(maxV, is_f32), [&](auto r, auto buf) {
        using R = decltype(r);
        constexpr bool B = decltype(buf)::value;
        const uint32_t chunk4 = (4u * total + grid - 1) / grid;
        if (T) {
            ens_plaq_ta_kernel<R, B><<<grid, kBS, 0, st>>>(tab, n, total, (total + grid - 1) / grid, (const R*)win, (R*)T);
            ens_flow_pf_kernel<R, B><<<grid, kBS, 0, st>>>(tab, n, total, chunk4, (const R*)win, (const R*)T, (R*)wout,
                                                           (R*)acc, eps, ca, cb, reunit);
        } else {
            ens_flow_kernel<R, B><<<grid, kBS, 0, st>>>(tab, n, total, chunk4, (const R*)win, (R*)wout, (R*)acc, eps,
                                                        ca, cb, reunit);
        }
    });
    return hipGetLastError();
}

hipError_t launch_ens_sum(const EnsEntry* tab, int n, uint32_t total, uint32_t maxV, const void* fld, int is_f32,
                          int obs, double* partials, uint32_t grid, hipStream_t st) {
    const uint32_t chunk = (total + grid - 1) / grid;
    with_real(is_f32, fits_buffer(maxV, is_f32), [&](auto r, auto buf) {
        using R = decltype(r);
        constexpr bool B = decltype(buf)::value;
        if (obs == 0)
            ens_sum_kernel<R, B, 0><<<grid, kBS, 0, st>>>(tab, n, total, chunk, (const R*)fld, partials);
        else
            ens_sum_kernel<R, B, 1><<<grid, kBS, 0, st>>>(tab, n, total, chunk, (const R*)fld, partials);
    });
    return hipGetLastError();
}

hipError_t launch_ens_finalize(const EnsEntry* tab, int n, const double* partials, double* out, int stride,
                               int per_volume, hipStream_t st) {
    ens_finalize_kernel<<<n, kBS, 0, st>>>(tab, partials, out, stride, per_volume);
    return hipGetLastError();
}

hipError_t launch_copy(size_t n_reals, int is_f32, const void* src, void* dst, hipStream_t st) {
    const uint32_t g = (uint32_t)std::min<size_t>((n_reals + kBS - 1) / kBS, 8192);
    if (is_f32)
        copy_kernel<float><<<g, kBS, 0, st>>>(n_reals, (const float*)src, (float*)dst);
    else
        copy_kernel<double><<<g, kBS, 0, st>>>(n_reals, (const double*)src, (double*)dst);
    return hipGetLastError();
}

}  // namespace flue
