// k_loops.hip -- plaquette family, Polyakov loop, generic-path products and the
// deterministic second pass of every lattice sum.
#include "flue_flow.hpp"
#include <algorithm>

#include "flue_launch.hpp"

namespace flue {

namespace {

constexpr int kBS = 256;
constexpr int kMaxPath = 64;

struct PathArg {
    int n;
    int8_t d[kMaxPath];  // signed directions, 1-based as in the reference (mu=1 is time)
};

__device__ __forceinline__ uint32_t site_of(const Lat& L, const int c[4]) {
    return (uint32_t)c[1] + L.st[2] * (uint32_t)c[2] + L.st[3] * (uint32_t)c[3] + L.st[0] * (uint32_t)c[0];
}

__device__ __forceinline__ void write_partial(double* partials, uint32_t blk, DD r) {
    if (threadIdx.x == 0) {
        partials[2 * blk] = r.hi;
        partials[2 * blk + 1] = r.lo;
    }
}

// sum over sites and planes mu0 <= mu <= mu1, mu < nu <= nu1 of
// ReTr[(U_mu(x) U_nu(x+mu)) (U_nu(x) U_mu(x+nu))^+]   (src/wloops.F90:947-1006)
//
// One thread walks kPlaqLT consecutive time slices of one spatial site.  The spatial links
// at x + t, which the temporal planes (0, nu) read as neighbours, are kept in registers as
// the own links of the next slice, so each site reads its own links once and, per slice,
// 1 temporal link, 3 links at x + t and 9 links of spatial neighbours (against 21 link
// reads for the planes taken one by one).  The walk length lt is one of kPlaqWalks (the
// autotuner's variants); longer walks carry more links, shorter ones give more blocks.
template <typename FV>
__global__ void __launch_bounds__(kBS) plaquette_kernel(Lat L, FV f, int mu0, int mu1, int nu1, uint32_t nbs, int lt,
                                                        double* __restrict__ partials) {
    using R = typename FV::real;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t tc = blk / nbs, pb = blk - tc * nbs;
    const uint32_t p = pb * kBS + threadIdx.x;
    double acc = 0.0;
    if (p < L.Vs) {
        const int NT = L.n[0];
        const int t0 = (int)tc * lt, t1 = min(t0 + lt, NT);
        int c[4];
        site_coords(L, p, c);
        // forward spatial neighbours within a slice
        uint32_t sf[4];
#pragma unroll
        for (int d = 1; d < 4; ++d) sf[d] = c[d] == L.n[d] - 1 ? p - (uint32_t)(L.n[d] - 1) * L.st[d] : p + L.st[d];
        const bool tplanes = mu0 == 0;
        Mat<R> Us[3];
        uint32_t s = p + L.Vs * (uint32_t)t0;
        for (int t = t0; t < t1; ++t) {
            const uint32_t ts = L.Vs * (uint32_t)t;
            if (!tplanes || t == t0) {
#pragma unroll
                for (int nu = 1; nu < 4; ++nu)
                    if (nu <= nu1) Us[nu - 1] = load_link(f, nu, s);
            }
#pragma unroll
            for (int mu = 1; mu < 3; ++mu) {
                if (mu < mu0 || mu > mu1) continue;
#pragma unroll
                for (int nu = mu + 1; nu < 4; ++nu) {
                    if (nu > nu1) continue;
                    const Mat<R> A = mul(Us[mu - 1], load_link(f, nu, sf[mu] + ts));
                    const Mat<R> B = mul(Us[nu - 1], load_link(f, mu, sf[nu] + ts));
                    acc += (double)retr_nd(A, B);
                }
            }
            if (tplanes) {
                const uint32_t sn = t == NT - 1 ? p : s + L.Vs;  // x + t
                const Mat<R> U0 = load_link(f, 0, s);
#pragma unroll
                for (int nu = 1; nu < 4; ++nu) {
                    if (nu > nu1) continue;
                    const Mat<R> N = load_link(f, nu, sn);
                    const Mat<R> A = mul(U0, N);
                    const Mat<R> B = mul(Us[nu - 1], load_link(f, 0, sf[nu] + ts));
                    acc += (double)retr_nd(A, B);
                    Us[nu - 1] = N;
                }
            }
            s += L.Vs;
        }
    }
    write_partial(partials, blk, block_sum<kBS>(DD{acc, 0.0}));
}

// Polyakov loop: ordered product of the time links over t = 0..NT-1 at each spatial
// site (src/wloops.F90:1043-1063).  A block holds 64 spatial sites x 4 time segments;
// segment products are combined left to right in shared memory.
template <typename FV>
__global__ void __launch_bounds__(kBS) polyakov_kernel(Lat L, FV f, double* __restrict__ partials) {
    using R = typename FV::real;
    constexpr int NSEG = kBS / 64;
    __shared__ R sh[NSEG][18][64];
    const int lane = threadIdx.x & 63, seg = threadIdx.x >> 6;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t sp = blk * 64 + lane;
    const bool live = sp < L.Vs;
    const int NT = L.n[0];
    const int t0 = (seg * NT) / NSEG, t1 = ((seg + 1) * NT) / NSEG;
    Mat<R> P = identity<R>();
    if (live) {
        for (int t = t0; t < t1; ++t) {
            const Mat<R> U = load_link(f, 0, sp + L.Vs * (uint32_t)t);
            P = (t == t0) ? U : mul(P, U);
        }
    }
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        sh[seg][e][lane] = P.r[e];
        sh[seg][9 + e][lane] = P.i[e];
    }
    __syncthreads();
    double val = 0.0;
    if (seg == 0 && live) {
        Mat<R> T = P;
        for (int k = 1; k < NSEG; ++k) {
            Mat<R> Q;
#pragma unroll
            for (int e = 0; e < 9; ++e) {
                Q.r[e] = sh[k][e][lane];
                Q.i[e] = sh[k][9 + e][lane];
            }
            T = mul(T, Q);
        }
        val = (double)retr(T);
    }
    __syncthreads();
    write_partial(partials, blk, block_sum<kBS>(DD{val, 0.0}));
}

// product along a signed path starting at the site s with coordinates c (modified):
// +mu multiplies U_mu(x) then steps, -mu steps back then multiplies U_mu(x)^+
// (src/wloops.F90:36-55)
template <typename FV>
__device__ __forceinline__ Mat<typename FV::real> walk_path(const Lat& L, const FV& f, const PathArg& p, uint32_t s,
                                                            int c[4]) {
    using R = typename FV::real;
    Mat<R> M = identity<R>();
    for (int k = 0; k < p.n; ++k) {
        const int d = p.d[k];
        const int ax = (d > 0 ? d : -d) - 1;
        Mat<R> U;
        if (d > 0) {
            U = load_link(f, ax, s);
            s = step_fwd(L, s, c, ax);
        } else {
            s = step_bwd(L, s, c, ax);
            U = load_link_dag(f, ax, s);
        }
        M = (k == 0) ? U : mul(M, U);
    }
    return M;
}

template <typename FV>
__global__ void __launch_bounds__(kBS) path_trace_kernel(Lat L, FV f, PathArg p, int tslice, int by_slice,
                                                         double* __restrict__ partials) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    uint32_t s;
    bool live;
    if (tslice >= 0) {
        const uint32_t idx = blk * kBS + threadIdx.x;
        live = idx < L.Vs;
        s = idx + L.Vs * (uint32_t)tslice;
    } else if (by_slice) {
        // the same sites of consecutive slices in consecutive blocks
        const uint32_t NT = (uint32_t)L.n[0];
        const uint32_t chunk = blk / NT, t = blk - chunk * NT;
        const uint32_t idx = chunk * kBS + threadIdx.x;
        live = idx < L.Vs;
        s = idx + L.Vs * t;
    } else {
        s = blk * kBS + threadIdx.x;
        live = s < L.V;
    }
    double acc = 0.0;
    if (live) {
        int c[4];
        site_coords(L, s, c);
        acc = (double)retr(walk_path(L, f, p, s, c));
    }
    write_partial(partials, blk, block_sum<kBS>(DD{acc, 0.0}));
}

// one path product at one base point, written transposed as a Fortran-ordered
// complex128 3x3 (src/wloops.F90:56)
template <typename FV>
__global__ void generic_path_kernel(Lat L, FV f, PathArg p, int t, int x, int y, int z, double* __restrict__ out) {
    using R = typename FV::real;
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int c[4] = {t, x, y, z};
    const uint32_t s = site_of(L, c);
    const Mat<R> M = walk_path(L, f, p, s, c);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            out[2 * (a + 3 * b)] = (double)M.r[b * 3 + a];
            out[2 * (a + 3 * b) + 1] = (double)M.i[b * 3 + a];
        }
}

// fixed-order compensated sum of per-block partials
__global__ void __launch_bounds__(kBS) finalize_kernel(const double* __restrict__ partials, uint32_t n,
                                                       double* __restrict__ out, double scale, double offset) {
    // thread t sums entries t, t + kBS, ... in that order (coalesced, fixed order)
    DD acc = {0.0, 0.0};
    for (uint32_t i = threadIdx.x; i < n; i += kBS) acc = dd_add(acc, DD{partials[2 * i], partials[2 * i + 1]});
    DD r = block_sum<kBS>(acc);
    if (threadIdx.x == 0) out[0] = offset + (r.hi + r.lo) * scale;
}

// stage one of a long reduction: block b sums entries [b*kFinalChunk, (b+1)*kFinalChunk)
__global__ void __launch_bounds__(kBS) finalize_stage_kernel(const double* __restrict__ partials, uint32_t n,
                                                             double* __restrict__ out) {
    const uint32_t lo = blockIdx.x * kFinalChunk;
    const uint32_t hi = lo + kFinalChunk < n ? lo + kFinalChunk : n;
    DD acc = {0.0, 0.0};
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBS) acc = dd_add(acc, DD{partials[2 * i], partials[2 * i + 1]});
    const DD r = block_sum<kBS>(acc);
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = r.hi;
        out[2 * blockIdx.x + 1] = r.lo;
    }
}

PathArg make_path(const int32_t* path, int npath) {
    PathArg p;
    p.n = npath;
    for (int i = 0; i < kMaxPath; ++i) p.d[i] = (int8_t)(i < npath ? path[i] : 0);
    return p;
}

}  // namespace

// blocks per time chunk (the spatial sites of one slice), times the time chunks
static uint32_t plaquette_slice_blocks(const Lat& L) { return (L.Vs + kBS - 1) / kBS; }

int plaquette_walk(int variant) {
    return (variant >= 1 && variant <= kPlaquetteVariants) ? kPlaqWalks[variant - 1] : kPlaqWalks[kPlaquetteDefault - 1];
}

uint32_t plaquette_blocks(const Lat& L, int variant) {
    const int lt = plaquette_walk(variant);
    return plaquette_slice_blocks(L) * (uint32_t)((L.n[0] + lt - 1) / lt);
}

uint32_t plaquette_max_blocks(const Lat& L) {
    uint32_t m = 0;
    for (int v = 1; v <= kPlaquetteVariants; ++v) m = std::max(m, plaquette_blocks(L, v));
    return m;
}

hipError_t launch_plaquette(const Lat& L, const void* U, int is_f32, int mu0, int mu1, int nu1, double* partials,
                            hipStream_t st, int variant, uint32_t* npart) {
    const int lt = plaquette_walk(variant);
    const uint32_t g = plaquette_blocks(L, variant), nbs = plaquette_slice_blocks(L);
    with_view(L.V, U, is_f32, [&](auto f) {
        plaquette_kernel<decltype(f)><<<g, kBS, 0, st>>>(L, f, mu0, mu1, nu1, nbs, lt, partials);
    });
    if (npart) *npart = g;
    return hipGetLastError();
}

uint32_t polyakov_blocks(const Lat& L) { return (L.Vs + 63) / 64; }

hipError_t launch_polyakov(const Lat& L, const void* U, int is_f32, double* partials, hipStream_t st) {
    const uint32_t g = polyakov_blocks(L);
    with_view(L.V, U, is_f32, [&](auto f) { polyakov_kernel<decltype(f)><<<g, kBS, 0, st>>>(L, f, partials); });
    return hipGetLastError();
}

// Over the whole lattice, a path with a time step takes its blocks slice-fastest: the
// blocks in flight then cover the same sites of neighbouring slices and share the links the
// path reaches there (in storage order one slice lies hundreds of blocks on).
static bool path_has_time(const int32_t* path, int npath) {
    for (int i = 0; i < npath; ++i)
        if (path[i] == 1 || path[i] == -1) return true;
    return false;
}

uint32_t path_blocks(const Lat& L, int tslice, const int32_t* path, int npath) {
    const uint32_t per_slice = (L.Vs + kBS - 1) / kBS;
    if (tslice >= 0) return per_slice;
    const uint32_t plain = (L.V + kBS - 1) / kBS, by_slice = (uint32_t)L.n[0] * per_slice;
    if (!path) return std::max(plain, by_slice);
    return path_has_time(path, npath) ? by_slice : plain;
}

hipError_t launch_path_trace(const Lat& L, const void* U, int is_f32, const int32_t* path, int npath, int tslice,
                             double* partials, hipStream_t st) {
    if (npath < 1 || npath > kMaxPath) return hipErrorInvalidValue;
    const PathArg p = make_path(path, npath);
    const uint32_t g = path_blocks(L, tslice, path, npath);
    const int by_slice = tslice < 0 && path_has_time(path, npath);
    with_view(L.V, U, is_f32, [&](auto f) {
        path_trace_kernel<decltype(f)><<<g, kBS, 0, st>>>(L, f, p, tslice, by_slice, partials);
    });
    return hipGetLastError();
}

hipError_t launch_generic_path(const Lat& L, const void* U, int is_f32, const int32_t base[4], const int32_t* path,
                               int npath, double* out18, hipStream_t st) {
    if (npath < 0 || npath > kMaxPath) return hipErrorInvalidValue;
    const PathArg p = make_path(path, npath);
    with_view(L.V, U, is_f32, [&](auto f) {
        generic_path_kernel<decltype(f)><<<1, 64, 0, st>>>(L, f, p, base[0], base[1], base[2], base[3], out18);
    });
    return hipGetLastError();
}

// Long partial lists are reduced in two fixed-order stages: blocks of kFinalChunk
// entries into the tail area behind the list (partial_bytes reserves it), then those.
hipError_t launch_finalize_affine(const double* partials, uint32_t n, double* out, double scale, double offset,
                                  hipStream_t st) {
    if (n <= 2 * kFinalChunk) {
        finalize_kernel<<<1, kBS, 0, st>>>(partials, n, out, scale, offset);
        return hipGetLastError();
    }
    const uint32_t n2 = (n + kFinalChunk - 1) / kFinalChunk;
    double* tail = const_cast<double*>(partials) + 2 * (size_t)n;
    finalize_stage_kernel<<<n2, kBS, 0, st>>>(partials, n, tail);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    finalize_kernel<<<1, kBS, 0, st>>>(tail, n2, out, scale, offset);
    return hipGetLastError();
}

hipError_t launch_finalize(const double* partials, uint32_t n, double* out, double scale, hipStream_t st) {
    return launch_finalize_affine(partials, n, out, scale, 0.0, st);
}

}  // namespace flue
