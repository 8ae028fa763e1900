// k_rect.hip -- the magnetic B with rectangle-first clover sums (magnetic variant 7).
//
// Planes (x, y) and (x, z): one wavefront per 16 x 16 tile of base sites in one plane slice,
// x along the lanes (16 consecutive sites, so link loads coalesce) and four rows of the second
// direction per loop iteration.  Every base site forms its seven rectangles once (rect_base) and adds
// the AH parts of the leaves at the four corners of each into the tile's footprint, 19 x 19
// cells in LDS (the tile plus the three rows and columns its rectangles reach beyond it).
// The footprint has one wavefront, whose LDS additions execute in program order, so each
// cell's sum is formed in a fixed order.  The footprints go to memory; the gather adds to
// each site of a tile the overlapping cells of the tiles before it along x, along the second
// direction and diagonally (in that fixed order), and sums |F|^2.  Plane (z, y) has no x
// direction to coalesce along: its z and y links are first copied with z fastest (an LDS-tiled
// transpose), which makes it plane (x', y') of the lattice with extents (NT, NZ, NY, NX), and
// the same two kernels run there.  The gather can also store each site's AH parts instead of
// summing |F|^2: the clover sums of the B^2 adjoint.
#include "flue_rect.hpp"
#include "flue_launch.hpp"

namespace flue {

namespace {

constexpr int kRT = 16;             // tile edge (base sites)
constexpr int kFP = kRT + 3;        // footprint edge
constexpr int kFC = kFP * kFP;      // footprint cells
constexpr int kRW = 64;             // one wavefront per tile

// the second direction and the spectator spatial direction of rectangle plane rp (0: (x, y),
// 1: (x, z))
__device__ __forceinline__ int rect_d2(int rp) { return rp == 0 ? 2 : 3; }

struct TileIdx {
    uint32_t tx, t2, cs, ct;  // tile along x, tile along d2, spectator coordinate, t
};

__device__ __forceinline__ TileIdx tile_idx(const Lat& L, int rp, uint32_t tile) {
    const uint32_t ntx = (uint32_t)L.n[1] / kRT, nt2 = (uint32_t)L.n[rect_d2(rp)] / kRT;
    const uint32_t ns = (uint32_t)L.n[rp == 0 ? 3 : 2];
    TileIdx t;
    t.tx = tile % ntx;
    uint32_t q = tile / ntx;
    t.t2 = q % nt2;
    q /= nt2;
    t.cs = q % ns;
    t.ct = q / ns;
    return t;
}

__device__ __forceinline__ uint32_t tile_of(const Lat& L, int rp, uint32_t tx, uint32_t t2, uint32_t cs,
                                            uint32_t ct) {
    const uint32_t ntx = (uint32_t)L.n[1] / kRT, nt2 = (uint32_t)L.n[rect_d2(rp)] / kRT;
    const uint32_t ns = (uint32_t)L.n[rp == 0 ? 3 : 2];
    return ((ct * ns + cs) * nt2 + t2) * ntx + tx;
}

// Footprint accumulation.  FP64 fields: LDS atomic adds, element e of cell c at [e * kFC + c].
// FP32 fields: the cell's 9 values padded to 12 floats at [c * 12], read, added and written
// back as three 16-byte LDS accesses, volatile so they stay in program order.  That is safe
// because the footprint belongs to one wavefront: its LDS operations execute in order and the
// lanes of one access address distinct cells.  (FP32 LDS atomic adds kept the waves waiting to
// issue LDS instructions for two thirds of their cycles; FP64 ones do not.)
template <typename R>
__device__ __forceinline__ void fp_add(R* fp, int cell, const AH<R>& p) {
    if constexpr (sizeof(R) == 4) {
        typedef float f4 __attribute__((ext_vector_type(4)));
        typedef __attribute__((address_space(3))) volatile f4 lds_f4;
        lds_f4* q = (lds_f4*)(fp + cell * 12);
        f4 a = q[0], b = q[1], c = q[2];
        a.x += p.d[0];
        a.y += p.d[1];
        a.z += p.d[2];
        a.w += p.or_[0];
        b.x += p.or_[1];
        b.y += p.or_[2];
        b.z += p.oi[0];
        b.w += p.oi[1];
        c.x += p.oi[2];
        q[0] = a;
        q[1] = b;
        q[2] = c;
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            __hip_atomic_fetch_add(fp + k * kFC + cell, p.d[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(fp + (3 + k) * kFC + cell, p.or_[k], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(fp + (6 + k) * kFC + cell, p.oi[k], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
}

// LDS reals per footprint cell
template <typename R>
constexpr int kCellReals = sizeof(R) == 4 ? 12 : 9;

// WPE == 1 (FP64) takes the chains with a one-link lookahead (rect_base); at two waves per
// SIMD (FP32) the other wave covers the load latency
template <typename FV, bool C2, int WPE>
__global__ void __launch_bounds__(kRW) __attribute__((amdgpu_waves_per_eu(WPE, WPE)))
rect_kernel(Lat L0, FV f0, Lat Lzy, FV fzy, uint32_t ntiles, typename FV::real* __restrict__ fpout) {
    using R = typename FV::real;
    constexpr int CR = kCellReals<R>;
    __shared__ __align__(16) R fp[CR * kFC];
    // blockIdx.y: 0 plane (x,y), 1 plane (x,z), 2 plane (z,y) as plane (x,y) of the copy
    const bool zy = blockIdx.y == 2;
    const Lat L = zy ? Lzy : L0;
    const FV f = zy ? fzy : f0;
    const int rp = zy ? 0 : (int)blockIdx.y;
    const int d2 = rect_d2(rp);
    const uint32_t tile = xcd_block(blockIdx.x, ntiles);
    const TileIdx ti = tile_idx(L, rp, tile);
    for (int k = threadIdx.x; k < CR * kFC; k += kRW) fp[k] = R(0);
    __syncthreads();
    const int lx = (int)threadIdx.x & 15;
#pragma unroll 1
    for (int r = 0; r < kRT / 4; ++r) {
        const int ly = ((int)threadIdx.x >> 4) + 4 * r;
        const int cx = (int)ti.tx * kRT + lx, c2 = (int)ti.t2 * kRT + ly;
        const int cy = rp == 0 ? c2 : (int)ti.cs, cz = rp == 0 ? (int)ti.cs : c2;
        const uint32_t s = (uint32_t)cx + L.st[2] * (uint32_t)cy + L.st[3] * (uint32_t)cz + L.st[0] * ti.ct;
        const PlaneGeom<FV> g = plane_geom(L, f, s, cx, c2, 1, d2);
        rect_base<C2, WPE == 1>(make_quad<1, 1, true>(g), [&](int da, int db, const AH<R>& p) {
            fp_add(fp, (ly + db) * kFP + lx + da, p);
        });
    }
    __syncthreads();
    // to memory as [e * kFC + c] in both precisions
    R* out = fpout + ((size_t)blockIdx.y * ntiles + tile) * (9 * kFC);
    if constexpr (CR == 12) {
        for (int c = threadIdx.x; c < kFC; c += kRW)
#pragma unroll
            for (int e = 0; e < 9; ++e) out[e * kFC + c] = fp[c * 12 + e];
    } else {
        for (int k = threadIdx.x; k < 9 * kFC; k += kRW) out[k] = fp[k];
    }
}

// per tile: the 256 sites' clover AH parts from the four footprints that cover them, then
// sum |F|^2 into one compensated partial per tile
// gather block: one wavefront, four sites per thread (measured faster than a thread per site)
constexpr int kGB = 64;

// the clover AH parts of tile cell (cx, cy): its own footprint's cell, then the overlapping
// cells of the tiles before it along x, along d2 and diagonally, in that order
template <typename R>
__device__ __forceinline__ AH<R> cell_sum(const R* own, const R* left, const R* below, const R* diag, int cx,
                                          int cy) {
    const int c = cy * kFP + cx;
    R v[9];
#pragma unroll
    for (int e = 0; e < 9; ++e) {
        R x = own[e * kFC + c];
        if (cx < 3) x += left[e * kFC + c + kRT];
        if (cy < 3) x += below[e * kFC + c + kRT * kFP];
        if (cx < 3 && cy < 3) x += diag[e * kFC + c + kRT * kFP + kRT];
        v[e] = x;
    }
    AH<R> h;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        h.d[k] = v[k];
        h.or_[k] = v[3 + k];
        h.oi[k] = v[6 + k];
    }
    return h;
}

// STORE = false: sum |F|^2 over the tile into one compensated partial per tile.  STORE = true:
// write each site's AH parts into Abuf in the B^2 adjoint's layout (9 planes of V values per
// magnetic plane (y,x), (x,z), (z,y): d0 d1 d2, o01r o02r o12r, o01i o02i o12i).
template <typename R, bool STORE>
__global__ void __launch_bounds__(kGB) rect_gather_kernel(Lat L0, Lat Lzy, const R* __restrict__ fpin,
                                                         uint32_t ntiles, double* __restrict__ partials,
                                                         R* __restrict__ Abuf) {
    const bool zy = blockIdx.y == 2;
    const Lat L = zy ? Lzy : L0;
    const int rp = zy ? 0 : (int)blockIdx.y;
    const uint32_t tile = xcd_block(blockIdx.x, ntiles);
    const TileIdx ti = tile_idx(L, rp, tile);
    const uint32_t ntx = (uint32_t)L.n[1] / kRT, nt2 = (uint32_t)L.n[rect_d2(rp)] / kRT;
    const uint32_t txl = (ti.tx + ntx - 1) % ntx, t2b = (ti.t2 + nt2 - 1) % nt2;
    const R* base = fpin + (size_t)blockIdx.y * ntiles * (9 * kFC);
    const R* own = base + (size_t)tile * (9 * kFC);
    const R* left = base + (size_t)tile_of(L, rp, txl, ti.t2, ti.cs, ti.ct) * (9 * kFC);
    const R* below = base + (size_t)tile_of(L, rp, ti.tx, t2b, ti.cs, ti.ct) * (9 * kFC);
    const R* diag = base + (size_t)tile_of(L, rp, txl, t2b, ti.cs, ti.ct) * (9 * kFC);
    const int cx = (int)threadIdx.x & 15;
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < kRT * kRT / kGB; ++r) {
        const int cy = ((int)threadIdx.x >> 4) + r * (kGB / kRT);
        if constexpr (STORE) {
            const AH<R> h = cell_sum(own, left, below, diag, cx, cy);
            // the site in the field's own lattice: for plane 2, the copy's x', y', z' are z, y, x
            const uint32_t c2 = ti.t2 * kRT + (uint32_t)cy, c1 = ti.tx * kRT + (uint32_t)cx;
            uint32_t s;
            if (zy)
                s = ti.cs + L0.st[2] * c2 + L0.st[3] * c1 + L0.st[0] * ti.ct;
            else if (rp == 0)
                s = c1 + L0.st[2] * c2 + L0.st[3] * ti.cs + L0.st[0] * ti.ct;
            else
                s = c1 + L0.st[2] * ti.cs + L0.st[3] * c2 + L0.st[0] * ti.ct;
            // the adjoint's planes are (y,x), (x,z), (z,y): the first is the tiles' (x,y) reversed
            const R sg = blockIdx.y == 0 ? R(-1) : R(1);
            R* o = Abuf + (size_t)blockIdx.y * 9 * L0.V + s;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                o[(size_t)k * L0.V] = sg * h.d[k];
                o[(size_t)(3 + k) * L0.V] = sg * h.or_[k];
                o[(size_t)(6 + k) * L0.V] = sg * h.oi[k];
            }
        } else {
            acc += f_norm2(cell_sum(own, left, below, diag, cx, cy));
        }
    }
    if constexpr (!STORE) {
        const DD s = block_sum<kGB>(DD{acc, 0.0});
        if (threadIdx.x == 0) {
            const uint32_t i = blockIdx.y * ntiles + tile;
            partials[2 * i] = s.hi;
            partials[2 * i + 1] = s.lo;
        }
    }
}

// U_z and U_y of the field (lattice L) into P, z fastest: P holds two direction blocks of
// 9 complex planes, [k][e][s'] with s' = z + NZ (y + NY (x + NX t)), k = 0 for U_z, 1 for
// U_y.  One block per 16 x 16 (x, z) tile at fixed (y, t), through a padded LDS tile so that
// both the reads (along x) and the writes (along z) coalesce.
template <typename R>
__global__ void __launch_bounds__(256) zy_relayout_kernel(Lat L, const R* __restrict__ U, R* __restrict__ P) {
    using P2 = typename Pair2<R>::type;
    __shared__ P2 tile[16][17];
    const uint32_t NX = (uint32_t)L.n[1], NY = (uint32_t)L.n[2], NZ = (uint32_t)L.n[3];
    uint32_t q = blockIdx.x;
    const uint32_t bx = q % (NX / 16);
    q /= NX / 16;
    const uint32_t bz = q % (NZ / 16);
    q /= NZ / 16;
    const uint32_t y = q % NY, t = q / NY;
    const uint32_t i = threadIdx.x >> 4, j = threadIdx.x & 15;
    const uint32_t s = bx * 16 + j + NX * (y + NY * (bz * 16 + i + NZ * t));
    const uint32_t s2 = bz * 16 + j + NZ * (y + NY * (bx * 16 + i + NX * t));
    const P2* src = reinterpret_cast<const P2*>(U);
    P2* dst = reinterpret_cast<P2*>(P);
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
        const int mu = k == 0 ? 3 : 2;
#pragma unroll 1
        for (int e = 0; e < 9; ++e) {
            tile[i][j] = src[(size_t)(mu * 9 + e) * L.V + s];
            __syncthreads();
            {
                typedef R r2 __attribute__((ext_vector_type(2)));
                const P2 v = tile[j][i];
                r2 w;
                w.x = v.x;
                w.y = v.y;
                st_stream(reinterpret_cast<r2*>(dst) + (size_t)(k * 9 + e) * L.V + s2, w);
            }
            __syncthreads();
        }
    }
}

// lattice whose x, y, z are the z, y, x of L: its plane (x, y) is L's plane (z, y)
Lat zy_lattice(const Lat& L) {
    const int32_t d[4] = {L.n[0], L.n[3], L.n[2], L.n[1]};
    return make_lat(d);
}

}  // namespace

bool rect_fits(const Lat& L) { return L.n[1] % kRT == 0 && L.n[2] % kRT == 0 && L.n[3] % kRT == 0; }

uint32_t rect_tiles(const Lat& L) { return L.V / (kRT * kRT); }

// the three planes' footprints, then the z-fastest copy of U_z and U_y
size_t rect_fp_bytes(const Lat& L, int is_f32) {
    const size_t sz = is_f32 ? 4 : 8;
    return (size_t)3 * rect_tiles(L) * 9 * kFC * sz + (size_t)36 * L.V * sz;
}

// partials: rect_tiles for each of the planes (x, y), (x, z), (z, y).  With Abuf, the per-site
// AH parts go there instead (the B^2 adjoint's clover sums) and no partials are written.
hipError_t launch_magnetic_rect(const Lat& L, const void* U, int is_f32, int su3, double* partials, void* fpbuf,
                                hipStream_t st, uint32_t* npart, void* Abuf) {
    if (!rect_fits(L) || !fits_buffer(L.V, is_f32)) return hipErrorInvalidValue;
    const uint32_t nt = rect_tiles(L);
    const Lat Lzy = zy_lattice(L);
    with_view(L.V, U, is_f32, [&](auto f) {
        using F = decltype(f);
        using R = typename F::real;
        if constexpr (F::buffered) {
            R* fp = (R*)fpbuf;
            R* P = fp + (size_t)3 * nt * 9 * kFC;
            zy_relayout_kernel<R><<<L.V / 256, 256, 0, st>>>(L, f.p, P);
            // the copy read as a field of Lzy whose directions 1 and 2 are its two blocks (the
            // view's base is one direction block before P; directions 0 and 3 are never read)
            const F fzy{(const R*)((uintptr_t)P - (uintptr_t)18 * L.V * sizeof(R)), L.V};
            const dim3 g(nt, 3);
            // FP32 and declared-SU(3) FP64 links: two-row chains; plain FP64: full products
            if constexpr (sizeof(R) == 4) {
                rect_kernel<F, true, 2><<<g, kRW, 0, st>>>(L, f, Lzy, fzy, nt, fp);
            } else {
                if (su3)
                    rect_kernel<F, true, 1><<<g, kRW, 0, st>>>(L, f, Lzy, fzy, nt, fp);
                else
                    rect_kernel<F, false, 1><<<g, kRW, 0, st>>>(L, f, Lzy, fzy, nt, fp);
            }
            if (Abuf)
                rect_gather_kernel<R, true><<<g, kGB, 0, st>>>(L, Lzy, fp, nt, nullptr, (R*)Abuf);
            else
                rect_gather_kernel<R, false><<<g, kGB, 0, st>>>(L, Lzy, fp, nt, partials, nullptr);
        }
    });
    if (npart) *npart = 3 * nt;
    return hipGetLastError();
}

}  // namespace flue
