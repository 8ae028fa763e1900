// flue_launch.hpp -- host-side launchers implemented by the kernel files and used by
// the C ABI (capi.hip).  All launchers are asynchronous on the given stream and
// return a hipError_t from the launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "flue_common.hpp"

namespace flue {

// field layout conversion: complex input/output with element strides (in complex
// elements) for the index order (t, x, y, z, mu, a, b)
hipError_t launch_import(const Lat& L, const void* src, int src_is_f32, const int64_t strides[7], void* dst,
                         int dst_is_f32, hipStream_t st);
hipError_t launch_export(const Lat& L, const void* src, int src_is_f32, void* dst, int dst_is_f32,
                         const int64_t strides[7], hipStream_t st);
hipError_t launch_random_field(const Lat& L, uint64_t seed, double eps, void* dst, int is_f32, hipStream_t st);
hipError_t launch_convert(const Lat& L, const void* src, int src_is_f32, void* dst, int dst_is_f32, hipStream_t st);
// SU(3) check of every link: part holds 3 doubles per block (check_blocks), out3 receives
// (links with a non-finite entry, max |U U^+ - 1|, max |det U - 1|)
uint32_t check_blocks(const Lat& L);
hipError_t launch_check_field(const Lat& L, const void* U, int is_f32, double* part, double* out3, hipStream_t st);

// reductions: each writes one compensated partial (hi, lo) per block to `partials`,
// then launch_finalize sums them in a fixed order into out[0] (double, device)
// plaquette walk lengths (time slices per thread), one per variant; variant 0 = the default
constexpr int kPlaquetteVariants = 6;
constexpr int kPlaqWalks[kPlaquetteVariants] = {2, 3, 4, 6, 8, 12};
constexpr int kPlaquetteDefault = 4;  // walks of 6 slices
int plaquette_walk(int variant);
uint32_t plaquette_blocks(const Lat& L, int variant);
uint32_t plaquette_max_blocks(const Lat& L);
// *npart (optional) receives the number of partials written
hipError_t launch_plaquette(const Lat& L, const void* U, int is_f32, int mu0, int mu1, int nu1, double* partials,
                            hipStream_t st, int variant, uint32_t* npart);
uint32_t polyakov_blocks(const Lat& L);
hipError_t launch_polyakov(const Lat& L, const void* U, int is_f32, double* partials, hipStream_t st);
// blocks of path_trace over a slice (tslice >= 0) or the whole lattice (-1); with path ==
// nullptr, enough for any path
uint32_t path_blocks(const Lat& L, int tslice, const int32_t* path, int npath);
hipError_t launch_path_trace(const Lat& L, const void* U, int is_f32, const int32_t* path, int npath, int tslice,
                             double* partials, hipStream_t st);
hipError_t launch_generic_path(const Lat& L, const void* U, int is_f32, const int32_t base[4], const int32_t* path,
                               int npath, double* out18, hipStream_t st);
hipError_t launch_finalize(const double* partials, uint32_t n, double* out, double scale, hipStream_t st);

// 5-loop (or 1x1 clover) field strength for one plane; out_layout 0 = complex128 with
// strides (t, x, y, z, a, b) in complex elements, 1 = SoA planes (9 elem x re/im) in the
// field precision
hipError_t launch_fmunu(const Lat& L, const void* U, int is_f32, int mu, int nu, int clover_only, void* out,
                        int out_layout, const int64_t strides[6], hipStream_t st);
uint32_t magnetic_blocks(const Lat& L, int is_f32);
int magnetic_default_variant(int is_f32);
// rectangular Wilson-loop table (k_wloops.hip): R, T up to kMaxLoopRT
constexpr int kMaxLoopRT = 6;
// tslice value selecting one table per time slice
constexpr int kEverySlice = -2;
uint32_t wloop_blocks(const Lat& L, int tslice, int mu, int nu);
size_t wloop_partial_bytes(const Lat& L, int tslice);
hipError_t launch_wloop_table(const Lat& L, const void* U, int is_f32, int mu, int nu, int Rmax, int Tmax,
                              int tslice, double* partials, double* out, double scale, hipStream_t st);

// scalar gluon propagator (k_gluon.hip): momentum-space D over 0 <= q < dext of the
// transformed field Uq (lattice Lq of its extents), complex128 at strides dstr (t,x,y,z);
// and the position-space form into D(NS,NS,NS,NS) (Fortran order, FP64)
hipError_t launch_mom_scalar(const Lat& Lq, const void* Uq, int is_f32, int mu_start, double xi,
                             const int32_t dext[4], const int64_t dstr[4], double* D, hipStream_t st);
hipError_t launch_scalar_gluon(const Lat& L, const void* U, int is_f32, int NS, double* D, hipStream_t st);

// partial sums per block of the first stage of a long fixed-order reduction
constexpr uint32_t kFinalChunk = 2048;

// number of magnetic kernel variants the autotuner chooses from: 1..6 forms x occupancy
// targets (k_clover.hip), 7 rectangle-first (k_rect.hip; NX, NY, NZ multiples of 16 and a
// footprint buffer)
constexpr int kMagneticVariants = 7;
constexpr int kMagneticRect = 7;
// static device scratch of the allocation-free magnetic entry point: partials for lattices
// of up to 44 M sites (48^3x96 needs 2 MiB)
constexpr size_t kMagneticScratchBytes = (size_t)8 << 20;
// fpbuf: the footprint buffer of variant 7 (rect_fp_bytes), unused by the others
hipError_t launch_magnetic(const Lat& L, const void* U, int is_f32, int su3, double* partials, hipStream_t st,
                           int variant, uint32_t* npart, void* fpbuf = nullptr);
bool rect_fits(const Lat& L);
uint32_t rect_tiles(const Lat& L);
size_t rect_fp_bytes(const Lat& L, int is_f32);
hipError_t launch_magnetic_rect(const Lat& L, const void* U, int is_f32, int su3, double* partials, void* fpbuf,
                                hipStream_t st, uint32_t* npart, void* Abuf = nullptr);

// gradients and flow (k_grad.hip, k_b2grad.hip); grad / Wout / Abuf are SoA in the field
// precision, and so is acc (8 su(3) coordinates x 4 directions x V)
uint32_t grad_blocks(const Lat& L);
// stencil variants (k_grad.hip): 1 = staple sums in storage order, 2..6 = staple sums in
// (z, t) brick orders, 7 = plaquette-first (the six plaquette parts of every site into T,
// plaq_ta_bytes, then each link from 3 links and 6 stored parts; unitary links only),
// 8 = plaquette-first with both passes walking along t, 9..13 = plaquette-first in the (z, t)
// brick orders of 2..6
constexpr int kFlowVariants = 13;
constexpr int kFlowPF = 7;
constexpr int kFlowPFWalk = 8;
constexpr int kFlowDefault = 1;
// whether a flow/gradient variant is a plaquette-first form (needs plaq_ta_bytes of workspace)
constexpr bool flow_is_pf(int variant) { return variant >= kFlowPF && variant <= kFlowVariants; }
// the staple-sum variant with the visit order of `variant` (storage order for 7 and 8)
constexpr int staple_order_of(int variant) {
    return variant > kFlowPFWalk ? variant - kFlowPF : (flow_is_pf(variant) ? kFlowDefault : variant);
}
size_t plaq_ta_bytes(const Lat& L, int is_f32);
hipError_t launch_wilson_grad(const Lat& L, const void* U, int is_f32, double beta, void* grad, hipStream_t st,
                              int variant, void* T);
hipError_t launch_flow_stage(const Lat& L, const void* Win, int is_f32, void* Wout, void* acc, double eps,
                             double ca, double cb, int reunit, hipStream_t st, int variant, void* T);
hipError_t launch_magnetic_ah(const Lat& L, const void* U, int is_f32, void* Abuf, hipStream_t st);
hipError_t launch_b2_grad(const Lat& L, const void* U, int is_f32, const void* Abuf, void* grad, hipStream_t st);
// reverse-mode B^2 adjoint: slab by slab of time slices, per-rectangle terms into `slots`
// (b2_slot_bytes), then a fixed-order sum per link; slabs are sized so the slot buffer holds
// at most kB2SlotReals reals (4 GiB in FP64, 2 GiB in FP32; at least one slice).  At 48^3
// that is 3 slices in both precisions, which ran faster than 1, 7 or 14.
constexpr uint64_t kB2SlotReals = (uint64_t)1 << 29;
// B^2 adjoint forms: 1 = per-link gather, 2 = reverse mode (default)
constexpr int kB2Variants = 2;
constexpr int kB2DefaultVariant = 2;
uint32_t b2_slab_slices(const Lat& L, int is_f32);
size_t b2_slot_bytes(const Lat& L, int is_f32);
hipError_t launch_b2_grad_rev(const Lat& L, const void* U, int is_f32, const void* Abuf, void* slots, void* grad,
                              hipStream_t st);
hipError_t launch_finalize_affine(const double* partials, uint32_t n, double* out, double scale, double offset,
                                  hipStream_t st);

// ensembles (k_ensemble.hip): configuration c's field starts at element `off` of the
// back-to-back fields, its flow accumulator (reals of the fields' precision) at element
// `acc_off`, and it owns logical
// blocks [blk0, blk0 + nblk) of 256 sites
struct EnsEntry {
    Lat L;
    uint64_t off;
    uint64_t acc_off;
    uint32_t blk0;
    uint32_t nblk;
};
// T != nullptr: the plaquette-first stencil with the plaquette parts of every field in T
// (48 reals per site, field c's at 2/3 of its field offset); T == nullptr: staple sums
hipError_t launch_ens_flow_stage(const EnsEntry* tab, int n, uint32_t total, uint32_t maxV, const void* win,
                                 int is_f32, void* wout, void* acc, double eps, double ca, double cb, int reunit,
                                 uint32_t grid, void* T, hipStream_t st);
hipError_t launch_ens_sum(const EnsEntry* tab, int n, uint32_t total, uint32_t maxV, const void* fld, int is_f32,
                          int obs, double* partials, uint32_t grid, hipStream_t st);
hipError_t launch_ens_finalize(const EnsEntry* tab, int n, const double* partials, double* out, int stride,
                               int per_volume, hipStream_t st);
hipError_t launch_copy(size_t n_reals, int is_f32, const void* src, void* dst, hipStream_t st);
// ensemble flow variants: persistent grid of 1, 2, 4 or 8 workgroups per CU with the staple
// stencil (1..4) or the plaquette-first stencil (5..8)
constexpr int kEnsembleVariants = 8;
constexpr int kEnsembleDefault = 3;

}  // namespace flue
