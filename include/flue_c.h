/*
 * flue_c.h -- C ABI of libflue_hip.so, the MI355X implementation of FLUE's
 * gluonic-observable path (plaquette family, Polyakov loop, generic path
 * products, Wilson loops, 5-loop F_munu, magnetic B^2, their gradients, the
 * Wilson flow, and the scalar gluon propagator with its momentum-space analysis).
 *
 * Two groups of entry points:
 *
 * 1. Host drop-ins named like the reference's FFI layer (c_wrapper/FLUE_c.f90,
 *    wrapped by f2py as FLUE.compiled.flue_c with --lower).  They follow its
 *    conventions: every argument by reference, lattice dims first, arrays in
 *    Fortran order as complex128 (double _Complex) host memory
 *    data(NT,NX,NY,NZ,4,3,3), scalars next, results last.  Where the
 *    reference's Fortran procedure stops or has undefined behaviour, these
 *    return a nonzero status instead.  The GPU does the work: the field is
 *    copied to the device, re-laid out and measured there.
 *
 * 2. Device entry points (flue_*) on device-resident fields in the library's
 *    SoA layout: 36 complex planes (direction mu, colour a, b) of V sites, each
 *    site an adjacent (re, im) pair: real index (((mu*3 + a)*3 + b)*V + s)*2 + ri
 *    with s = x + NX*(y + NY*(z + NZ*t)), FP64 or FP32 (72 V reals in all).
 *    They are asynchronous on the given hipStream_t (NULL = default stream) and
 *    write results to device memory; the few that wait for the stream say so
 *    (flue_check_field, flue_b2_grad, and the graph-launching flow calls).  Lattice
 *    sums are deterministic (fixed reduction order) and use compensated FP64
 *    partial sums; `ws` is a device workspace of at least
 *    flue_workspace_bytes(dims) bytes (NULL = a buffer the library keeps per
 *    device, grown on demand -- do not pass NULL under stream capture).
 *
 * Directions follow the reference: mu = 1 is time, 2..4 are x, y, z
 * (1-based, as in src/wloops.F90), and dims are {NT, NX, NY, NZ}.
 * Every function returns FLUE_OK (0) or one of the FLUE_E* codes.
 */
#ifndef FLUE_C_H
#define FLUE_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
typedef struct {
    double re, im;
} flue_c128;
#else
typedef double _Complex flue_c128;
#endif

enum {
    FLUE_OK = 0,
    FLUE_EINVAL = 1,   /* argument value out of range */
    FLUE_ENULL = 2,    /* required pointer is NULL */
    FLUE_EPATH = 3,    /* path entry 0 or |mu| > 4, or path too long */
    FLUE_ENT = 4,      /* Polyakov loop needs NT >= 3 */
    FLUE_EDIM = 5,     /* lattice extent < 1 or volume too large */
    FLUE_EHIP = 6,     /* HIP runtime error */
    FLUE_ENOMEM = 7,   /* device allocation failed */
    FLUE_ENOTIMPL = 8, /* outside this library's scope (kept for the ABI; no entry returns it) */
    FLUE_ENODEV = 9,   /* no GPU visible */
    FLUE_EWS = 10,     /* workspace too small */
    FLUE_ESU3 = 11     /* links are not SU(3) within the bound (see flue_check_field) */
};

/* link precision of a SoA field.  FLUE_F64 makes no assumption on the links, as the
 * reference does not, except flue_b2_grad (below), which needs SU(3) links in every
 * precision and checks for them.  FLUE_F32 fields must hold SU(3) links (unitary, det 1,
 * to FP32 rounding; flue_check_field tests it): the clover kernels
 * carry two rows of each link product and rebuild the third as conj(row0 x row1).
 * Sums and per-site traces are FP64 in both.  FLUE_F64_SU3 is an FP64 field whose links
 * the caller declares SU(3) (to FP64 rounding): stored and computed as FLUE_F64, except
 * that the magnetic measurement (flue_magnetic, variants 4 and 5) carries two rows as FP32
 * does.  No reference counterpart: an opt-in for fields known to be gauge fields. */
enum { FLUE_F64 = 0, FLUE_F32 = 1, FLUE_F64_SU3 = 2 };
enum { FLUE_C128 = 0, FLUE_C64 = 1 };       /* element type of a complex array */
enum { FLUE_OUT_STRIDED = 0, FLUE_OUT_SOA = 1 };

const char* flue_strerror(int status);
int flue_version(int32_t* major, int32_t* minor, int32_t* patch);

/* ------------------------------------------------------------------ */
/* 1. host drop-ins (c_wrapper/FLUE_c.f90 conventions)                  */
/* ------------------------------------------------------------------ */

/* MultiplyMatMat_c, c_wrapper/FLUE_c.f90:31-36 (src/SU3MatrixOps.f90:19-37): MM = left*right */
int multiplymatmat_c(flue_c128 MM[9], const flue_c128 left[9], const flue_c128 right[9]);

/* get_qhat_c, c_wrapper/FLUE_c.f90:49-56 (src/gluonAna/momentum.f90:8-24):
 * qhat = (2 pi / a) coord / shape; a == NULL means a = 1 (the reference divides by
 * the absent argument, momentum.f90:23) */
int get_qhat_c(const double coord[4], const int32_t shape[4], double qhat[4], const double* a);

/* calc_mom_space_scalarD_c, c_wrapper/FLUE_c.f90:18-29 (src/gluonCalc/gluonProp.f90:112-250):
 * scalar gluon propagator from the Fourier-transformed field
 * U(0:NT, 0:NS, 0:NS, 0:NS, 0:3, 0:2, 0:2) (complex128, Fortran order) into
 * D(0:NT/2, 0:NS/2, 0:NS/2, 0:NS/2).  Computed on the GPU (flue_gluon_prop_mom).  Where the
 * reference's result is undefined: U(-q) is read at (E - q) mod E for the array extent E
 * (the reference reads index E - q, one past the end of the dimension when q = 0,
 * gluonProp.f90:177); the mu = 0 phase uses the t component of qhat (the reference reads
 * qhat(0), outside its 1-based array, :185); momenta past D's extent, which the reference
 * writes out of bounds when an extent E is even, are not computed.  The phase of
 * directions 1..3 and the prefactor conditions (:216, :225) are the reference's. */
int calc_mom_space_scalard_c(const int32_t* NS, const int32_t* NT, const flue_c128* U, flue_c128* D,
                             const int32_t* mu_start, const double* xi);

/* scalarGluonProp_c, c_wrapper/FLUE_c.f90:9-16 (src/gluonCalc/gluonProp.f90:17-110):
 * position-space field U(NT,NS,NS,NS,4,3,3) into D(NS,NS,NS,NS), on the GPU
 * (flue_scalar_gluon_prop).  Needs NT <= NS (the reference writes D at every (t,x,y,z)
 * of the field: FLUE_EDIM otherwise).  D is overwritten (the reference adds into its
 * intent(out) D without clearing it), entries with t > NT are 0, xi has no effect (it
 * enters only for mu == 0, which the reference's 1-based direction loop never reaches),
 * and the volume sum of identical terms (:69-73) is taken as V times the term. */
int scalargluonprop_c(const int32_t* NS, const int32_t* NT, const flue_c128* U, const double* xi, double* D);

/* Q_Average_c, c_wrapper/FLUE_c.f90:38-47 (src/gluonAna/gpManip.f90:10-57), host code:
 * averages q over the open bins (k dq, (k+1) dq), dq = max(q)/q_slices, until the next
 * bin's upper edge reaches max(q); DOUT takes D(1) for every kept bin and Qcount is one
 * more than the number of kept bins, as in the reference.  DOUT and QOUT hold NQ + 1
 * values.  FLUE_EINVAL for NQ < 1, q_slices < 1, a NaN in q, and inputs on which the
 * reference's loop does not end (max(q) < 0 with one slice). */
int q_average_c(const int32_t* NQ, const double* q, const double* D, const int32_t* q_slices, double* DOUT,
                double* QOUT, int32_t* Qcount);

/* cone_cut_c, c_wrapper/FLUE_c.f90:58-71 (src/gluonAna/gpManip.f90:59-163), host code:
 * keeps the momenta Q(NQ,4) (Fortran order) within IRRadiusIN of the (1,1,1,1) diagonal
 * and, above |q| = IRCutIN, at an angle below angleIN to it (the first component scaled
 * by xiIN after |q| is taken); packs them with D and D4 into QOUT, Dout, D4out.  radius,
 * NT and NS are accepted and unused, as in the reference when every optional is given. */
int cone_cut_c(const int32_t* NQ, const int32_t* radius, const double* Q, const double* D, const double* D4,
               const int32_t* NT, const int32_t* NS, const double* angleIN, const double* xiIN,
               const double* IRCutIN, const double* IRRadiusIN, double* QOUT, double* Dout, double* D4out,
               int32_t* qcount);

/* periodCoord, src/wloops.F90:546-565: single-step periodic wrap of a 1-based
 * coordinate; dataShape is the 7-entry SHAPE(data) */
int periodcoord_c(const int32_t coord[4], const int32_t dataShape[7], int32_t out[4]);

/* genericPath, src/wloops.F90:20-57: ordered product along path(1:npath) from the
 * 1-based coordBase, returned transposed as a Fortran 3x3 */
int genericpath_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ,
                  const flue_c128* data, const int32_t coordBase[4], const int32_t* npath, const int32_t* path,
                  flue_c128 out[9]);

/* genPlaquette, src/wloops.F90:59-127 */
int genplaquette_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ,
                   const flue_c128* data, const int32_t* muStart, const int32_t* muEnd, const int32_t* nuEnd,
                   double* sumTrP, int32_t* nP, double* time);

/* plaquette, src/wloops.F90:915-1017 (assumed-shape there; dims passed here) */
int plaquette_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
                const int32_t* muStart, const int32_t* muEnd, const int32_t* nuEnd, double* sumTrP, int32_t* nP,
                double* time);

/* polyakov, src/wloops.F90:1019-1069; NT < 3 returns FLUE_ENT */
int polyakov_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
               double* sumTrP, int32_t* nP, double* time);

/* cloverLoopMuNuCoord, src/wloops.F90:129-199: out(NT,NX,NY,NZ,3,3) */
int cloverloopmunucoord_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ,
                          const flue_c128* data, const int32_t* mu, const int32_t* nu, flue_c128* out);

/* Loop5MuNuCoord, src/wloops.F90:201-432: out(NT,NX,NY,NZ,3,3) */
int loop5munucoord_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ,
                     const flue_c128* data, const int32_t* mu, const int32_t* nu, flue_c128* out);

/* magnetic, src/wloops.F90:479-544 */
int magnetic_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
               double* B);

/* wmunu, src/wloops.F90:679-791 */
int wmunu_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
            const int32_t* mu, const int32_t* nu, const int32_t* nMu, const int32_t* nNu, double* loopVal);

/* wmunutau, src/wloops.F90:567-677 */
int wmunutau_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
               const int32_t* tau, const int32_t* mu, const int32_t* nu, const int32_t* nMu, const int32_t* nNu,
               double* loopVal);

/* ------------------------------------------------------------------ */
/* 2. device entry points (SoA fields)                                  */
/* ------------------------------------------------------------------ */

size_t flue_field_bytes(const int32_t dims[4], int32_t prec);
size_t flue_workspace_bytes(const int32_t dims[4]);

/* relayout between a complex array with element strides (in complex elements) for
 * the index order (t, x, y, z, mu, a, b) and a SoA field; the reference's Fortran
 * array has strides {1, NT, NT*NX, NT*NX*NY, V, 4V, 12V} (src/wloops.F90:60) */
int flue_import_field(const int32_t dims[4], const void* src, int32_t src_type, const int64_t strides[7],
                      void* dst, int32_t prec, void* stream);
int flue_export_field(const int32_t dims[4], const void* src, int32_t prec, void* dst, int32_t dst_type,
                      const int64_t strides[7], void* stream);
int flue_convert_field(const int32_t dims[4], const void* src, int32_t src_prec, void* dst, int32_t dst_prec,
                       void* stream);

/* SU(3) check of every link, computed in FP64 (no reference counterpart; the reference
 * readers reunitarise every link instead, src/gaugeIO/ILDG_bin.f90:138-151): stats[0] = the
 * number of links with a non-finite entry, stats[1] = max over links and entries of
 * |(U U^+ - 1)_ab|, stats[2] = max |det U - 1| (stats may be NULL).  Deterministic.  Waits
 * for `stream` and returns FLUE_ESU3 if stats[0] > 0 or stats[1] or stats[2] exceeds tol;
 * tol = 0 selects the bound of the precision: 1e-5 for FLUE_F32, 1e-10 for FLUE_F64 and
 * FLUE_F64_SU3.  ws needs flue_workspace_bytes(dims) bytes. */
int flue_check_field(const int32_t dims[4], const void* U, int32_t prec, double tol, double stats[3], void* ws,
                     size_t ws_bytes, void* stream);

/* seeded synthetic field U = fix(exp(i eps sum_a c_a lambda_a/2)), c_a ~ N(0,1) keyed
 * by (seed, link) through Philox4x32-10; no reference counterpart (the reference
 * reads fields from disk, src/gaugeIO/) */
int flue_random_field(const int32_t dims[4], uint64_t seed, double eps, void* dst, int32_t prec, void* stream);

/* genPlaquette / plaquette (src/wloops.F90:59-127, 915-1017): *d_sum = sum ReTr P */
int flue_plaquette(const int32_t dims[4], const void* U, int32_t prec, int32_t muStart, int32_t muEnd,
                   int32_t nuEnd, double* d_sum, void* ws, size_t ws_bytes, void* stream);

/* polyakov (src/wloops.F90:1019-1069): *d_sum = sum_xyz ReTr prod_t U_t */
int flue_polyakov(const int32_t dims[4], const void* U, int32_t prec, double* d_sum, void* ws, size_t ws_bytes,
                  void* stream);

/* sum over sites (tau = 0) or over the time slice tau (1-based) of ReTr genericPath
 * (src/wloops.F90:20-57); with the rectangular path this is wmunu / wmunutau times
 * the volume (src/wloops.F90:567-791) */
int flue_path_trace(const int32_t dims[4], const void* U, int32_t prec, const int32_t* path, int32_t npath,
                    int32_t tau, double* d_sum, void* ws, size_t ws_bytes, void* stream);

/* every rectangular Wilson loop of the plane (mu, nu) with 1 <= R <= Rmax steps along mu
 * and 1 <= T <= Tmax along nu (Rmax, Tmax <= 6), in one pass:
 * d_out[(R-1)*Tmax + (T-1)] = wmunu(data, mu, nu, R, T) for tau = 0 (src/wloops.F90:679-791)
 * or wmunutau(data, tau, mu, nu, R, T) for the 1-based time slice tau (:567-677);
 * mu != nu.  ws needs flue_workspace_bytes(dims) bytes. */
int flue_wilson_loops(const int32_t dims[4], const void* U, int32_t prec, int32_t mu, int32_t nu, int32_t Rmax,
                      int32_t Tmax, int32_t tau, double* d_out, void* ws, size_t ws_bytes, void* stream);

/* the same table for every time slice in one pass (static potential W(R, T, t)):
 * d_out[((R-1)*Tmax + (T-1))*NT + (t-1)] = wmunutau(data, t, mu, nu, R, T) for t = 1..NT
 * (src/wloops.F90:567-677) */
int flue_wilson_loops_slices(const int32_t dims[4], const void* U, int32_t prec, int32_t mu, int32_t nu,
                             int32_t Rmax, int32_t Tmax, double* d_out, void* ws, size_t ws_bytes, void* stream);

/* genericPath at one 1-based base point, transposed, as 9 complex128 in Fortran order
 * written to device memory d_out */
int flue_generic_path(const int32_t dims[4], const void* U, int32_t prec, const int32_t coordBase[4],
                      const int32_t* path, int32_t npath, double* d_out, void* stream);

/* Loop5MuNuCoord (variant 0, src/wloops.F90:201-432) or cloverLoopMuNuCoord
 * (variant 1, :129-199) for the plane (mu, nu).  layout FLUE_OUT_STRIDED: complex128
 * at element strides for (t, x, y, z, a, b); FLUE_OUT_SOA: 18 planes of V reals in the
 * field precision (element a*3+b, re/im) */
int flue_fmunu(const int32_t dims[4], const void* U, int32_t prec, int32_t mu, int32_t nu, int32_t variant,
               void* out, int32_t layout, const int64_t strides[6], void* stream);

/* calc_mom_space_scalarD (src/gluonCalc/gluonProp.f90:112-250) on a Fourier-transformed
 * field Uq held like a SoA field over its extents dims (36 complex planes of V elements,
 * the momenta q in place of the sites): within a plane, momentum q sits at
 * sum_i q_i qstrides[i] (complex elements, order t, x, y, z; NULL = the site order
 * x + NX(y + NY(z + NZ t)); the strides must keep every q inside the plane).  d_D
 * (complex128) at element strides dstrides for (t, x, y, z) over 0 <= q < dext,
 * dext[i] <= dims[i]/2 + 1 (the reference's loop range).  U(-q) is read at
 * (dims - q) mod dims; the phase of direction mu >= 1 uses the momentum component of
 * direction mu - 1 as the reference's qhat(mu) does, and that of mu = 0 the t component.
 * mu_start >= 0; directions mu_start..3 are summed. */
int flue_gluon_prop_mom(const int32_t dims[4], const void* Uq, int32_t prec, const int64_t qstrides[4],
                        int32_t mu_start, double xi, const int32_t dext[4], void* d_D, const int64_t dstrides[4],
                        void* stream);

/* scalarGluonProp (src/gluonCalc/gluonProp.f90:17-110) on a position-space field with
 * NT <= NX = NY = NZ = NS: d_D holds NS^4 doubles, D(t,x,y,z) in Fortran order, 0 for
 * t >= NT */
int flue_scalar_gluon_prop(const int32_t dims[4], const void* U, int32_t prec, double* d_D, void* stream);

/* magnetic (src/wloops.F90:479-544): *d_B = -(1/V) sum_x sum_i ReTr F_i(x)^2 */
int flue_magnetic(const int32_t dims[4], const void* U, int32_t prec, double* d_B, void* ws, size_t ws_bytes,
                  void* stream);

/* flue_magnetic with an explicit kernel variant (0 = the tuned choice, else 7 where it
 * applies, else the precision's default, 5 for FP32 and 4 for FP64):
 *   1..3  site-major (one thread does the three planes of its site), 1..3 waves per SIMD;
 *   4..6  plane-major (planes run one after another, sites in per-plane L2-friendly
 *         orders), 1..3 waves per SIMD;
 *   7     rectangle-first: each rectangle of the clover is formed once and its leaves at
 *         the other three corners follow by conjugation (plane (z,y) from a z-fastest copy
 *         of its links).  It needs NX, NY, NZ divisible by 16 and a workspace of
 *         flue_magnetic_work_bytes (about 300 bytes per site in FP32, 600 in FP64);
 *         otherwise the call runs the default form instead.
 * Every variant is deterministic; they agree to rounding (their summation orders differ). */
int flue_magnetic_ex(const int32_t dims[4], const void* U, int32_t prec, int32_t variant, double* d_B, void* ws,
                     size_t ws_bytes, void* stream);

/* device bytes of the workspace flue_magnetic_ex needs for every variant (flue_workspace_bytes
 * covers variants 1..6 only); 0 for invalid arguments */
size_t flue_magnetic_work_bytes(const int32_t dims[4], int32_t prec);

/* flue_magnetic_ex without a workspace and without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The per-block partial sums
 * live in a static device array of 8 MiB that each GPU holds until the process ends, enough
 * for lattices of up to about 44 M sites (48^3x96 needs 2 MiB); larger lattices return
 * FLUE_EWS before any launch.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (on different streams or host threads) race, so serialise them.  The
 * kernels write every partial the reduction reads, so no clearing is needed between calls.
 * Results are identical to flue_magnetic_ex with the same variant and an 8 MiB workspace:
 * variant 7 (also as the tuned choice) runs where its footprints fit the array (extents
 * divisible by 16 and up to about 28 k sites in FP32, 14 k in FP64) and as the default form
 * elsewhere. */
int flue_magnetic_noalloc(const int32_t dims[4], const void* U, int32_t prec, int32_t variant, double* d_B,
                          void* stream);

/* ------------------------------------------------------------------ */
/* 3. gradients and Wilson flow (no counterpart in the reference)       */
/* ------------------------------------------------------------------ */
/* Gradient convention: for a link U = U_mu(x), D^a f = d/ds f(exp(s T^a) U) at s = 0
 * with T^a = -i lambda_a / 2; the gradient field holds G = sum_a T^a D^a f per link
 * (traceless anti-Hermitian, SoA layout and precision of the input field), so
 * D^a f = -2 Tr(T^a G).  The Wilson action is S = beta sum_p (1 - ReTr P / 3), the
 * plaquette sum of genPlaquette (src/wloops.F90:59-127) rescaled. */

size_t flue_b2_work_bytes(const int32_t dims[4], int32_t prec);
size_t flue_flow_work_bytes(const int32_t dims[4], int32_t prec);

/* *d_S = beta * sum_x sum_{mu<nu} (1 - ReTr P_munu(x) / 3) */
int flue_wilson_action(const int32_t dims[4], const void* U, int32_t prec, double beta, double* d_S, void* ws,
                       size_t ws_bytes, void* stream);

/* G = (beta/6) P_TA(U Sigma) per link, Sigma the staple sum (the tuned FLUE_OP_FLOW visit
 * order of the staple form) */
int flue_wilson_grad(const int32_t dims[4], const void* U, int32_t prec, double beta, void* grad, void* stream);

/* bytes of the plaquette parts the plaquette-first stencil keeps (48 reals per site) */
size_t flue_pf_work_bytes(const int32_t dims[4], int32_t prec);

/* flue_wilson_grad with an explicit FLUE_OP_FLOW variant (0 = tuned): 1..6 staple sums in
 * the visit orders of the autotuner table, 7..13 plaquette-first: the six plaquettes of every
 * site are formed once and their traceless anti-Hermitian parts stored in ws
 * (flue_pf_work_bytes), then each link takes its up plaquettes at x and its down plaquettes
 * at x - nu conjugated by U_nu(x - nu) (7 in storage order, 8 walking along t, 9..13 in the
 * orders of 2..6).  That identity holds for unitary links, so these variants run for FLUE_F32
 * and FLUE_F64_SU3 fields given a workspace; FLUE_F64 fields and ws == NULL take the staple
 * sums in the same visit order (variant 1 for 7 and 8). */
int flue_wilson_grad_ex(const int32_t dims[4], const void* U, int32_t prec, double beta, int32_t variant, void* grad,
                        void* ws, size_t ws_bytes, void* stream);

/* flue_b2_grad with an explicit form: 1 = per-link gather, 2 = reverse mode (each leaf
 * rectangle walked once, its terms summed per link from fixed slots); 0 = tuned or 2 */
int flue_b2_grad_ex(const int32_t dims[4], const void* U, int32_t prec, int32_t variant, void* grad, void* work,
                    size_t work_bytes, void* stream);

/* gradient of the magnetic B (flue_magnetic); temporal links get 0.  `work` holds the
 * per-site anti-Hermitian parts of the three spatial-plane leaf sums, then the reverse form's
 * slots (and, on extents divisible by 16, the rectangle-first tiles that form those sums)
 * (flue_b2_work_bytes; NULL = library buffer).  Precondition in every precision: SU(3)
 * links, since the gather takes the product from a leaf corner to the walk's end as
 * X^+ F.  The call first runs flue_check_field's scan with the precision's bound (waiting
 * for `stream` once) and returns FLUE_ESU3, with no gradient written, on a field that
 * fails it. */
int flue_b2_grad(const int32_t dims[4], const void* U, int32_t prec, void* grad, void* work, size_t work_bytes,
                 void* stream);

/* one stage (0, 1, 2) of Luescher's RK3 Wilson flow with step eps, reading Win and
 * writing Wout (distinct buffers); acc is the accumulator the three stages of a step
 * share, 32 * V reals of the field's precision (doubles for FLUE_F64 and FLUE_F64_SU3,
 * floats for FLUE_F32; a buffer of 32 * V doubles serves every precision): stage 0 writes
 * it without reading it, stage 1 reads and writes it, stage 2 only reads it.  The
 * exponential map, the product and the reunitarisation after stage 2 are done in FP64. */
int flue_flow_stage(const int32_t dims[4], const void* Win, int32_t prec, void* Wout, double* acc, double eps,
                    int32_t stage, void* stream);

/* nsteps RK3 steps in place on U.  use_graph != 0 captures two steps in a hipGraph and
 * replays it (needs a non-default stream); `work` is flue_flow_work_bytes (NULL = library
 * buffer, which must then exist before any capture): two more link buffers, the
 * accumulator and the plaquette parts.  The stages alternate between U and the second
 * buffer, and an odd last step passes through the third, so every step ends in U without
 * a copy.  The flow moves SU(3) links along the
 * group (exponential map, reunitarisation after every step), so it takes SU(3) links in
 * every precision: the tuned FLUE_OP_FLOW variant, including the plaquette-first variants
 * 7..13, is used.  flue_flow_stage (no workspace) uses the staple sums in the tuned
 * variant's visit order. */
int flue_wilson_flow(const int32_t dims[4], void* U, int32_t prec, double eps, int32_t nsteps, int32_t use_graph,
                     void* work, size_t work_bytes, void* stream);

/* ------------------------------------------------------------------ */
/* autotuner (no reference counterpart): the variants of an operation are timed with       */
/* hipEvents on the caller's field; the fastest is remembered per (device, op, precision,  */
/* extents) and used by every later call of that operation (and by the *_ex entries when  */
/* their variant argument is 0).  The table persists through save/load.  Variants of one   */
/* operation compute the same quantity: the flow orders and the ensemble grids give        */
/* bit-identical results, the others agree to rounding (their summation orders differ).    */
/*   FLUE_OP_MAGNETIC       1..7  kernel form x waves per SIMD, or rectangle-first (7, timed  */
/*                                only where it applies) (flue_magnetic_ex)                 */
/*   FLUE_OP_B2_GRAD        1..2  per-link gather / reverse mode with slots (flue_b2_grad_ex) */
/*   FLUE_OP_PLAQUETTE      1..6  time slices per thread 2, 3, 4, 6, 8, 12 (flue_plaquette,   */
/*                                flue_wilson_action)                                        */
/*   FLUE_OP_FLOW          1..13  stencil of the flow and the Wilson gradient: staple sums   */
/*                                in storage order or (z, t) bricks 1x2, 2x2, 1x4, 2x4, 4x4   */
/*                                (1..6, bit-identical); plaquette-first (unitary links, see  */
/*                                flue_wilson_grad_ex) in storage order (7), walking along t  */
/*                                (8) or in the same five brick orders (9..13); 7 and 9..13   */
/*                                are bit-identical, the two stencils agree to rounding       */
/*   FLUE_OP_ENSEMBLE_FLOW  1..8  persistent grid of 1, 2, 4, 8 workgroups per CU with the   */
/*                                staple stencil (1..4) or the plaquette-first one (5..8)    */
/*                                (flue_ensemble_flow; tuned by flue_ensemble_autotune)      */
/* ------------------------------------------------------------------ */
enum { FLUE_OP_MAGNETIC = 0, FLUE_OP_B2_GRAD = 1, FLUE_OP_PLAQUETTE = 2, FLUE_OP_FLOW = 3, FLUE_OP_ENSEMBLE_FLOW = 4 };
/* device bytes flue_autotune needs as `ws` for op (0 for an invalid op) */
size_t flue_autotune_work_bytes(int32_t op, const int32_t dims[4], int32_t prec);
/* times every variant of op (not FLUE_OP_ENSEMBLE_FLOW) on U, reps calls each after a
 * warm-up (FLUE_OP_FLOW: one RK3 step into scratch links); U is only read.  ws:
 * flue_autotune_work_bytes(op, dims, prec) (NULL = library buffer).  Waits for the stream.
 * With the environment variable FLUE_TUNE_LOG set, every variant's time is printed to
 * stderr. */
int flue_autotune(int32_t op, const int32_t dims[4], const void* U, int32_t prec, int32_t reps, int32_t* best,
                  double* best_ms, void* ws, size_t ws_bytes, void* stream);
int flue_tuned_variant(int32_t op, const int32_t dims[4], int32_t prec, int32_t* variant);
int flue_tune_save(const char* path);
int flue_tune_load(const char* path);

/* ------------------------------------------------------------------ */
/* 4. ensembles: n fields with their own extents (dims is n x {NT,NX,NY,NZ}) stored     */
/*    back to back in `fields`, each in the SoA layout, one launch per operation       */
/* ------------------------------------------------------------------ */
size_t flue_ensemble_field_bytes(int32_t n, const int32_t* dims, int32_t prec);
size_t flue_ensemble_work_bytes(int32_t n, const int32_t* dims, int32_t prec);

/* field c from the seeded generator with seeds[c] (see flue_random_field) */
int flue_ensemble_random(int32_t n, const int32_t* dims, const uint64_t* seeds, double eps, void* fields,
                         int32_t prec, void* stream);

/* nsteps RK3 Wilson-flow steps of every field in place.  With measure_every > 0 the
 * plaquette sum (sum of ReTr P over the 6V plaquettes, as genPlaquette(1,4,4)) and the
 * magnetic B are written to d_plaq[m*n + c] and d_B[m*n + c] at steps
 * m*measure_every, m = 0 .. nsteps/measure_every.  use_graph != 0 captures the whole run
 * in one hipGraph and launches it (needs a non-default stream); the call returns after
 * the run completes in that case. */
int flue_ensemble_flow(int32_t n, const int32_t* dims, void* fields, int32_t prec, double eps, int32_t nsteps,
                       int32_t measure_every, double* d_plaq, double* d_B, int32_t use_graph, void* work,
                       size_t work_bytes, void* stream);

/* FLUE_OP_ENSEMBLE_FLOW for this ensemble: times one flow stage of every field per grid
 * variant (the fields are only read; work as for flue_ensemble_flow) and remembers the
 * fastest for later flue_ensemble_flow calls on ensembles of the same shape.  Waits for
 * the stream. */
int flue_ensemble_autotune(int32_t n, const int32_t* dims, const void* fields, int32_t prec, int32_t reps,
                           int32_t* best, double* best_ms, void* work, size_t work_bytes, void* stream);

#ifdef __cplusplus
}
#endif
#endif
