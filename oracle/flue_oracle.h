/*
 * flue_oracle.h -- CPU FP64 restatement of FLUE's gluonic-observable path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in flue_lat_amd/: only the test suite, the smoke check and
 * the CPU-baseline leg of the benchmark load it.  The product path never
 * links, loads or calls it.
 *
 * Conventions (all follow the reference):
 *   - gauge field `data` is complex128 in Fortran order
 *     data(t,x,y,z,mu,a,b), i.e. 0-based linear offset
 *     t + NT*(x + NX*(y + NY*(z + NZ*(mu + 4*(a + 3*b)))))
 *     (src/wloops.F90:21,60 and src/gaugeIO/ILDG_bin.f90:45-55);
 *   - dims[4] = {NT, NX, NY, NZ};
 *   - coordinates passed to genericPath-style functions are 1-based
 *     (t,x,y,z) and path entries are signed 1-based directions, mu=1 is time;
 *   - 3x3 matrices are 9 complex128 in Fortran column-major order.
 * Every function returns 0 on success or a nonzero ORC_* code.
 */
#ifndef FLUE_ORACLE_H
#define FLUE_ORACLE_H
#include <stdint.h>
#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex or_cplx;

enum { ORC_OK = 0, ORC_EINVAL = 1, ORC_EPATH = 3, ORC_ENT = 4 };

/* SU(3) primitives (src/SU3MatrixOps.f90) */
void or_matmul(or_cplx MM[9], const or_cplx A[9], const or_cplx B[9]);
void or_matdagmatdag(or_cplx MM[9], const or_cplx A[9], const or_cplx B[9]);
double or_retrace_mult(const or_cplx A[9], const or_cplx B[9]);
void or_period_coord(const int32_t coord[4], const int32_t dims[4], int32_t out[4]);

/* Observables (src/wloops.F90) */
int or_generic_path(const or_cplx* data, const int32_t dims[4], const int32_t coordBase[4],
                    const int32_t* path, int32_t npath, or_cplx out[9]);
int or_gen_plaquette(const or_cplx* data, const int32_t dims[4], int32_t muStart, int32_t muEnd,
                     int32_t nuEnd, double* sumTrP, int32_t* nP);
int or_plaquette(const or_cplx* data, const int32_t dims[4], int32_t muStart, int32_t muEnd,
                 int32_t nuEnd, double* sumTrP, int32_t* nP);
int or_polyakov(const or_cplx* data, const int32_t dims[4], double* sumTrP, int32_t* nP);
int or_clover_munu(const or_cplx* data, const int32_t dims[4], int32_t mu, int32_t nu, or_cplx* out);
int or_loop5_munu(const or_cplx* data, const int32_t dims[4], int32_t mu, int32_t nu, or_cplx* out);
int or_loop5_munu_sites(const or_cplx* data, const int32_t dims[4], int32_t mu, int32_t nu,
                        const int64_t* sites, int64_t nsites, or_cplx* out);
int or_magnetic(const or_cplx* data, const int32_t dims[4], double* B);
int or_magnetic_sum_tslices(const or_cplx* data, const int32_t dims[4], int32_t t0, int32_t t1,
                            double* sum);
int or_wmunu(const or_cplx* data, const int32_t dims[4], int32_t mu, int32_t nu, int32_t nMu,
             int32_t nNu, double* loopVal);
int or_wmunutau(const or_cplx* data, const int32_t dims[4], int32_t tau, int32_t mu, int32_t nu,
                int32_t nMu, int32_t nNu, double* loopVal);

/* Synthetic field and su(3) helpers (no reference counterpart) */
void or_su3_exp(const or_cplx X[9], or_cplx out[9]);
void or_fix_su3(or_cplx U[9]);
void or_gellmann(int32_t a, or_cplx out[9]);
int or_random_field(const int32_t dims[4], uint64_t seed, double eps, or_cplx* out);
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Gradients: D^a f = d/ds f(U_mu(x) -> exp(s T^a) U_mu(x)), T^a = -i lambda_a/2.
 * grad(t,x,y,z,mu,a,b) = sum_a T^a D^a f  (traceless anti-Hermitian). */
int or_wilson_action(const or_cplx* data, const int32_t dims[4], double beta, double* S);
int or_wilson_grad(const or_cplx* data, const int32_t dims[4], double beta, or_cplx* grad);
int or_b2_grad(const or_cplx* data, const int32_t dims[4], or_cplx* grad);
int or_fd_link(or_cplx* data, const int32_t dims[4], int32_t which, double beta,
               const int32_t coord[4], int32_t mu, int32_t a, double h, double* deriv);

/* Wilson flow, Luescher RK3 with Z = -P_TA(U * staple-sum) */
int or_wilson_flow(or_cplx* data, const int32_t dims[4], double eps, int32_t nsteps, int32_t reunit);

/* Gluon propagator and momentum-space analysis (src/gluonCalc/gluonProp.f90,
 * src/gluonAna/gpManip.f90).  Uq is the Fourier-transformed field with extents E in the
 * gauge-field order; D is complex128 over dext in Fortran order.  Where the reference's
 * result is undefined (U(-q) read past the end for q = 0, qhat(0) for mu = 0, D accumulated
 * uninitialised, Q_Average's endless loop) these restate a fixed reading, described at each
 * function; the tests pin that reading by closed forms, not against a reference answer. */
int or_mom_scalar_d(const or_cplx* Uq, const int32_t E[4], const int32_t dext[4], int32_t mu_start, double xi,
                    or_cplx* D);
int or_scalar_gluon_prop(const or_cplx* data, const int32_t dims[4], double xi, double* D);
int or_q_average(int32_t NQ, const double* q, const double* D, int32_t q_slices, double* DOUT, double* QOUT,
                 int32_t* Qcount);
int or_cone_cut(int32_t NQ, const double* Q, const double* D, const double* D4, double angle, double xi,
                double IRCut, double IRRadius, double* QOUT, double* DOUT, double* D4OUT, int32_t* qcount);

#ifdef __cplusplus
}
#endif
#endif
