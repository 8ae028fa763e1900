/*
 * flue_oracle.c -- CPU FP64 restatement of FLUE's gluonic-observable path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in flue_lat_amd/.
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg may load
 * it; the product path never does.
 *
 * Each function restates the reference algorithm it cites (file:line in the
 * FLUE_Lat sources), keeping its loop order, its path tables and its quirks
 * (genericPath returns the transpose, the anti-Hermitian projection is not
 * traceless, traces are summed without the 1/3).  Parts with no reference
 * counterpart (field generator, su(3) exponential, gradients, flow) are
 * written from the mathematics and say so.
 */
#include "flue_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef or_cplx cplx;

#define M3(m, i, j) (m)[(i) + 3 * (j)]

/* ------------------------------------------------------------------ */
/* layout helpers                                                      */
/* ------------------------------------------------------------------ */
static inline int64_t vol4(const int32_t d[4]) { return (int64_t)d[0] * d[1] * d[2] * d[3]; }

/* 0-based site offset in Fortran order (t fastest) */
static inline int64_t site0(const int32_t d[4], int t, int x, int y, int z) {
    return t + (int64_t)d[0] * (x + (int64_t)d[1] * (y + (int64_t)d[2] * z));
}

/* load link U_mu(coord) with 1-based coord and 1-based mu as a Fortran 3x3 */
static inline void load_link(const cplx* data, const int32_t d[4], const int32_t c[4], int mu, cplx out[9]) {
    int64_t V = vol4(d);
    int64_t s = site0(d, c[0] - 1, c[1] - 1, c[2] - 1, c[3] - 1);
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
            M3(out, a, b) = data[s + V * ((mu - 1) + 4 * (a + 3 * b))];
}

static inline void store_link(cplx* data, const int32_t d[4], const int32_t c[4], int mu, const cplx in[9]) {
    int64_t V = vol4(d);
    int64_t s = site0(d, c[0] - 1, c[1] - 1, c[2] - 1, c[3] - 1);
    for (int b = 0; b < 3; ++b)
        for (int a = 0; a < 3; ++a)
            data[s + V * ((mu - 1) + 4 * (a + 3 * b))] = M3(in, a, b);
}

static inline void ident(cplx m[9]) {
    for (int i = 0; i < 9; ++i) m[i] = 0.0;
    m[0] = m[4] = m[8] = 1.0;
}

static inline void dagger(cplx o[9], const cplx m[9]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M3(o, i, j) = conj(M3(m, j, i));
}

static inline void transpose(cplx o[9], const cplx m[9]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M3(o, i, j) = M3(m, j, i);
}

static inline void copy9(cplx o[9], const cplx m[9]) { memcpy(o, m, 9 * sizeof(cplx)); }

/* ------------------------------------------------------------------ */
/* SU(3) primitives                                                    */
/* ------------------------------------------------------------------ */

/* MM = A*B.  src/SU3MatrixOps.f90:19-37 */
void or_matmul(cplx MM[9], const cplx A[9], const cplx B[9]) {
    cplx r[9];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            M3(r, i, j) = M3(A, i, 0) * M3(B, 0, j) + M3(A, i, 1) * M3(B, 1, j) + M3(A, i, 2) * M3(B, 2, j);
    copy9(MM, r);
}

/* MM = A^dagger * B^dagger (what the code computes; its comment says
 * (A*B)^dagger).  src/SU3MatrixOps.f90:39-58 */
void or_matdagmatdag(cplx MM[9], const cplx A[9], const cplx B[9]) {
    cplx r[9];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            M3(r, i, j) = conj(M3(A, 0, i) * M3(B, j, 0) + M3(A, 1, i) * M3(B, j, 1) + M3(A, 2, i) * M3(B, j, 2));
    copy9(MM, r);
}

/* Re Tr(A*B).  src/SU3MatrixOps.f90:70-92 */
double or_retrace_mult(const cplx A[9], const cplx B[9]) {
    cplx tr = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) tr += M3(A, i, k) * M3(B, k, i);
    return creal(tr);
}

/* single-step periodic wrap of a 1-based coordinate.  src/wloops.F90:546-565 */
void or_period_coord(const int32_t coord[4], const int32_t dims[4], int32_t out[4]) {
    for (int c = 0; c < 4; ++c) {
        out[c] = coord[c];
        if (coord[c] > dims[c])
            out[c] = 1;
        else if (coord[c] == 0)
            out[c] = dims[c];
    }
}

static int check_dims(const int32_t d[4]) {
    if (!d) return ORC_EINVAL;
    for (int i = 0; i < 4; ++i)
        if (d[i] < 1) return ORC_EINVAL;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* genericPath: ordered product along a signed path, returned transposed */
/* src/wloops.F90:20-57                                                 */
/* ------------------------------------------------------------------ */
static void path_product(const cplx* data, const int32_t d[4], const int32_t base[4], const int32_t* path,
                         int npath, cplx U[9]) {
    int32_t coord[4] = {base[0], base[1], base[2], base[3]};
    cplx amat[9], bmat[9];
    ident(U);
    for (int pp = 0; pp < npath; ++pp) {
        copy9(bmat, U);
        int mu = path[pp];
        int amu = mu < 0 ? -mu : mu;
        if (mu < 0) {
            coord[amu - 1] -= 1; /* step back first (:41-44) */
            or_period_coord(coord, d, coord);
            cplx l[9];
            load_link(data, d, coord, amu, l);
            dagger(amat, l); /* CONJG(TRANSPOSE(.)) (:47) */
        } else {
            load_link(data, d, coord, amu, amat); /* (:49) */
            coord[amu - 1] += 1;                  /* then step (:50-51) */
            or_period_coord(coord, d, coord);
        }
        or_matmul(U, bmat, amat); /* right-multiply (:54) */
    }
}

static int check_path(const int32_t* path, int npath) {
    if (npath < 0 || (npath > 0 && !path)) return ORC_EPATH;
    for (int i = 0; i < npath; ++i)
        if (path[i] == 0 || path[i] > 4 || path[i] < -4) return ORC_EPATH;
    return ORC_OK;
}

int or_generic_path(const cplx* data, const int32_t dims[4], const int32_t coordBase[4], const int32_t* path,
                    int32_t npath, cplx out[9]) {
    if (check_dims(dims)) return ORC_EINVAL;
    if (check_path(path, npath)) return ORC_EPATH;
    for (int i = 0; i < 4; ++i)
        if (coordBase[i] < 1 || coordBase[i] > dims[i]) return ORC_EINVAL;
    cplx U[9];
    path_product(data, dims, coordBase, path, npath, U);
    transpose(out, U); /* (:56) */
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* genPlaquette: sum of ReTr genericPath(x,(mu,nu,-mu,-nu))              */
/* src/wloops.F90:59-127 (loop nest x,y,z,t with t innermost :90-98)    */
/* ------------------------------------------------------------------ */
int or_gen_plaquette(const cplx* data, const int32_t dims[4], int32_t muStart, int32_t muEnd, int32_t nuEnd,
                     double* sumTrP, int32_t* nP) {
    if (check_dims(dims)) return ORC_EINVAL;
    if (muStart < 1 || muEnd > 4 || nuEnd > 4) return ORC_EINVAL;
    double s = 0.0;
    int32_t n = 0;
    for (int mu = muStart; mu <= muEnd; ++mu)
        for (int nu = mu + 1; nu <= nuEnd; ++nu) {
            int32_t path[4] = {mu, nu, -mu, -nu}; /* (:84) */
            for (int x = 1; x <= dims[1]; ++x)
                for (int y = 1; y <= dims[2]; ++y)
                    for (int z = 1; z <= dims[3]; ++z)
                        for (int t = 1; t <= dims[0]; ++t) {
                            int32_t base[4] = {t, x, y, z};
                            cplx U[9];
                            path_product(data, dims, base, path, 4, U);
                            /* RealTraceMultMatMat(P, Ident, plaq) = Re Tr plaq (:103) */
                            s += creal(U[0] + U[4] + U[8]);
                            n += 1;
                        }
        }
    *sumTrP = s;
    *nP = n;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* plaquette: hard-coded ReTr[(U_mu(x)U_nu(x+mu)) (U_mu(x+nu)^+ U_nu(x)^+)] */
/* src/wloops.F90:915-1017                                               */
/* ------------------------------------------------------------------ */
int or_plaquette(const cplx* data, const int32_t dims[4], int32_t muStart, int32_t muEnd, int32_t nuEnd,
                 double* sumTrP, int32_t* nP) {
    if (check_dims(dims)) return ORC_EINVAL;
    if (muStart < 1 || muEnd > 4 || nuEnd > 4) return ORC_EINVAL;
    /* The reference runs this loop nest serially.  Here the (plane, x) rows run in
     * parallel, each summed serially in the reference's (y, z, t) order into its own
     * slot, and the slots are added in the reference's (plane, x) order afterwards, so
     * the result does not depend on the thread count. */
    int npl = 0, pmu[6], pnu[6];
    for (int mu = muStart; mu <= muEnd; ++mu)
        for (int nu = mu + 1; nu <= nuEnd; ++nu) {
            pmu[npl] = mu;
            pnu[npl] = nu;
            ++npl;
        }
    const int nrow = npl * dims[1];
    double* row = (double*)calloc(nrow > 0 ? (size_t)nrow : 1, sizeof(double));
    if (!row) return ORC_EINVAL;
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < nrow; ++r) {
        const int mu = pmu[r / dims[1]], nu = pnu[r / dims[1]], x = r % dims[1] + 1;
        double s = 0.0;
        for (int y = 1; y <= dims[2]; ++y)
            for (int z = 1; z <= dims[3]; ++z)
                for (int t = 1; t <= dims[0]; ++t) {
                    int32_t base[4] = {t, x, y, z}, c[4];
                    cplx Umu_x[9], Unu_xmu[9], Umu_xnu[9], Unu_x[9], A[9], B[9];
                    load_link(data, dims, base, mu, Umu_x); /* :963 */
                    memcpy(c, base, sizeof c);
                    c[mu - 1] += 1;
                    or_period_coord(c, dims, c); /* :968-974 */
                    load_link(data, dims, c, nu, Unu_xmu);
                    memcpy(c, base, sizeof c);
                    c[nu - 1] += 1;
                    or_period_coord(c, dims, c); /* :980-985 */
                    load_link(data, dims, c, mu, Umu_xnu);
                    load_link(data, dims, base, nu, Unu_x); /* :992 */
                    or_matmul(A, Umu_x, Unu_xmu);          /* :997 */
                    or_matdagmatdag(B, Umu_xnu, Unu_x);    /* :999 */
                    s += or_retrace_mult(A, B);            /* :1001,1005 */
                }
        row[r] = s;
    }
    double s = 0.0;
    for (int r = 0; r < nrow; ++r) s += row[r];
    free(row);
    *sumTrP = s;
    *nP = (int32_t)((int64_t)npl * vol4(dims));
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* polyakov: sum_{x,y,z} ReTr prod_t U_1(t,x,y,z).  src/wloops.F90:1019-1069 */
/* NT <= 2 leaves Unu_Temp undefined in the reference (:1048-1060), so it  */
/* is rejected here.                                                      */
/* ------------------------------------------------------------------ */
int or_polyakov(const cplx* data, const int32_t dims[4], double* sumTrP, int32_t* nP) {
    if (check_dims(dims)) return ORC_EINVAL;
    if (dims[0] < 3) return ORC_ENT;
    double s = 0.0;
    int32_t n = 0;
    for (int x = 1; x <= dims[1]; ++x)
        for (int y = 1; y <= dims[2]; ++y)
            for (int z = 1; z <= dims[3]; ++z) {
                int32_t c[4] = {1, x, y, z};
                cplx Ux[9], Uxt[9], Utmp[9];
                load_link(data, dims, c, 1, Ux); /* :1047-1048 */
                for (int t = 2; t <= dims[0] - 1; ++t) {
                    c[0] = t;
                    load_link(data, dims, c, 1, Uxt);
                    or_matmul(Utmp, Ux, Uxt); /* :1053 */
                    copy9(Ux, Utmp);
                }
                c[0] = dims[0];
                load_link(data, dims, c, 1, Ux);  /* :1057-1058 */
                s += or_retrace_mult(Utmp, Ux);    /* :1060 */
                n += 1;
            }
    *sumTrP = s;
    *nP = n;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* anti-Hermitian (not traceless) projection times 1/4, applied to the  */
/* transposed leaf sum.  src/wloops.F90:178-190 and :411-423             */
/* ------------------------------------------------------------------ */
static void ah_quarter(cplx thisU[9]) {
    cplx c12 = 0.5 * (M3(thisU, 0, 1) - conj(M3(thisU, 1, 0)));
    cplx c13 = 0.5 * (M3(thisU, 0, 2) - conj(M3(thisU, 2, 0)));
    cplx c23 = 0.5 * (M3(thisU, 1, 2) - conj(M3(thisU, 2, 1)));
    M3(thisU, 0, 1) = c12;
    M3(thisU, 0, 2) = c13;
    M3(thisU, 1, 2) = c23;
    M3(thisU, 1, 0) = -conj(c12);
    M3(thisU, 2, 0) = -conj(c13);
    M3(thisU, 2, 1) = -conj(c23);
    M3(thisU, 0, 0) = I * cimag(M3(thisU, 0, 0));
    M3(thisU, 1, 1) = I * cimag(M3(thisU, 1, 1));
    M3(thisU, 2, 2) = I * cimag(M3(thisU, 2, 2));
    for (int i = 0; i < 9; ++i) thisU[i] *= 0.25;
}

static void add_path_T(const cplx* data, const int32_t d[4], const int32_t base[4], const int32_t* path, int n,
                       cplx acc[9]) {
    cplx U[9];
    path_product(data, d, base, path, n, U);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M3(acc, i, j) += M3(U, j, i); /* genericPath transposes */
}

static inline void site_coord(const int32_t d[4], int64_t s, int32_t c[4]) {
    c[0] = (int32_t)(s % d[0]) + 1;
    s /= d[0];
    c[1] = (int32_t)(s % d[1]) + 1;
    s /= d[1];
    c[2] = (int32_t)(s % d[2]) + 1;
    s /= d[2];
    c[3] = (int32_t)s + 1;
}

static int check_plane(const int32_t dims[4], int mu, int nu) {
    if (check_dims(dims)) return ORC_EINVAL;
    if (mu < 1 || mu > 4 || nu < 1 || nu > 4) return ORC_EINVAL;
    return ORC_OK;
}

/* cloverLoopMuNuCoord: 4 1x1 leaves.  src/wloops.F90:129-199 */
static void clover_site(const cplx* data, const int32_t d[4], int mu, int nu, const int32_t base[4], cplx out[9]) {
    int32_t p[4][4] = {{nu, -mu, -nu, mu}, {mu, nu, -mu, -nu}, {-mu, -nu, mu, nu}, {-nu, mu, nu, -mu}}; /* :143-149 */
    cplx thisU[9] = {0};
    for (int l = 0; l < 4; ++l) add_path_T(data, d, base, p[l], 4, thisU);
    ah_quarter(thisU);
    copy9(out, thisU);
}

int or_clover_munu(const cplx* data, const int32_t dims[4], int32_t mu, int32_t nu, cplx* out) {
    if (check_plane(dims, mu, nu)) return ORC_EINVAL;
    int64_t V = vol4(dims);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < V; ++s) {
        int32_t c[4];
        cplx F[9];
        site_coord(dims, s, c);
        clover_site(data, dims, mu, nu, c, F);
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) out[s + V * (a + 3 * b)] = M3(F, a, b);
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Loop5MuNuCoord: 5-loop improved F_munu, hep-lat/0203008.              */
/* src/wloops.F90:201-432.  Coefficients :219-224, path tables :234-290, */
/* weighted sums :313-405, projection :411-423.                          */
/* ------------------------------------------------------------------ */
static void loop5_site(const cplx* data, const int32_t d[4], int mu, int nu, const int32_t base[4], cplx out[9]) {
    const double k5 = 1.0 / 180.0;
    const double k1 = (19.0 / 9.0) - 55.0 * k5;
    const double k2 = (1.0 / 36.0) - 16.0 * k5;
    const double k3 = 64.0 * k5 - (32.0 / 45.0);
    const double k4 = (1.0 / 15.0) - 6.0 * k5;
    int32_t p1x1[4][4] = {{nu, -mu, -nu, mu}, {mu, nu, -mu, -nu}, {-mu, -nu, mu, nu}, {-nu, mu, nu, -mu}};
    int32_t p2x2[4][8] = {{nu, nu, -mu, -mu, -nu, -nu, mu, mu},
                          {mu, mu, nu, nu, -mu, -mu, -nu, -nu},
                          {-mu, -mu, -nu, -nu, mu, mu, nu, nu},
                          {-nu, -nu, mu, mu, nu, nu, -mu, -mu}};
    int32_t p3x3[4][12] = {{nu, nu, nu, -mu, -mu, -mu, -nu, -nu, -nu, mu, mu, mu},
                           {mu, mu, mu, nu, nu, nu, -mu, -mu, -mu, -nu, -nu, -nu},
                           {-mu, -mu, -mu, -nu, -nu, -nu, mu, mu, mu, nu, nu, nu},
                           {-nu, -nu, -nu, mu, mu, mu, nu, nu, nu, -mu, -mu, -mu}};
    /* rows in the order the reference sums them (:362-379): 1,5,2,6,3,7,4,8 */
    int32_t p1x2[8][6] = {{nu, nu, -mu, -nu, -nu, mu},   {nu, -mu, -mu, -nu, mu, mu},
                          {mu, nu, nu, -mu, -nu, -nu},   {mu, mu, nu, -mu, -mu, -nu},
                          {-mu, -nu, -nu, mu, nu, nu},   {-mu, -mu, -nu, mu, mu, nu},
                          {-nu, -nu, mu, nu, nu, -mu},   {-nu, mu, mu, nu, -mu, -mu}};
    int32_t p1x3[8][8] = {{nu, nu, nu, -mu, -nu, -nu, -nu, mu},     {nu, -mu, -mu, -mu, -nu, mu, mu, mu},
                          {mu, nu, nu, nu, -mu, -nu, -nu, -nu},     {mu, mu, mu, nu, -mu, -mu, -mu, -nu},
                          {-mu, -nu, -nu, -nu, mu, nu, nu, nu},     {-mu, -mu, -mu, -nu, mu, mu, mu, nu},
                          {-nu, -nu, -nu, mu, nu, nu, nu, -mu},     {-nu, mu, mu, mu, nu, -mu, -mu, -mu}};
    cplx thisU[9] = {0}, workU[9];
    for (int l = 0; l < 4; ++l) add_path_T(data, d, base, p1x1[l], 4, thisU);
    for (int i = 0; i < 9; ++i) thisU[i] *= k1; /* :327 */
    memset(workU, 0, sizeof workU);
    for (int l = 0; l < 4; ++l) add_path_T(data, d, base, p2x2[l], 8, workU);
    for (int i = 0; i < 9; ++i) thisU[i] += workU[i] * k2; /* :341 */
    memset(workU, 0, sizeof workU);
    for (int l = 0; l < 4; ++l) add_path_T(data, d, base, p3x3[l], 12, workU);
    for (int i = 0; i < 9; ++i) thisU[i] += workU[i] * k5; /* :358 */
    memset(workU, 0, sizeof workU);
    for (int l = 0; l < 8; ++l) add_path_T(data, d, base, p1x2[l], 6, workU);
    for (int i = 0; i < 9; ++i) thisU[i] += workU[i] * k3 * 0.5; /* :382 */
    memset(workU, 0, sizeof workU);
    for (int l = 0; l < 8; ++l) add_path_T(data, d, base, p1x3[l], 8, workU);
    for (int i = 0; i < 9; ++i) thisU[i] += workU[i] * k4 * 0.5; /* :405 */
    ah_quarter(thisU);                                          /* :411-423 */
    copy9(out, thisU);
}

int or_loop5_munu(const cplx* data, const int32_t dims[4], int32_t mu, int32_t nu, cplx* out) {
    if (check_plane(dims, mu, nu)) return ORC_EINVAL;
    int64_t V = vol4(dims);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < V; ++s) {
        int32_t c[4];
        cplx F[9];
        site_coord(dims, s, c);
        loop5_site(data, dims, mu, nu, c, F);
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) out[s + V * (a + 3 * b)] = M3(F, a, b);
    }
    return ORC_OK;
}

/* per-site F at a list of 0-based Fortran site offsets; out is nsites x 9 (Fortran 3x3 each) */
int or_loop5_munu_sites(const cplx* data, const int32_t dims[4], int32_t mu, int32_t nu, const int64_t* sites,
                        int64_t nsites, cplx* out) {
    if (check_plane(dims, mu, nu)) return ORC_EINVAL;
    int64_t V = vol4(dims);
    for (int64_t i = 0; i < nsites; ++i)
        if (sites[i] < 0 || sites[i] >= V) return ORC_EINVAL;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nsites; ++i) {
        int32_t c[4];
        site_coord(dims, sites[i], c);
        loop5_site(data, dims, mu, nu, c, out + 9 * i);
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* magnetic: B = -(1/V) sum_x sum_i ReTr(F_i F_i) with F_1=Loop5(3,2),   */
/* F_2=Loop5(2,4), F_3=Loop5(4,3).  src/wloops.F90:479-544               */
/* ------------------------------------------------------------------ */
static const int32_t kMagPlanes[3][2] = {{3, 2}, {2, 4}, {4, 3}}; /* :502-506 */

static double magnetic_site(const cplx* data, const int32_t dims[4], const int32_t c[4]) {
    double acc = 0.0;
    for (int i = 0; i < 3; ++i) {
        cplx F[9];
        loop5_site(data, dims, kMagPlanes[i][0], kMagPlanes[i][1], c, F);
        acc -= or_retrace_mult(F, F); /* :530-533 */
    }
    return acc;
}

int or_magnetic_sum_tslices(const cplx* data, const int32_t dims[4], int32_t t0, int32_t t1, double* sum) {
    if (check_dims(dims)) return ORC_EINVAL;
    if (t0 < 1 || t1 > dims[0] || t0 > t1) return ORC_EINVAL;
    int64_t Vs = (int64_t)dims[1] * dims[2] * dims[3];
    int64_t nt = t1 - t0 + 1;
    int64_t n = Vs * nt;
    double* per = (double*)malloc(sizeof(double) * n);
    if (!per) return ORC_EINVAL;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        int64_t sp = i / nt;
        int32_t c[4];
        c[0] = t0 + (int32_t)(i % nt);
        c[1] = (int32_t)(sp % dims[1]) + 1;
        c[2] = (int32_t)((sp / dims[1]) % dims[2]) + 1;
        c[3] = (int32_t)(sp / ((int64_t)dims[1] * dims[2])) + 1;
        per[i] = magnetic_site(data, dims, c);
    }
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += per[i];
    free(per);
    *sum = s;
    return ORC_OK;
}

int or_magnetic(const cplx* data, const int32_t dims[4], double* B) {
    double s;
    int rc = or_magnetic_sum_tslices(data, dims, 1, dims[0], &s);
    if (rc) return rc;
    *B = s / (double)vol4(dims); /* :543 */
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* wmunu / wmunutau: nMu x nNu rectangular loop ReTr[A B C^+ D^+]         */
/* src/wloops.F90:679-791 and :567-677                                   */
/* ------------------------------------------------------------------ */
static double rect_loop(const cplx* data, const int32_t d[4], const int32_t base[4], int mu, int nu, int nMu,
                        int nNu) {
    int32_t c[4];
    cplx Umu_x[9], Unu_xpamuh[9], Umu_xpanuh[9], Unu_x[9], Ut[9], A[9], B[9];
    /* U_nMu(x) :717-730 */
    memcpy(c, base, sizeof c);
    ident(Umu_x);
    for (int n = 0; n < nMu; ++n) {
        load_link(data, d, c, mu, Ut);
        or_matmul(Umu_x, Umu_x, Ut);
        c[mu - 1] += 1;
        or_period_coord(c, d, c);
    }
    /* U_nNu(x + nMu mu) :731-747 */
    memcpy(c, base, sizeof c);
    for (int n = 0; n < nMu; ++n) {
        c[mu - 1] += 1;
        or_period_coord(c, d, c);
    }
    ident(Unu_xpamuh);
    for (int n = 0; n < nNu; ++n) {
        load_link(data, d, c, nu, Ut);
        or_matmul(Unu_xpamuh, Unu_xpamuh, Ut);
        c[nu - 1] += 1;
        or_period_coord(c, d, c);
    }
    /* U_nMu(x + nNu nu) :748-764 */
    memcpy(c, base, sizeof c);
    for (int n = 0; n < nNu; ++n) {
        c[nu - 1] += 1;
        or_period_coord(c, d, c);
    }
    ident(Umu_xpanuh);
    for (int n = 0; n < nMu; ++n) {
        load_link(data, d, c, mu, Ut);
        or_matmul(Umu_xpanuh, Umu_xpanuh, Ut);
        c[mu - 1] += 1;
        or_period_coord(c, d, c);
    }
    /* U_nNu(x) :765-778 */
    memcpy(c, base, sizeof c);
    ident(Unu_x);
    for (int n = 0; n < nNu; ++n) {
        load_link(data, d, c, nu, Ut);
        or_matmul(Unu_x, Unu_x, Ut);
        c[nu - 1] += 1;
        or_period_coord(c, d, c);
    }
    or_matmul(A, Umu_x, Unu_xpamuh);         /* :780 */
    or_matdagmatdag(B, Umu_xpanuh, Unu_x);   /* :782 */
    return or_retrace_mult(A, B);            /* :784 */
}

static int check_rect(const int32_t dims[4], int mu, int nu, int nMu, int nNu) {
    if (check_plane(dims, mu, nu)) return ORC_EINVAL;
    if (nMu < 1 || nNu < 1) return ORC_EINVAL;
    return ORC_OK;
}

int or_wmunu(const cplx* data, const int32_t dims[4], int32_t mu, int32_t nu, int32_t nMu, int32_t nNu,
             double* loopVal) {
    if (check_rect(dims, mu, nu, nMu, nNu)) return ORC_EINVAL;
    double s = 0.0;
#pragma omp parallel for collapse(3) reduction(+ : s) schedule(static)
    for (int x = 1; x <= dims[1]; ++x)
        for (int y = 1; y <= dims[2]; ++y)
            for (int z = 1; z <= dims[3]; ++z)
                for (int t = 1; t <= dims[0]; ++t) {
                    int32_t base[4] = {t, x, y, z};
                    s += rect_loop(data, dims, base, mu, nu, nMu, nNu);
                }
    *loopVal = s / (double)vol4(dims); /* SIZE(data)/36 (:790) */
    return ORC_OK;
}

int or_wmunutau(const cplx* data, const int32_t dims[4], int32_t tau, int32_t mu, int32_t nu, int32_t nMu,
                int32_t nNu, double* loopVal) {
    if (check_rect(dims, mu, nu, nMu, nNu)) return ORC_EINVAL;
    if (tau < 1 || tau > dims[0]) return ORC_EINVAL;
    double s = 0.0;
#pragma omp parallel for collapse(3) reduction(+ : s) schedule(static)
    for (int x = 1; x <= dims[1]; ++x)
        for (int y = 1; y <= dims[2]; ++y)
            for (int z = 1; z <= dims[3]; ++z) {
                int32_t base[4] = {tau, x, y, z};
                s += rect_loop(data, dims, base, mu, nu, nMu, nNu);
            }
    *loopVal = s / (double)((int64_t)dims[1] * dims[2] * dims[3]); /* :676 */
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* su(3) helpers -- no reference counterpart                            */
/* ------------------------------------------------------------------ */

/* Gell-Mann matrices lambda_1..lambda_8 (textbook), a = 1..8 */
void or_gellmann(int32_t a, cplx m[9]) {
    for (int i = 0; i < 9; ++i) m[i] = 0.0;
    const double r3 = 1.0 / sqrt(3.0);
    switch (a) {
    case 1: M3(m, 0, 1) = 1; M3(m, 1, 0) = 1; break;
    case 2: M3(m, 0, 1) = -I; M3(m, 1, 0) = I; break;
    case 3: M3(m, 0, 0) = 1; M3(m, 1, 1) = -1; break;
    case 4: M3(m, 0, 2) = 1; M3(m, 2, 0) = 1; break;
    case 5: M3(m, 0, 2) = -I; M3(m, 2, 0) = I; break;
    case 6: M3(m, 1, 2) = 1; M3(m, 2, 1) = 1; break;
    case 7: M3(m, 1, 2) = -I; M3(m, 2, 1) = I; break;
    case 8: M3(m, 0, 0) = r3; M3(m, 1, 1) = r3; M3(m, 2, 2) = -2 * r3; break;
    default: break;
    }
}

/* exp(X) by scaling and squaring of a degree-16 Taylor polynomial (Horner). */
void or_su3_exp(const cplx X[9], cplx out[9]) {
    double nrm = 0.0;
    for (int i = 0; i < 9; ++i) nrm += creal(X[i] * conj(X[i]));
    nrm = sqrt(nrm);
    int s = 0;
    while (nrm > 0.125 && s < 60) {
        nrm *= 0.5;
        ++s;
    }
    double scale = ldexp(1.0, -s);
    cplx Y[9], E[9], T[9];
    for (int i = 0; i < 9; ++i) Y[i] = X[i] * scale;
    ident(E);
    for (int k = 16; k >= 1; --k) {
        or_matmul(T, Y, E);
        ident(E);
        for (int i = 0; i < 9; ++i) E[i] += T[i] / (double)k;
    }
    for (int i = 0; i < s; ++i) or_matmul(E, E, E);
    copy9(out, E);
}

/* Gram-Schmidt on rows, third row = conj(row1 x row2).
 * Follows FixSU3Matrix, src/gaugeIO/ILDG_bin.f90:108-159. */
static void normalise3(cplx v[3]) {
    double n = sqrt(creal(v[0] * conj(v[0])) + creal(v[1] * conj(v[1])) + creal(v[2] * conj(v[2])));
    v[0] /= n;
    v[1] /= n;
    v[2] /= n;
}
void or_fix_su3(cplx U[9]) {
    cplx v1[3], v2[3], v3[3];
    for (int j = 0; j < 3; ++j) {
        v1[j] = M3(U, 0, j);
        v2[j] = M3(U, 1, j);
    }
    normalise3(v1); /* :142 */
    cplx vdotw = conj(v1[0]) * v2[0] + conj(v1[1]) * v2[1] + conj(v1[2]) * v2[2]; /* :114 */
    for (int j = 0; j < 3; ++j) v2[j] -= v1[j] * vdotw;                            /* :116 */
    normalise3(v2);
    for (int ic = 0; ic < 3; ++ic) { /* :128-133 */
        int jc = (ic + 1) % 3, kc = (jc + 1) % 3;
        v3[ic] = conj(v1[jc] * v2[kc] - v1[kc] * v2[jc]);
    }
    normalise3(v3);
    for (int j = 0; j < 3; ++j) {
        M3(U, 0, j) = v1[j];
        M3(U, 1, j) = v2[j];
        M3(U, 2, j) = v3[j];
    }
}

/* Philox4x32-10 (Salmon et al., SC'11) */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, &hi0, &lo0);
        mulhilo32(0xCD9E8D57u, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 8 Gaussians per link, keyed by (seed, link id, block) -- order independent */
static void link_gaussians(uint64_t seed, uint64_t link, double g[8]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int blk = 0; blk < 2; ++blk) {
        uint32_t ctr[4] = {(uint32_t)link, (uint32_t)(link >> 32), (uint32_t)blk, 0x464C5545u}, r[4];
        or_philox4x32(ctr, key, r);
        for (int p = 0; p < 2; ++p) {
            double u1 = ((double)r[2 * p] + 0.5) * 0x1p-32;
            double u2 = ((double)r[2 * p + 1] + 0.5) * 0x1p-32;
            double rad = sqrt(-2.0 * log(u1));
            double ang = 6.283185307179586476925286766559 * u2;
            g[4 * blk + 2 * p] = rad * cos(ang);
            g[4 * blk + 2 * p + 1] = rad * sin(ang);
        }
    }
}

/* U = fix(exp(i eps sum_a c_a lambda_a / 2)), c_a ~ N(0,1), written in Fortran order */
int or_random_field(const int32_t dims[4], uint64_t seed, double eps, cplx* out) {
    if (check_dims(dims)) return ORC_EINVAL;
    int64_t V = vol4(dims);
#pragma omp parallel for schedule(static)
    for (int64_t L = 0; L < 4 * V; ++L) {
        double g[8];
        link_gaussians(seed, (uint64_t)L, g);
        cplx X[9] = {0}, lam[9], U[9];
        for (int a = 0; a < 8; ++a) {
            or_gellmann(a + 1, lam);
            for (int i = 0; i < 9; ++i) X[i] += (I * eps * 0.5 * g[a]) * lam[i];
        }
        or_su3_exp(X, U);
        or_fix_su3(U);
        int64_t s = L % V, mu = L / V;
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a) out[s + V * (mu + 4 * (a + 3 * b))] = M3(U, a, b);
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Gradients -- no reference counterpart; written from the mathematics.  */
/* ------------------------------------------------------------------ */

/* P_TA(M) = (M - M^+)/2 - Tr(M - M^+)/6 * 1 */
static void proj_ta(cplx o[9], const cplx M[9]) {
    cplx t[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M3(t, i, j) = 0.5 * (M3(M, i, j) - conj(M3(M, j, i)));
    cplx tr = (M3(t, 0, 0) + M3(t, 1, 1) + M3(t, 2, 2)) / 3.0;
    M3(t, 0, 0) -= tr;
    M3(t, 1, 1) -= tr;
    M3(t, 2, 2) -= tr;
    copy9(o, t);
}

/* S = beta * sum_x sum_{mu<nu} (1 - ReTr P_munu(x)/3) */
int or_wilson_action(const cplx* data, const int32_t dims[4], double beta, double* S) {
    double s;
    int32_t n;
    int rc = or_plaquette(data, dims, 1, 4, 4, &s, &n);
    if (rc) return rc;
    *S = beta * ((double)n - s / 3.0);
    return ORC_OK;
}

/* staple sum Sigma with U_mu(x) * Sigma = sum of the 6 plaquettes through the link, 1-based coords */
static void staple_sum(const cplx* data, const int32_t d[4], const int32_t c[4], int mu, cplx St[9]) {
    for (int i = 0; i < 9; ++i) St[i] = 0.0;
    for (int nu = 1; nu <= 4; ++nu) {
        if (nu == mu) continue;
        /* up: U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ */
        int32_t up[3] = {nu, -mu, -nu};
        int32_t cm[4];
        memcpy(cm, c, sizeof cm);
        cm[mu - 1] += 1;
        or_period_coord(cm, d, cm);
        cplx P[9];
        path_product(data, d, cm, up, 3, P);
        for (int i = 0; i < 9; ++i) St[i] += P[i];
        /* down: U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) */
        int32_t dn[3] = {-nu, -mu, nu};
        path_product(data, d, cm, dn, 3, P);
        for (int i = 0; i < 9; ++i) St[i] += P[i];
    }
}

/* dS/dU_mu(x) = (beta/6) P_TA(U_mu(x) Sigma_mu(x)) */
int or_wilson_grad(const cplx* data, const int32_t dims[4], double beta, cplx* grad) {
    if (check_dims(dims)) return ORC_EINVAL;
    int64_t V = vol4(dims);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < V; ++s) {
        int32_t c[4];
        site_coord(dims, s, c);
        for (int mu = 1; mu <= 4; ++mu) {
            cplx U[9], St[9], M[9], G[9];
            load_link(data, dims, c, mu, U);
            staple_sum(data, dims, c, mu, St);
            or_matmul(M, U, St);
            proj_ta(G, M);
            for (int i = 0; i < 9; ++i) G[i] *= beta / 6.0;
            store_link(grad, dims, c, mu, G);
        }
    }
    return ORC_OK;
}

/* The 28 leaves of Loop5 for plane (mu,nu) with their weights: lengths and paths as in
 * src/wloops.F90:234-290, weights k1,k2,k5,k3/2,k4/2 (:327-405). */
typedef struct {
    int n;
    int32_t p[12];
    double k;
} leaf_t;

static int loop5_leaves(int mu, int nu, leaf_t L[28]) {
    const double k5 = 1.0 / 180.0, k1 = (19.0 / 9.0) - 55.0 * k5, k2 = (1.0 / 36.0) - 16.0 * k5,
                 k3 = 64.0 * k5 - (32.0 / 45.0), k4 = (1.0 / 15.0) - 6.0 * k5;
    int32_t s[4][2] = {{nu, -mu}, {mu, nu}, {-mu, -nu}, {-nu, mu}}; /* first two directions of each quadrant */
    int idx = 0;
    /* rectangle a (steps along d1) x b (steps along d2), CCW: d1^a d2^b (-d1)^a (-d2)^b */
    int shapes[7][2] = {{1, 1}, {2, 2}, {3, 3}, {1, 2}, {2, 1}, {1, 3}, {3, 1}};
    double ks[7] = {k1, k2, k5, 0.5 * k3, 0.5 * k3, 0.5 * k4, 0.5 * k4};
    for (int sh = 0; sh < 7; ++sh)
        for (int q = 0; q < 4; ++q) {
            int d1 = s[q][0], d2 = s[q][1], a = shapes[sh][0], b = shapes[sh][1];
            leaf_t* l = &L[idx++];
            l->k = ks[sh];
            l->n = 0;
            for (int i = 0; i < a; ++i) l->p[l->n++] = d1;
            for (int i = 0; i < b; ++i) l->p[l->n++] = d2;
            for (int i = 0; i < a; ++i) l->p[l->n++] = -d1;
            for (int i = 0; i < b; ++i) l->p[l->n++] = -d2;
        }
    return idx;
}

/* walk a path from base collecting (coord, |mu|, dagger) per step and the link matrices */
static void walk(const cplx* data, const int32_t d[4], const int32_t base[4], const int32_t* path, int n,
                 int32_t coords[][4], int* dirs, int* dag, cplx mats[][9]) {
    int32_t c[4] = {base[0], base[1], base[2], base[3]};
    for (int j = 0; j < n; ++j) {
        int mu = path[j], amu = mu < 0 ? -mu : mu;
        if (mu < 0) {
            c[amu - 1] -= 1;
            or_period_coord(c, d, c);
            memcpy(coords[j], c, sizeof c);
            cplx l[9];
            load_link(data, d, c, amu, l);
            dagger(mats[j], l);
            dag[j] = 1;
        } else {
            memcpy(coords[j], c, sizeof c);
            load_link(data, d, c, amu, mats[j]);
            c[amu - 1] += 1;
            or_period_coord(c, d, c);
            dag[j] = 0;
        }
        dirs[j] = amu;
    }
}

/* Gradient of B (magnetic) by scattering every leaf occurrence to its link.
 * B = -(1/(16V)) sum_x sum_i ReTr(A_i(x)^2), A = (Q - Q^+)/2 with Q the untransposed
 * weighted leaf sum; dB = -(1/(8V)) sum ReTr(A dQ).  For a leaf W = P U S:
 * forward link contributes +k U S A P, a daggered link -k S A P U^+, and
 * grad = (1/(16V)) P_TA(sum). */
int or_b2_grad(const cplx* data, const int32_t dims[4], cplx* grad) {
    if (check_dims(dims)) return ORC_EINVAL;
    int64_t V = vol4(dims);
    cplx* acc = (cplx*)calloc((size_t)(36 * V), sizeof(cplx));
    if (!acc) return ORC_EINVAL;
    for (int64_t s = 0; s < V; ++s) {
        int32_t c[4];
        site_coord(dims, s, c);
        for (int pl = 0; pl < 3; ++pl) {
            leaf_t L[28];
            loop5_leaves(kMagPlanes[pl][0], kMagPlanes[pl][1], L);
            cplx Q[9] = {0}, A[9];
            for (int l = 0; l < 28; ++l) {
                cplx W[9];
                path_product(data, dims, c, L[l].p, L[l].n, W);
                for (int i = 0; i < 9; ++i) Q[i] += L[l].k * W[i];
            }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) M3(A, i, j) = 0.5 * (M3(Q, i, j) - conj(M3(Q, j, i)));
            for (int l = 0; l < 28; ++l) {
                int n = L[l].n;
                int32_t coords[12][4];
                int dirs[12], dag[12];
                cplx mats[12][9], pre[13][9], suf[13][9];
                walk(data, dims, c, L[l].p, n, coords, dirs, dag, mats);
                ident(pre[0]);
                for (int j = 0; j < n; ++j) or_matmul(pre[j + 1], pre[j], mats[j]);
                ident(suf[n]);
                for (int j = n - 1; j >= 0; --j) or_matmul(suf[j], mats[j], suf[j + 1]);
                for (int j = 0; j < n; ++j) {
                    cplx SA[9], SAP[9], M[9];
                    or_matmul(SA, suf[j + 1], A);
                    or_matmul(SAP, SA, pre[j]);
                    double sign = 1.0;
                    if (!dag[j]) {
                        or_matmul(M, mats[j], SAP); /* U S A P */
                    } else {
                        or_matmul(M, SAP, mats[j]); /* S A P U^+ (mats holds U^+) */
                        sign = -1.0;
                    }
                    int64_t ls = site0(dims, coords[j][0] - 1, coords[j][1] - 1, coords[j][2] - 1, coords[j][3] - 1);
                    cplx* dst = acc + 36 * ls + 9 * (dirs[j] - 1);
                    for (int i = 0; i < 9; ++i) dst[i] += sign * L[l].k * M[i];
                }
            }
        }
    }
    for (int64_t s = 0; s < V; ++s)
        for (int mu = 0; mu < 4; ++mu) {
            cplx G[9];
            proj_ta(G, acc + 36 * s + 9 * mu);
            for (int i = 0; i < 9; ++i) G[i] /= (16.0 * (double)V);
            for (int b = 0; b < 3; ++b)
                for (int a = 0; a < 3; ++a) grad[s + V * (mu + 4 * (a + 3 * b))] = M3(G, a, b);
        }
    free(acc);
    return ORC_OK;
}

/* local parts of the observables around a link, for finite differences */
static double local_wilson(const cplx* data, const int32_t d[4], const int32_t c[4], int mu, double beta) {
    /* the 6 plaquettes through U_mu(c): ReTr(U Sigma) */
    cplx U[9], St[9];
    load_link(data, d, c, mu, U);
    staple_sum(data, d, c, mu, St);
    return -beta / 3.0 * or_retrace_mult(U, St);
}

/* -ReTr F^2 (three magnetic planes) at each distinct site within 3 of c in each spatial
 * direction at the same t, written to vals in a fixed order; returns the count */
static int local_b2_values(const cplx* data, const int32_t d[4], const int32_t c[4], double* vals) {
    int64_t Vs = (int64_t)d[1] * d[2] * d[3];
    unsigned char* seen = (unsigned char*)calloc((size_t)Vs, 1);
    int n = 0;
    for (int dz = -3; dz <= 3; ++dz)
        for (int dy = -3; dy <= 3; ++dy)
            for (int dx = -3; dx <= 3; ++dx) {
                int32_t e[4];
                e[0] = c[0];
                e[1] = ((c[1] - 1 + dx) % d[1] + d[1]) % d[1] + 1;
                e[2] = ((c[2] - 1 + dy) % d[2] + d[2]) % d[2] + 1;
                e[3] = ((c[3] - 1 + dz) % d[3] + d[3]) % d[3] + 1;
                int64_t sp = (e[1] - 1) + (int64_t)d[1] * ((e[2] - 1) + (int64_t)d[2] * (e[3] - 1));
                if (seen[sp]) continue;
                seen[sp] = 1;
                vals[n++] = magnetic_site(data, d, e);
            }
    free(seen);
    return n;
}

/* central finite difference of observable `which` (0: Wilson action with beta, 1: B) along
 * exp(+-h T^a) U_mu(coord), coord 1-based, mu 1-based, a = 1..8.  The link is perturbed in
 * place and restored, so `data` must be writable. */
int or_fd_link(cplx* data, const int32_t dims[4], int32_t which, double beta, const int32_t coord[4],
               int32_t mu, int32_t a, double h, double* deriv) {
    if (check_dims(dims) || mu < 1 || mu > 4 || a < 1 || a > 8) return ORC_EINVAL;
    int64_t V = vol4(dims);
    cplx U0[9], lam[9], X[9], E[9], U1[9];
    load_link(data, dims, coord, mu, U0);
    or_gellmann(a, lam);
    /* B is differenced site by site before summing, so the result is not limited by the
     * rounding of the whole local sum */
    double val[2], b2v[2][343];
    int nb = 0;
    for (int sgn = 0; sgn < 2; ++sgn) {
        double hh = sgn == 0 ? h : -h;
        for (int i = 0; i < 9; ++i) X[i] = -I * 0.5 * hh * lam[i]; /* h T^a */
        or_su3_exp(X, E);
        or_matmul(U1, E, U0);
        store_link(data, dims, coord, mu, U1);
        if (which == 0)
            val[sgn] = local_wilson(data, dims, coord, mu, beta);
        else
            nb = local_b2_values(data, dims, coord, b2v[sgn]);
    }
    store_link(data, dims, coord, mu, U0);
    if (which == 0) {
        *deriv = (val[0] - val[1]) / (2.0 * h);
    } else {
        double s = 0.0;
        for (int i = 0; i < nb; ++i) s += b2v[0][i] - b2v[1][i];
        *deriv = s / (double)V / (2.0 * h);
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Wilson flow (Luescher, JHEP 2010, RK3 scheme of eqs. (C.1)-(C.3))     */
/* ------------------------------------------------------------------ */
static void flow_force(const cplx* data, const int32_t d[4], cplx* Z) {
    /* Z_mu(x) = -P_TA(U_mu(x) Sigma_mu(x)), stored as 36*V (site-major) */
    int64_t V = vol4(d);
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < V; ++s) {
        int32_t c[4];
        site_coord(d, s, c);
        for (int mu = 1; mu <= 4; ++mu) {
            cplx U[9], St[9], M[9], G[9];
            load_link(data, d, c, mu, U);
            staple_sum(data, d, c, mu, St);
            or_matmul(M, U, St);
            proj_ta(G, M);
            for (int i = 0; i < 9; ++i) Z[36 * s + 9 * (mu - 1) + i] = -G[i];
        }
    }
}

int or_wilson_flow(cplx* data, const int32_t dims[4], double eps, int32_t nsteps, int32_t reunit) {
    if (check_dims(dims) || nsteps < 0) return ORC_EINVAL;
    int64_t V = vol4(dims);
    cplx* Z = (cplx*)malloc(sizeof(cplx) * 36 * (size_t)V);
    cplx* acc = (cplx*)calloc((size_t)36 * (size_t)V, sizeof(cplx));
    if (!Z || !acc) {
        free(Z);
        free(acc);
        return ORC_EINVAL;
    }
    /* stage coefficients: X_1 = 1/4 Z0; X_2 = 8/9 Z1 - 17/36 Z0; X_3 = 3/4 Z2 - 8/9 Z1 + 17/36 Z0,
     * written as acc <- a*acc + b*Z_i, W <- exp(acc) W with (a,b) = (0,1/4), (-17/9, 8/9), (-1, 3/4) */
    const double ca[3] = {0.0, -17.0 / 9.0, -1.0};
    const double cb[3] = {0.25, 8.0 / 9.0, 0.75};
    for (int step = 0; step < nsteps; ++step) {
        for (int st = 0; st < 3; ++st) {
            flow_force(data, dims, Z);
#pragma omp parallel for schedule(static)
            for (int64_t s = 0; s < V; ++s) {
                int32_t c[4];
                site_coord(dims, s, c);
                for (int mu = 1; mu <= 4; ++mu) {
                    cplx* A = acc + 36 * s + 9 * (mu - 1);
                    const cplx* z = Z + 36 * s + 9 * (mu - 1);
                    /* stage 0 starts a fresh accumulator (never scale what is in it) */
                    for (int i = 0; i < 9; ++i) A[i] = (st == 0 ? 0.0 : ca[st] * A[i]) + cb[st] * eps * z[i];
                    cplx E[9], U[9], W[9];
                    or_su3_exp(A, E);
                    load_link(data, dims, c, mu, U);
                    or_matmul(W, E, U);
                    if (reunit && st == 2) or_fix_su3(W);
                    store_link(data, dims, c, mu, W);
                }
            }
        }
    }
    free(Z);
    free(acc);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Gluon propagator and its momentum-space analysis                    */
/* (src/gluonCalc/gluonProp.f90, src/gluonAna/gpManip.f90)             */
/* ------------------------------------------------------------------ */

/* TrSub = left - right^+ - Tr(left - right^+)/3, the trace taken as Tr(TrSub * Ident) and
 * the third subtracted from all nine entries, as the array expression of
 * src/SU3MatrixOps.f90:94-105 does */
static void traceless_conjg_subtract(cplx o[9], const cplx l[9], const cplx r[9]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M3(o, i, j) = M3(l, i, j) - conj(M3(r, j, i));
    const cplx tr = (M3(o, 0, 0) + M3(o, 1, 1)) + M3(o, 2, 2);
    for (int i = 0; i < 9; ++i) o[i] = o[i] - tr / 3.0;
}

/* Gell-Mann components, src/SU3MatrixOps.f90:107-122 (com(3) keeps the imaginary part,
 * com(8) has no 1/sqrt(3) on its first two terms) */
static void colour_decomp(cplx com[8], const cplx A[9]) {
    com[0] = creal(M3(A, 0, 1)) + creal(M3(A, 1, 0));
    com[1] = -cimag(M3(A, 0, 1)) + cimag(M3(A, 1, 0));
    com[2] = M3(A, 0, 0) - M3(A, 1, 1);
    com[3] = creal(M3(A, 0, 2)) + creal(M3(A, 2, 0));
    com[4] = -cimag(M3(A, 0, 2)) + cimag(M3(A, 2, 0));
    com[5] = creal(M3(A, 1, 2)) + creal(M3(A, 2, 1));
    com[6] = -cimag(M3(A, 1, 2)) + cimag(M3(A, 2, 1));
    com[7] = creal(M3(A, 0, 0)) + creal(M3(A, 1, 1)) - (2.0 / sqrt(3.0)) * creal(M3(A, 2, 2));
}

/* calc_mom_space_scalarD, src/gluonCalc/gluonProp.f90:112-250, over the momenta
 * 0 <= q < dext of the Fourier-transformed field Uq (extents E, Fortran order like a gauge
 * field).  Readings where the reference's result is undefined:
 *   - U(-q) is taken at (E - q) mod E; the reference reads index E - q (:177), one past
 *     the dimension's end when a component of q is 0;
 *   - the phase of direction mu (0-based, :185-186) uses qhat(mu) with qhat 1-based, as
 *     the reference does, i.e. the momentum component of direction mu - 1; for mu = 0 the
 *     reference reads qhat(0), outside the array, and this uses the t component.
 * The prefactor keeps the reference's conditions (:216, :225), including qx counted twice. */
int or_mom_scalar_d(const cplx* Uq, const int32_t E[4], const int32_t dext[4], int32_t mu_start, double xi,
                    cplx* D) {
    if (check_dims(E) || check_dims(dext) || mu_start < 0) return ORC_EINVAL;
    for (int i = 0; i < 4; ++i)
        if (dext[i] > E[i] / 2 + 1) return ORC_EINVAL;
    const int64_t V = vol4(E);
    const int64_t nD = vol4(dext);
#pragma omp parallel for schedule(static)
    for (int64_t o = 0; o < nD; ++o) {
        int32_t q[4];
        int64_t r = o;
        for (int i = 0; i < 4; ++i) {
            q[i] = (int32_t)(r % dext[i]);
            r /= dext[i];
        }
        double qhat[4];
        for (int i = 0; i < 4; ++i) qhat[i] = (2.0 * M_PI / 1.0) * (double)q[i] / (double)E[i];
        int32_t qn[4];
        for (int i = 0; i < 4; ++i) qn[i] = (E[i] - q[i]) % E[i];
        const int64_t sp = site0(E, q[0], q[1], q[2], q[3]);
        const int64_t sn = site0(E, qn[0], qn[1], qn[2], qn[3]);
        cplx acc = 0.0;
        for (int mu = mu_start; mu <= 3; ++mu) {
            cplx Up[9], Un[9], Ap[9], An[9];
            for (int b = 0; b < 3; ++b)
                for (int a = 0; a < 3; ++a) {
                    M3(Up, a, b) = Uq[sp + V * (mu + 4 * (a + 3 * b))];
                    M3(Un, a, b) = Uq[sn + V * (mu + 4 * (a + 3 * b))];
                }
            traceless_conjg_subtract(Ap, Up, Un);
            traceless_conjg_subtract(An, Un, Up);
            const double th = qhat[mu >= 1 ? mu - 1 : 0];
            const cplx pp = (0.5 * I) * cexp(-I * 0.5 * th);
            const cplx pn = (0.5 * I) * cexp(I * 0.5 * th);
            for (int i = 0; i < 9; ++i) {
                Ap[i] = pp * Ap[i];
                An[i] = pn * An[i];
            }
            cplx cp[8], cn[8];
            colour_decomp(cp, Ap);
            colour_decomp(cn, An);
            double s = 0.0;
            for (int a = 0; a < 8; ++a) s += creal(mu == 0 ? xi * xi * cp[a] * cn[a] : cp[a] * cn[a]);
            acc += s;
        }
        const double Vd = (double)E[0] * (double)E[1] * (double)E[2] * (double)E[3];
        double prefac;
        if ((mu_start == 1 && q[1] + q[2] + q[1] < 1) || (mu_start == 0 && q[0] + q[1] + q[2] + q[3] + q[1] < 1))
            prefac = 2.0 / (8.0 * Vd * 4.0);
        else
            prefac = 2.0 / (8.0 * Vd * 3.0);
        D[o] = acc * prefac;
    }
    return ORC_OK;
}

/* scalarGluonProp, src/gluonCalc/gluonProp.f90:17-110, on a position-space field with
 * NT <= NS (the reference writes D(NS,NS,NS,NS) at every (t,x,y,z) of the field).  D starts
 * at zero (the reference accumulates into its intent(out) D without clearing it, :103-124),
 * and the sum over the volume of identical terms (:69-73) is carried out term by term as
 * the reference writes it. */
int or_scalar_gluon_prop(const cplx* data, const int32_t dims[4], double xi, double* D) {
    (void)xi; /* enters only for mu == 0, which the 1-based direction loop never reaches */
    if (check_dims(dims)) return ORC_EINVAL;
    const int32_t NT = dims[0], NS = dims[1];
    if (dims[2] != NS || dims[3] != NS || NT > NS) return ORC_EINVAL;
    const int64_t V = vol4(dims);
    const int32_t dD[4] = {NS, NS, NS, NS};
    memset(D, 0, sizeof(double) * (size_t)vol4(dD));
    for (int mu = 1; mu <= 4; ++mu) {
#pragma omp parallel for schedule(static)
        for (int64_t s = 0; s < V; ++s) {
            int32_t c[4];
            site_coord(dims, s, c);
            double qhat[4];
            for (int i = 0; i < 4; ++i) qhat[i] = (2.0 * M_PI / 1.0) * (double)c[i] / (double)dims[i];
            cplx Umu[9], T1[9], T2[9], A[9];
            load_link(data, dims, c, mu, Umu);
            for (int i = 0; i < 9; ++i) T1[i] = T2[i] = 0.0;
            const double ph = (double)c[mu - 1] * qhat[mu - 1];
            const cplx e1 = cexp(-I * ph), e2 = cexp(I * ph);
            for (int64_t k = 0; k < V; ++k)
                for (int i = 0; i < 9; ++i) {
                    T1[i] += e1 * Umu[i];
                    T2[i] += e2 * Umu[i];
                }
            traceless_conjg_subtract(A, T2, T1);
            const cplx p1 = (0.5 * I) * cexp(-I * qhat[mu - 1] * 0.5);
            const cplx p2 = (0.5 * I) * cexp(I * qhat[mu - 1] * 0.5);
            for (int i = 0; i < 9; ++i) A[i] = p2 * (p1 * A[i]);
            cplx com[8];
            colour_decomp(com, A);
            cplx cs = 0.0;
            for (int a = 0; a < 8; ++a) cs += com[a] * com[a];
            D[site0(dD, c[0] - 1, c[1] - 1, c[2] - 1, c[3] - 1)] += creal(cs);
        }
    }
    const double f = (1.0 / 3.0) * (1.0 / 8.0) * (double)V;
    for (int64_t i = 0; i < vol4(dD); ++i) D[i] = f * D[i];
    return ORC_OK;
}

/* Q_Average, src/gluonAna/gpManip.f90:10-57 (DOUT takes D(1) for every slice, as there) */
int or_q_average(int32_t NQ, const double* q, const double* D, int32_t q_slices, double* DOUT, double* QOUT,
                 int32_t* Qcount) {
    if (NQ < 1 || q_slices < 1) return ORC_EINVAL;
    for (int32_t i = 0; i < NQ; ++i)
        if (isnan(q[i])) return ORC_EINVAL;
    for (int32_t i = 0; i <= NQ; ++i) DOUT[i] = QOUT[i] = 0.0;
    double maxq = q[0];
    for (int32_t i = 1; i < NQ; ++i)
        if (q[i] > maxq) maxq = q[i];
    const double dq = maxq / (double)q_slices;
    /* inputs on which the reference's loop does not end */
    if (dq < 0.0 ? !(dq + dq >= maxq) : (dq == 0.0 && maxq > 0.0)) return ORC_EINVAL;
    double q_pos = 0.0;
    int32_t qcount = 1;
    for (int64_t it = 0;; ++it) {
        if (it > (int64_t)q_slices + 2) return ORC_EINVAL;
        double sum = 0.0;
        int32_t n = 0;
        for (int32_t i = 0; i < NQ; ++i)
            if (q[i] > q_pos && q[i] < q_pos + dq) {
                sum += q[i];
                ++n;
            }
        if (n > 0) {
            QOUT[qcount - 1] = sum / (double)n;
            DOUT[qcount - 1] = D[0];
            ++qcount;
        }
        q_pos = q_pos + dq;
        if (q_pos + dq >= maxq) break;
    }
    *Qcount = qcount;
    return ORC_OK;
}

/* cone_cut, src/gluonAna/gpManip.f90:59-163, with every optional argument present (as the
 * C wrapper c_wrapper/FLUE_c.f90:58-71 passes them); Q, QOUT are (NQ, 4) in Fortran order */
int or_cone_cut(int32_t NQ, const double* Q, const double* D, const double* D4, double angle, double xi,
                double IRCut, double IRRadius, double* QOUT, double* DOUT, double* D4OUT, int32_t* qcount) {
    if (NQ < 0) return ORC_EINVAL;
    int32_t n = 0;
    for (int32_t i = 0; i < 4 * NQ; ++i) QOUT[i] = 0.0;
    for (int32_t i = 0; i < NQ; ++i) DOUT[i] = D4OUT[i] = 0.0;
    for (int32_t k = 0; k < NQ; ++k) {
        double qh[4];
        for (int i = 0; i < 4; ++i) qh[i] = Q[k + (int64_t)NQ * i];
        const double q_norm = sqrt(((qh[0] * qh[0] + qh[1] * qh[1]) + qh[2] * qh[2]) + qh[3] * qh[3]);
        qh[0] = qh[0] * xi;
        double r, theta;
        if (qh[1] == qh[0] && qh[2] == qh[0] && qh[3] == qh[0]) {
            r = 0.0;
            theta = 0.0;
        } else {
            theta = acos((((0.5 * qh[0] + 0.5 * qh[1]) + 0.5 * qh[2]) + 0.5 * qh[3]) / q_norm);
            r = q_norm * sin(theta);
        }
        const int keep = q_norm <= IRCut ? (r <= IRRadius) : (r <= IRRadius && theta < angle);
        if (keep) {
            for (int i = 0; i < 4; ++i) QOUT[n + (int64_t)NQ * i] = Q[k + (int64_t)NQ * i];
            DOUT[n] = D[k];
            D4OUT[n] = D4[k];
            ++n;
        }
    }
    *qcount = n;
    return ORC_OK;
}
