// capi.hip -- the extern "C" surface of libflue_hip.so (declared in include/flue_c.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/flue_c.h"
#include "flue_launch.hpp"

using namespace flue;

namespace {

constexpr uint64_t kMaxSites = (1ull << 31) - 1;

struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};
std::mutex g_ws_mutex;
Workspace g_ws[64];

int hip_status(hipError_t e) {
    if (e == hipSuccess) return FLUE_OK;
    if (e == hipErrorOutOfMemory) return FLUE_ENOMEM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return FLUE_ENODEV;
    return FLUE_EHIP;
}

#define FLUE_TRY(expr)                  \
    do {                                \
        hipError_t e__ = (expr);        \
        if (e__ != hipSuccess) return hip_status(e__); \
    } while (0)

int check_dims(const int32_t* d) {
    if (!d) return FLUE_ENULL;
    uint64_t v = 1;
    for (int i = 0; i < 4; ++i) {
        if (d[i] < 1) return FLUE_EDIM;
        v *= (uint64_t)d[i];
        if (v > kMaxSites) return FLUE_EDIM;
    }
    return FLUE_OK;
}

int check_prec(int32_t p) { return (p == FLUE_F64 || p == FLUE_F32 || p == FLUE_F64_SU3) ? FLUE_OK : FLUE_EINVAL; }

int check_path(const int32_t* path, int32_t npath) {
    if (npath < 0 || npath > 64) return FLUE_EPATH;
    if (npath > 0 && !path) return FLUE_ENULL;
    for (int i = 0; i < npath; ++i)
        if (path[i] == 0 || path[i] > 4 || path[i] < -4) return FLUE_EPATH;
    return FLUE_OK;
}

// workspace: caller-provided, or the library's per-device buffer
int get_ws(void* ws, size_t ws_bytes, size_t need, void** out) {
    if (ws) {
        if (ws_bytes < need) return FLUE_EWS;
        *out = ws;
        return FLUE_OK;
    }
    int dev = 0;
    FLUE_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return FLUE_EINVAL;
    std::lock_guard<std::mutex> lk(g_ws_mutex);
    Workspace& w = g_ws[dev];
    if (w.bytes < need) {
        if (w.ptr) FLUE_TRY(hipFree(w.ptr));
        w.ptr = nullptr;
        w.bytes = 0;
        FLUE_TRY(hipMalloc(&w.ptr, need));
        w.bytes = need;
    }
    *out = w.ptr;
    return FLUE_OK;
}

// per-block compensated partials plus the tail area of the two-stage finalize
size_t partial_bytes(uint32_t nblocks) {
    return ((size_t)nblocks + (nblocks + kFinalChunk - 1) / kFinalChunk) * 2 * sizeof(double) + 256;
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// whether the rectangle-first magnetic form (variant 7) applies to the lattice
bool magnetic_rect_ok(const Lat& L, int32_t prec) { return rect_fits(L) && fits_buffer(L.V, prec == FLUE_F32); }

// magnetic workspace of a variant: the partials, and for variant 7 the footprints after them
size_t magnetic_rect_partials(const Lat& L, int32_t prec) {
    return align256(partial_bytes(3 * rect_tiles(L)));
}
size_t magnetic_ws_bytes(const Lat& L, int32_t prec, int variant) {
    if (variant == kMagneticRect && magnetic_rect_ok(L, prec))
        return magnetic_rect_partials(L, prec) + rect_fp_bytes(L, prec == FLUE_F32);
    return partial_bytes(magnetic_blocks(L, prec == FLUE_F32));
}

// runs magnetic `variant` (1..7) on workspace w and finalizes B into d_B
int magnetic_run(const Lat& L, const void* U, int32_t prec, int variant, void* w, double* d_B, hipStream_t st) {
    uint32_t np = 0;
    void* fp = variant == kMagneticRect ? (char*)w + magnetic_rect_partials(L, prec) : nullptr;
    FLUE_TRY(launch_magnetic(L, U, prec == FLUE_F32, prec == FLUE_F64_SU3, (double*)w, st, variant, &np, fp));
    if (d_B) FLUE_TRY(launch_finalize((double*)w, np, d_B, 1.0 / (double)L.V, st));
    return FLUE_OK;
}

// ---------------------------------------------------------------- SU(3) check
// per-block (count, max, max) triples and the combined triple
size_t check_bytes(const Lat& L) { return ((size_t)check_blocks(L) * 3 + 3) * sizeof(double) + 256; }

// B^2 adjoint work buffer: the anti-Hermitian clover sums (27 reals per site), 256-aligned
size_t b2_ah_bytes(const Lat& L, int32_t prec) {
    return ((size_t)27 * L.V * (prec == FLUE_F32 ? 4 : 8) + 255) & ~(size_t)255;
}

// the region after the clover sums: the reverse form's slot buffer, and before it, where the
// rectangle-first magnetic form applies, that form's footprints and link copy
size_t b2_region_bytes(const Lat& L, int32_t prec) {
    const size_t slots = b2_slot_bytes(L, prec == FLUE_F32);
    return magnetic_rect_ok(L, prec) ? std::max(slots, rect_fp_bytes(L, prec == FLUE_F32)) : slots;
}

// the B^2 adjoint after the SU(3) check: clover sums into the start of `work` (rectangle-first
// where it applies, with two-row chains in both precisions since the links are SU(3)), then
// variant 1 (per-link gather) or 2 (reverse mode with the slot buffer after the clover sums)
int b2_grad_run(const Lat& L, const void* U, int32_t prec, int variant, void* work, void* grad, hipStream_t st) {
    const int f32 = prec == FLUE_F32;
    void* region = (char*)work + b2_ah_bytes(L, prec);
    if (magnetic_rect_ok(L, prec))
        FLUE_TRY(launch_magnetic_rect(L, U, f32, 1, nullptr, region, st, nullptr, work));
    else
        FLUE_TRY(launch_magnetic_ah(L, U, f32, work, st));
    if (variant == 1)
        FLUE_TRY(launch_b2_grad(L, U, f32, work, grad, st));
    else
        FLUE_TRY(launch_b2_grad_rev(L, U, f32, work, region, grad, st));
    return FLUE_OK;
}

// flow work buffer: the second link buffer, then the accumulator, then the plaquette
// parts, then a third link buffer
size_t flow_acc_offset(const Lat& L, int32_t prec) {
    return ((size_t)L.V * kPlanes * (prec == FLUE_F32 ? 4 : 8) + 255) & ~(size_t)255;
}
size_t flow_w3_offset(const Lat& L, int32_t prec) {
    return flow_acc_offset(L, prec) + (size_t)L.V * 32 * 8 + align256(plaq_ta_bytes(L, prec == FLUE_F32));
}

// one RK3 stage (Luescher: acc <- ca acc + cb Z, (ca, cb) = (0, 1/4), (-17/9, 8/9), (-1, 3/4))
int flow_stage_run(const Lat& L, const void* Win, int32_t prec, void* Wout, void* acc, double eps, int stage,
                   int variant, void* T, hipStream_t st) {
    static const double ca[3] = {0.0, -17.0 / 9.0, -1.0};
    static const double cb[3] = {0.25, 8.0 / 9.0, 0.75};
    return hip_status(launch_flow_stage(L, Win, prec == FLUE_F32, Wout, acc, eps, ca[stage], cb[stage], stage == 2, st,
                                        variant, T));
}

// largest accepted max|U U^+ - 1| and |det U - 1|: a few hundred FP32 roundings of a unit
// entry for FP32 links, far above FP64 rounding for FP64 links
double su3_bound(int32_t prec) { return prec == FLUE_F32 ? 1e-5 : 1e-10; }

// scans U with `w` (check_bytes) as scratch, waits for the result, and returns FLUE_ESU3
// when a link has a non-finite entry or exceeds the bound (tol == 0: su3_bound(prec))
int run_su3_check(const Lat& L, const void* U, int32_t prec, double tol, double* stats, void* w, hipStream_t st) {
    double* part = (double*)w;
    double* out = part + 3 * (size_t)check_blocks(L);
    FLUE_TRY(launch_check_field(L, U, prec == FLUE_F32, part, out, st));
    double h[3];
    FLUE_TRY(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, st));
    FLUE_TRY(hipStreamSynchronize(st));
    if (stats) std::memcpy(stats, h, sizeof h);
    const double b = tol > 0.0 ? tol : su3_bound(prec);
    return (h[0] > 0.0 || !(h[1] <= b) || !(h[2] <= b)) ? FLUE_ESU3 : FLUE_OK;
}

// ---------------------------------------------------------------- autotuner table
// (device, op, precision, NT, NX, NY, NZ) -> fastest kernel variant measured on it
struct TuneKey {
    int dev, op, prec, d[4];
    bool operator<(const TuneKey& o) const {
        if (dev != o.dev) return dev < o.dev;
        if (op != o.op) return op < o.op;
        if (prec != o.prec) return prec < o.prec;
        for (int i = 0; i < 4; ++i)
            if (d[i] != o.d[i]) return d[i] < o.d[i];
        return false;
    }
};
struct TuneVal {
    int variant;
    double ms;
};
std::mutex g_tune_mutex;
std::map<TuneKey, TuneVal> g_tune;

TuneKey tune_key(int op, const int32_t dims[4], int prec) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    return TuneKey{dev, op, prec, {dims[0], dims[1], dims[2], dims[3]}};
}

int tuned_variant(int op, const int32_t dims[4], int prec) {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    auto it = g_tune.find(tune_key(op, dims, prec));
    return it == g_tune.end() ? 0 : it->second.variant;
}

// number of variants of each tunable operation (FLUE_OP_*)
int op_variants(int op) {
    switch (op) {
    case FLUE_OP_MAGNETIC: return kMagneticVariants;
    case FLUE_OP_B2_GRAD: return kB2Variants;
    case FLUE_OP_PLAQUETTE: return kPlaquetteVariants;
    case FLUE_OP_FLOW: return kFlowVariants;
    case FLUE_OP_ENSEMBLE_FLOW: return kEnsembleVariants;
    default: return 0;
    }
}

// the tuned variant of `op` for (device, prec, key), else `dflt`
int variant_or(int op, const int32_t key[4], int prec, int dflt) {
    const int v = tuned_variant(op, key, prec);
    return v > 0 ? v : dflt;
}

hipStream_t S(void* s) { return (hipStream_t)s; }

int fortran_strides(const int32_t d[4], int64_t s[7]) {
    const int64_t V = (int64_t)d[0] * d[1] * d[2] * d[3];
    s[0] = 1;
    s[1] = d[0];
    s[2] = (int64_t)d[0] * d[1];
    s[3] = (int64_t)d[0] * d[1] * d[2];
    s[4] = V;
    s[5] = 4 * V;
    s[6] = 12 * V;
    return FLUE_OK;
}

// RAII device buffer for the host drop-ins
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// host drop-in prologue: dims by reference, field copied to the device as FP64 SoA
struct HostField {
    int32_t dims[4];
    Lat L;
    DevBuf raw, soa;
};

int load_host_field(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ,
                    const flue_c128* data, HostField& h) {
    if (!NT || !NX || !NY || !NZ || !data) return FLUE_ENULL;
    h.dims[0] = *NT;
    h.dims[1] = *NX;
    h.dims[2] = *NY;
    h.dims[3] = *NZ;
    int rc = check_dims(h.dims);
    if (rc) return rc;
    h.L = make_lat(h.dims);
    const size_t n = (size_t)h.L.V * 36;
    FLUE_TRY(hipMalloc(&h.raw.p, n * sizeof(flue_c128)));
    FLUE_TRY(hipMalloc(&h.soa.p, flue_field_bytes(h.dims, FLUE_F64)));
    FLUE_TRY(hipMemcpy(h.raw.p, data, n * sizeof(flue_c128), hipMemcpyHostToDevice));
    int64_t st[7];
    fortran_strides(h.dims, st);
    FLUE_TRY(launch_import(h.L, h.raw.p, 0, st, h.soa.p, 0, nullptr));
    return FLUE_OK;
}

int fetch_scalar(const double* d, double* h) {
    FLUE_TRY(hipMemcpy(h, d, sizeof(double), hipMemcpyDeviceToHost));
    return FLUE_OK;
}

int plaquette_count(const int32_t dims[4], int32_t mu0, int32_t mu1, int32_t nu1, int32_t* nP) {
    int64_t planes = 0;
    for (int mu = mu0; mu <= mu1; ++mu)
        for (int nu = mu + 1; nu <= nu1; ++nu) ++planes;
    const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
    const int64_t n = planes * V;
    if (n > INT32_MAX) return FLUE_EDIM;  // nP is a default integer in the reference
    *nP = (int32_t)n;
    return FLUE_OK;
}

int check_plaq_args(int32_t muStart, int32_t muEnd, int32_t nuEnd) {
    if (muStart < 1 || muStart > 4 || muEnd < 1 || muEnd > 4 || nuEnd < 1 || nuEnd > 4) return FLUE_EINVAL;
    return FLUE_OK;
}

}  // namespace

extern "C" {

const char* flue_strerror(int s) {
    switch (s) {
    case FLUE_OK: return "ok";
    case FLUE_EINVAL: return "argument value out of range";
    case FLUE_ENULL: return "required pointer is NULL";
    case FLUE_EPATH: return "path entry must be +-1..+-4 and the path at most 64 steps";
    case FLUE_ENT: return "Polyakov loop needs NT >= 3";
    case FLUE_EDIM: return "lattice extents must be >= 1 and the volume below 2^31 sites";
    case FLUE_EHIP: return "HIP runtime error";
    case FLUE_ENOMEM: return "device allocation failed";
    case FLUE_ENOTIMPL: return "outside the scope of this library";
    case FLUE_ENODEV: return "no GPU visible";
    case FLUE_EWS: return "workspace too small";
    case FLUE_ESU3: return "links are not SU(3): a non-finite entry, or |U U^+ - 1| or |det U - 1| above the bound";
    default: return "unknown status";
    }
}

int flue_version(int32_t* major, int32_t* minor, int32_t* patch) {
    if (!major || !minor || !patch) return FLUE_ENULL;
    *major = 0;
    *minor = 1;
    *patch = 0;
    return FLUE_OK;
}

// ---------------------------------------------------------------- host helpers

int multiplymatmat_c(flue_c128 MM[9], const flue_c128 left[9], const flue_c128 right[9]) {
    if (!MM || !left || !right) return FLUE_ENULL;
    flue_c128 r[9];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            double re = 0, im = 0;
            for (int k = 0; k < 3; ++k) {
                const flue_c128 a = left[i + 3 * k], b = right[k + 3 * j];
                re += a.re * b.re - a.im * b.im;
                im += a.re * b.im + a.im * b.re;
            }
            r[i + 3 * j] = {re, im};
        }
    std::memcpy(MM, r, sizeof r);
    return FLUE_OK;
}

int get_qhat_c(const double coord[4], const int32_t shape[4], double qhat[4], const double* a) {
    if (!coord || !shape || !qhat) return FLUE_ENULL;
    const double lat = a ? *a : 1.0;
    for (int i = 0; i < 4; ++i) {
        if (shape[i] == 0) return FLUE_EINVAL;
        qhat[i] = (2.0 * M_PI / lat) * coord[i] / (double)shape[i];
    }
    return FLUE_OK;
}

int q_average_c(const int32_t* NQ, const double* q, const double* D, const int32_t* q_slices, double* DOUT,
                double* QOUT, int32_t* Qcount) {
    if (!NQ || !q || !D || !q_slices || !DOUT || !QOUT || !Qcount) return FLUE_ENULL;
    const int32_t n = *NQ;
    if (n < 1 || *q_slices < 1) return FLUE_EINVAL;
    for (int32_t i = 0; i < n; ++i)
        if (std::isnan(q[i])) return FLUE_EINVAL;
    std::fill(DOUT, DOUT + n + 1, 0.0);
    std::fill(QOUT, QOUT + n + 1, 0.0);
    const double maxq = *std::max_element(q, q + n);
    const double dq = maxq / (double)*q_slices;
    // the reference's loop ends only if the running bin edge reaches max(q): refuse the
    // inputs on which it would not (a negative step with one slice, a step of 0 below a
    // positive maximum)
    if (dq < 0.0 ? !(dq + dq >= maxq) : (dq == 0.0 && maxq > 0.0)) return FLUE_EINVAL;
    double lo = 0.0;
    int32_t count = 1;
    for (int64_t it = 0;; ++it) {
        if (it > (int64_t)*q_slices + 2) return FLUE_EINVAL;
        const double hi = lo + dq;
        double sum = 0.0;
        int32_t k = 0;
        for (int32_t i = 0; i < n; ++i)
            if (q[i] > lo && q[i] < hi) {
                sum += q[i];
                ++k;
            }
        // open bins (lo, lo + dq) are disjoint and each kept one holds a value: count <= n + 1
        if (k > 0) {
            QOUT[count - 1] = sum / (double)k;
            DOUT[count - 1] = D[0];
            ++count;
        }
        lo = hi;
        if (lo + dq >= maxq) break;
    }
    *Qcount = count;
    return FLUE_OK;
}

int cone_cut_c(const int32_t* NQ, const int32_t* radius, const double* Q, const double* D, const double* D4,
               const int32_t* NT, const int32_t* NS, const double* angleIN, const double* xiIN,
               const double* IRCutIN, const double* IRRadiusIN, double* QOUT, double* Dout, double* D4out,
               int32_t* qcount) {
    (void)radius;  // IRRadiusIN is always given here, and radius only stands in for it when absent
    (void)NT;
    (void)NS;
    if (!NQ || !Q || !D || !D4 || !angleIN || !xiIN || !IRCutIN || !IRRadiusIN || !QOUT || !Dout || !D4out || !qcount)
        return FLUE_ENULL;
    const int64_t n = *NQ;
    if (n < 0) return FLUE_EINVAL;
    const double angle = *angleIN, xi = *xiIN, ircut = *IRCutIN, irradius = *IRRadiusIN;
    std::fill(QOUT, QOUT + 4 * n, 0.0);
    std::fill(Dout, Dout + n, 0.0);
    std::fill(D4out, D4out + n, 0.0);
    int32_t kept = 0;
    for (int64_t k = 0; k < n; ++k) {
        double v[4];
        for (int i = 0; i < 4; ++i) v[i] = Q[k + n * i];
        // the norm is taken before the anisotropy rescales the first component
        const double norm = std::sqrt(((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3]);
        v[0] *= xi;
        double r = 0.0, theta = 0.0;
        if (!(v[1] == v[0] && v[2] == v[0] && v[3] == v[0])) {
            // angle to the (1,1,1,1)/2 diagonal
            theta = std::acos((((0.5 * v[0] + 0.5 * v[1]) + 0.5 * v[2]) + 0.5 * v[3]) / norm);
            r = norm * std::sin(theta);
        }
        const bool keep = norm <= ircut ? (r <= irradius) : (r <= irradius && theta < angle);
        if (keep) {
            for (int i = 0; i < 4; ++i) QOUT[kept + n * i] = Q[k + n * i];
            Dout[kept] = D[k];
            D4out[kept] = D4[k];
            ++kept;
        }
    }
    *qcount = kept;
    return FLUE_OK;
}

int periodcoord_c(const int32_t coord[4], const int32_t dataShape[7], int32_t out[4]) {
    if (!coord || !dataShape || !out) return FLUE_ENULL;
    for (int c = 0; c < 4; ++c) {
        out[c] = coord[c];
        if (coord[c] > dataShape[c])
            out[c] = 1;
        else if (coord[c] == 0)
            out[c] = dataShape[c];
    }
    return FLUE_OK;
}

// ---------------------------------------------------------------- device entry points

size_t flue_field_bytes(const int32_t dims[4], int32_t prec) {
    if (check_dims(dims) || check_prec(prec)) return 0;
    const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2] * dims[3];
    return (size_t)V * kPlanes * (prec == FLUE_F32 ? 4 : 8);
}

size_t flue_workspace_bytes(const int32_t dims[4]) {
    if (check_dims(dims)) return 0;
    const Lat L = make_lat(dims);
    uint32_t nb = plaquette_max_blocks(L);
    nb = std::max(nb, polyakov_blocks(L));
    nb = std::max(nb, path_blocks(L, -1, nullptr, 0));
    nb = std::max(nb, magnetic_blocks(L, 0));
    nb = std::max(nb, magnetic_blocks(L, 1));
    return std::max({partial_bytes(nb), wloop_partial_bytes(L, -1) + 256, wloop_partial_bytes(L, kEverySlice) + 256,
                     check_bytes(L)});
}

int flue_check_field(const int32_t dims[4], const void* U, int32_t prec, double tol, double stats[3], void* ws,
                     size_t ws_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U) return FLUE_ENULL;
    if (check_prec(prec) || std::isnan(tol) || tol < 0.0) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    void* w;
    if ((rc = get_ws(ws, ws_bytes, check_bytes(L), &w))) return rc;
    return run_su3_check(L, U, prec, tol, stats, w, S(stream));
}

int flue_import_field(const int32_t dims[4], const void* src, int32_t src_type, const int64_t strides[7], void* dst,
                      int32_t prec, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!src || !dst || !strides) return FLUE_ENULL;
    if (check_prec(prec) || (src_type != FLUE_C128 && src_type != FLUE_C64)) return FLUE_EINVAL;
    return hip_status(launch_import(make_lat(dims), src, src_type == FLUE_C64, strides, dst, prec == FLUE_F32, S(stream)));
}

int flue_export_field(const int32_t dims[4], const void* src, int32_t prec, void* dst, int32_t dst_type,
                      const int64_t strides[7], void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!src || !dst || !strides) return FLUE_ENULL;
    if (check_prec(prec) || (dst_type != FLUE_C128 && dst_type != FLUE_C64)) return FLUE_EINVAL;
    return hip_status(launch_export(make_lat(dims), src, prec == FLUE_F32, dst, dst_type == FLUE_C64, strides, S(stream)));
}

int flue_convert_field(const int32_t dims[4], const void* src, int32_t src_prec, void* dst, int32_t dst_prec,
                       void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!src || !dst) return FLUE_ENULL;
    if (check_prec(src_prec) || check_prec(dst_prec)) return FLUE_EINVAL;
    return hip_status(launch_convert(make_lat(dims), src, src_prec == FLUE_F32, dst, dst_prec == FLUE_F32, S(stream)));
}

int flue_random_field(const int32_t dims[4], uint64_t seed, double eps, void* dst, int32_t prec, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!dst) return FLUE_ENULL;
    if (check_prec(prec) || !std::isfinite(eps)) return FLUE_EINVAL;
    return hip_status(launch_random_field(make_lat(dims), seed, eps, dst, prec == FLUE_F32, S(stream)));
}

int flue_plaquette(const int32_t dims[4], const void* U, int32_t prec, int32_t muStart, int32_t muEnd, int32_t nuEnd,
                   double* d_sum, void* ws, size_t ws_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !d_sum) return FLUE_ENULL;
    if (check_prec(prec) || check_plaq_args(muStart, muEnd, nuEnd)) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    const int variant = variant_or(FLUE_OP_PLAQUETTE, dims, prec, kPlaquetteDefault);
    uint32_t nb = plaquette_blocks(L, variant);
    void* w;
    if ((rc = get_ws(ws, ws_bytes, partial_bytes(nb), &w))) return rc;
    FLUE_TRY(launch_plaquette(L, U, prec == FLUE_F32, muStart - 1, muEnd - 1, nuEnd - 1, (double*)w, S(stream),
                              variant, &nb));
    FLUE_TRY(launch_finalize((double*)w, nb, d_sum, 1.0, S(stream)));
    return FLUE_OK;
}

int flue_polyakov(const int32_t dims[4], const void* U, int32_t prec, double* d_sum, void* ws, size_t ws_bytes,
                  void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !d_sum) return FLUE_ENULL;
    if (check_prec(prec)) return FLUE_EINVAL;
    if (dims[0] < 3) return FLUE_ENT;
    const Lat L = make_lat(dims);
    const uint32_t nb = polyakov_blocks(L);
    void* w;
    if ((rc = get_ws(ws, ws_bytes, partial_bytes(nb), &w))) return rc;
    FLUE_TRY(launch_polyakov(L, U, prec == FLUE_F32, (double*)w, S(stream)));
    FLUE_TRY(launch_finalize((double*)w, nb, d_sum, 1.0, S(stream)));
    return FLUE_OK;
}

int flue_path_trace(const int32_t dims[4], const void* U, int32_t prec, const int32_t* path, int32_t npath,
                    int32_t tau, double* d_sum, void* ws, size_t ws_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !d_sum) return FLUE_ENULL;
    if (check_prec(prec)) return FLUE_EINVAL;
    if ((rc = check_path(path, npath))) return rc;
    if (npath < 1) return FLUE_EPATH;
    if (tau < 0 || tau > dims[0]) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    const int ts = tau - 1;  // -1 = all slices
    const uint32_t nb = path_blocks(L, ts, path, npath);
    void* w;
    if ((rc = get_ws(ws, ws_bytes, partial_bytes(nb), &w))) return rc;
    FLUE_TRY(launch_path_trace(L, U, prec == FLUE_F32, path, npath, ts, (double*)w, S(stream)));
    FLUE_TRY(launch_finalize((double*)w, nb, d_sum, 1.0, S(stream)));
    return FLUE_OK;
}

int flue_wilson_loops(const int32_t dims[4], const void* U, int32_t prec, int32_t mu, int32_t nu, int32_t Rmax,
                      int32_t Tmax, int32_t tau, double* d_out, void* ws, size_t ws_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !d_out) return FLUE_ENULL;
    if (check_prec(prec) || mu < 1 || mu > 4 || nu < 1 || nu > 4 || mu == nu) return FLUE_EINVAL;
    if (Rmax < 1 || Rmax > kMaxLoopRT || Tmax < 1 || Tmax > kMaxLoopRT) return FLUE_EINVAL;
    if (tau < 0 || tau > dims[0]) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    const int ts = tau - 1;  // -1 = all slices
    void* w;
    if ((rc = get_ws(ws, ws_bytes, wloop_partial_bytes(L, ts), &w))) return rc;
    const double scale = 1.0 / (double)(ts >= 0 ? L.Vs : L.V);
    FLUE_TRY(launch_wloop_table(L, U, prec == FLUE_F32, mu - 1, nu - 1, Rmax, Tmax, ts, (double*)w, d_out, scale,
                                S(stream)));
    return FLUE_OK;
}

int flue_wilson_loops_slices(const int32_t dims[4], const void* U, int32_t prec, int32_t mu, int32_t nu,
                             int32_t Rmax, int32_t Tmax, double* d_out, void* ws, size_t ws_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !d_out) return FLUE_ENULL;
    if (check_prec(prec) || mu < 1 || mu > 4 || nu < 1 || nu > 4 || mu == nu) return FLUE_EINVAL;
    if (Rmax < 1 || Rmax > kMaxLoopRT || Tmax < 1 || Tmax > kMaxLoopRT) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    void* w;
    if ((rc = get_ws(ws, ws_bytes, wloop_partial_bytes(L, kEverySlice), &w))) return rc;
    FLUE_TRY(launch_wloop_table(L, U, prec == FLUE_F32, mu - 1, nu - 1, Rmax, Tmax, kEverySlice, (double*)w, d_out,
                                1.0 / (double)L.Vs, S(stream)));
    return FLUE_OK;
}

int flue_generic_path(const int32_t dims[4], const void* U, int32_t prec, const int32_t coordBase[4],
                      const int32_t* path, int32_t npath, double* d_out, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !d_out || !coordBase) return FLUE_ENULL;
    if (check_prec(prec)) return FLUE_EINVAL;
    if ((rc = check_path(path, npath))) return rc;
    int32_t base0[4];
    for (int i = 0; i < 4; ++i) {
        if (coordBase[i] < 1 || coordBase[i] > dims[i]) return FLUE_EINVAL;
        base0[i] = coordBase[i] - 1;
    }
    return hip_status(launch_generic_path(make_lat(dims), U, prec == FLUE_F32, base0, path, npath, d_out, S(stream)));
}

int flue_fmunu(const int32_t dims[4], const void* U, int32_t prec, int32_t mu, int32_t nu, int32_t variant, void* out,
               int32_t layout, const int64_t strides[6], void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !out) return FLUE_ENULL;
    if (check_prec(prec) || mu < 1 || mu > 4 || nu < 1 || nu > 4 || (variant != 0 && variant != 1))
        return FLUE_EINVAL;
    if (layout != FLUE_OUT_STRIDED && layout != FLUE_OUT_SOA) return FLUE_EINVAL;
    if (layout == FLUE_OUT_STRIDED && !strides) return FLUE_ENULL;
    return hip_status(launch_fmunu(make_lat(dims), U, prec == FLUE_F32, mu - 1, nu - 1, variant, out,
                                   layout == FLUE_OUT_SOA, strides, S(stream)));
}

int flue_gluon_prop_mom(const int32_t dims[4], const void* Uq, int32_t prec, const int64_t qstrides[4],
                        int32_t mu_start, double xi, const int32_t dext[4], void* d_D, const int64_t dstrides[4],
                        void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!Uq || !d_D || !dext || !dstrides) return FLUE_ENULL;
    if (check_prec(prec) || mu_start < 0) return FLUE_EINVAL;
    for (int i = 0; i < 4; ++i)
        if (dext[i] < 1 || dext[i] > dims[i] / 2 + 1) return FLUE_EINVAL;
    Lat Lq = make_lat(dims);
    if (qstrides) {
        // any strides that keep every momentum inside its plane of V elements
        int64_t last = 0;
        for (int i = 0; i < 4; ++i) {
            if (qstrides[i] < 1) return FLUE_EINVAL;
            last += (int64_t)(dims[i] - 1) * qstrides[i];
        }
        if (last >= (int64_t)Lq.V) return FLUE_EINVAL;
        for (int i = 0; i < 4; ++i) Lq.st[i] = (uint32_t)qstrides[i];
    }
    return hip_status(launch_mom_scalar(Lq, Uq, prec == FLUE_F32, mu_start, xi, dext, dstrides, (double*)d_D,
                                        S(stream)));
}

int flue_scalar_gluon_prop(const int32_t dims[4], const void* U, int32_t prec, double* d_D, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    // sizes first: D(NS,NS,NS,NS) is written at every (t,x,y,z) of the field
    if (dims[2] != dims[1] || dims[3] != dims[1] || dims[0] > dims[1]) return FLUE_EDIM;
    const uint64_t ns = (uint64_t)dims[1];
    if (ns * ns * ns * ns > kMaxSites) return FLUE_EDIM;
    if (!U || !d_D) return FLUE_ENULL;
    if (check_prec(prec)) return FLUE_EINVAL;
    return hip_status(launch_scalar_gluon(make_lat(dims), U, prec == FLUE_F32, dims[1], d_D, S(stream)));
}

int flue_magnetic(const int32_t dims[4], const void* U, int32_t prec, double* d_B, void* ws, size_t ws_bytes,
                  void* stream) {
    return flue_magnetic_ex(dims, U, prec, 0, d_B, ws, ws_bytes, stream);
}

size_t flue_magnetic_work_bytes(const int32_t dims[4], int32_t prec) {
    if (check_dims(dims) || check_prec(prec)) return 0;
    const Lat L = make_lat(dims);
    return std::max(magnetic_ws_bytes(L, prec, kMagneticRect), magnetic_ws_bytes(L, prec, 1));
}

int flue_magnetic_ex(const int32_t dims[4], const void* U, int32_t prec, int32_t variant, double* d_B, void* ws,
                     size_t ws_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !d_B) return FLUE_ENULL;
    if (check_prec(prec) || variant < 0 || variant > kMagneticVariants) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    // untuned: the rectangle-first form where it applies, else the precision's quadrant form
    if (variant == 0) variant = tuned_variant(FLUE_OP_MAGNETIC, dims, prec);
    if (variant == 0) variant = magnetic_rect_ok(L, prec) ? kMagneticRect : magnetic_default_variant(prec == FLUE_F32);
    // the rectangle-first form needs tileable extents and room for its footprints; otherwise
    // the precision's default form
    if (variant == kMagneticRect &&
        (!magnetic_rect_ok(L, prec) || (ws && ws_bytes < magnetic_ws_bytes(L, prec, kMagneticRect))))
        variant = magnetic_default_variant(prec == FLUE_F32);
    void* w;
    if ((rc = get_ws(ws, ws_bytes, magnetic_ws_bytes(L, prec, variant), &w))) return rc;
    return magnetic_run(L, U, prec, variant, w, d_B, S(stream));
}

// ---------------------------------------------------------------- allocation-free magnetic B

}  // extern "C"

// Scratch of flue_magnetic_noalloc: the per-block compensated partials of the magnetic
// kernel and the tail of the two-stage finalize.  Each GPU holds its own copy at its own
// address until the process ends.  Host code never uses this name as an address (on the
// host it is only a stand-in for the device symbol); it uses the per-device address from
// hipGetSymbolAddress instead.
static __device__ __align__(256) unsigned char scratch[kMagneticScratchBytes];

namespace {

std::mutex g_scratch_mutex;
void* g_scratch_dev[64];  // device address of `scratch` per device, looked up on first use

// device address of `scratch` on the current device, looked up once per device with that
// device current
int scratch_address(void** out) {
    int dev = 0;
    FLUE_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return FLUE_EINVAL;
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        FLUE_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return FLUE_EHIP;
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return FLUE_OK;
}

}  // namespace

extern "C" {

int flue_magnetic_noalloc(const int32_t dims[4], const void* U, int32_t prec, int32_t variant, double* d_B,
                          void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (check_prec(prec) || variant < 0 || variant > kMagneticVariants) return FLUE_EINVAL;
    // the partials of the largest form (three per block of 256 sites) must fit the scratch
    const Lat L = make_lat(dims);
    if (partial_bytes(magnetic_blocks(L, prec == FLUE_F32)) > kMagneticScratchBytes) return FLUE_EWS;
    if (!U || !d_B) return FLUE_ENULL;
    void* w;
    if ((rc = scratch_address(&w))) return rc;
    // every partial the finalize reads is written by the magnetic kernel first, so nothing
    // needs clearing between calls
    return flue_magnetic_ex(dims, U, prec, variant, d_B, w, kMagneticScratchBytes, stream);
}

// ---------------------------------------------------------------- gradients and flow

size_t flue_b2_work_bytes(const int32_t dims[4], int32_t prec) {
    if (check_dims(dims) || check_prec(prec)) return 0;
    const Lat L = make_lat(dims);
    // the per-site anti-Hermitian clover sums, then the slot buffer of the reverse form (one
    // slab of time slices); the SU(3) check uses the start of the buffer first
    return std::max(b2_ah_bytes(L, prec) + b2_region_bytes(L, prec), check_bytes(L));
}

size_t flue_flow_work_bytes(const int32_t dims[4], int32_t prec) {
    if (check_dims(dims) || check_prec(prec)) return 0;
    const Lat L = make_lat(dims);
    // a second link buffer, the RK3 accumulator (8 su(3) coordinates per link, room for FP64), the
    // plaquette parts of the plaquette-first stencil and a third link buffer (an odd step)
    return flow_w3_offset(L, prec) + flow_acc_offset(L, prec);
}

int flue_wilson_action(const int32_t dims[4], const void* U, int32_t prec, double beta, double* d_S, void* ws,
                       size_t ws_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !d_S) return FLUE_ENULL;
    if (check_prec(prec) || !std::isfinite(beta)) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    const int variant = variant_or(FLUE_OP_PLAQUETTE, dims, prec, kPlaquetteDefault);
    uint32_t nb = plaquette_blocks(L, variant);
    void* w;
    if ((rc = get_ws(ws, ws_bytes, partial_bytes(nb), &w))) return rc;
    FLUE_TRY(launch_plaquette(L, U, prec == FLUE_F32, 0, 3, 3, (double*)w, S(stream), variant, &nb));
    // S = beta * sum_p (1 - ReTr P / 3) over the 6V plaquettes
    FLUE_TRY(launch_finalize_affine((double*)w, nb, d_S, -beta / 3.0, beta * 6.0 * (double)L.V, S(stream)));
    return FLUE_OK;
}

int flue_wilson_grad(const int32_t dims[4], const void* U, int32_t prec, double beta, void* grad, void* stream) {
    return flue_wilson_grad_ex(dims, U, prec, beta, 0, grad, nullptr, 0, stream);
}

size_t flue_pf_work_bytes(const int32_t dims[4], int32_t prec) {
    if (check_dims(dims) || check_prec(prec)) return 0;
    return plaq_ta_bytes(make_lat(dims), prec == FLUE_F32);
}

int flue_wilson_grad_ex(const int32_t dims[4], const void* U, int32_t prec, double beta, int32_t variant, void* grad,
                        void* ws, size_t ws_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !grad) return FLUE_ENULL;
    if (check_prec(prec) || !std::isfinite(beta) || variant < 0 || variant > kFlowVariants) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    if (variant == 0) variant = variant_or(FLUE_OP_FLOW, dims, prec, kFlowDefault);
    // the plaquette-first form needs unitary links (FP32 and declared-SU(3) fields) and a
    // workspace; otherwise the staple form in the same visit order
    void* T = nullptr;
    if (flow_is_pf(variant)) {
        if (prec == FLUE_F64 || !ws) variant = staple_order_of(variant);
        else if (ws_bytes < plaq_ta_bytes(L, prec == FLUE_F32)) return FLUE_EWS;
        else T = ws;
    }
    return hip_status(launch_wilson_grad(L, U, prec == FLUE_F32, beta, grad, S(stream), variant, T));
}

int flue_b2_grad(const int32_t dims[4], const void* U, int32_t prec, void* grad, void* work, size_t work_bytes,
                 void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !grad) return FLUE_ENULL;
    if (check_prec(prec)) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    void* A;
    if ((rc = get_ws(work, work_bytes, flue_b2_work_bytes(dims, prec), &A))) return rc;
    // the gather takes the product from a corner to the walk's end as X_c^+ F, which holds
    // for unitary links only: refuse fields that are not SU(3) to the precision's bound
    if ((rc = run_su3_check(L, U, prec, 0.0, nullptr, A, S(stream)))) return rc;
    return b2_grad_run(L, U, prec, variant_or(FLUE_OP_B2_GRAD, dims, prec, kB2DefaultVariant), A, grad, S(stream));
}

int flue_b2_grad_ex(const int32_t dims[4], const void* U, int32_t prec, int32_t variant, void* grad, void* work,
                    size_t work_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U || !grad) return FLUE_ENULL;
    if (check_prec(prec) || variant < 0 || variant > kB2Variants) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    void* A;
    if ((rc = get_ws(work, work_bytes, flue_b2_work_bytes(dims, prec), &A))) return rc;
    if ((rc = run_su3_check(L, U, prec, 0.0, nullptr, A, S(stream)))) return rc;
    if (variant == 0) variant = variant_or(FLUE_OP_B2_GRAD, dims, prec, kB2DefaultVariant);
    return b2_grad_run(L, U, prec, variant, A, grad, S(stream));
}

int flue_flow_stage(const int32_t dims[4], const void* Win, int32_t prec, void* Wout, double* acc, double eps,
                    int32_t stage, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!Win || !Wout || !acc) return FLUE_ENULL;
    if (Win == Wout) return FLUE_EINVAL;
    if (check_prec(prec) || stage < 0 || stage > 2 || !std::isfinite(eps)) return FLUE_EINVAL;
    // Luescher RK3: acc <- ca acc + cb Z, (ca, cb) = (0, 1/4), (-17/9, 8/9), (-1, 3/4)
    // a single stage has no workspace for the plaquette-first form: staple sums
    const int variant = staple_order_of(variant_or(FLUE_OP_FLOW, dims, prec, kFlowDefault));
    return flow_stage_run(make_lat(dims), Win, prec, Wout, acc, eps, stage, variant, nullptr, S(stream));
}

int flue_wilson_flow(const int32_t dims[4], void* U, int32_t prec, double eps, int32_t nsteps, int32_t use_graph,
                     void* work, size_t work_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U) return FLUE_ENULL;
    if (check_prec(prec) || nsteps < 0 || !std::isfinite(eps)) return FLUE_EINVAL;
    void* w;
    if ((rc = get_ws(work, work_bytes, flue_flow_work_bytes(dims, prec), &w))) return rc;
    const Lat L = make_lat(dims);
    void* W2 = w;
    double* acc = (double*)((char*)w + flow_acc_offset(L, prec));
    void* T = (char*)acc + (size_t)L.V * 32 * 8;
    const int variant = variant_or(FLUE_OP_FLOW, dims, prec, kFlowDefault);
    hipStream_t st = S(stream);
    // three stages per step alternate between U and W2, so two steps return the field to U
    auto two_steps = [&]() -> int {
        void* bufs[2] = {U, W2};
        for (int k = 0; k < 6; ++k) {
            int r = flow_stage_run(L, bufs[k & 1], prec, bufs[(k + 1) & 1], acc, eps, k % 3, variant, T, st);
            if (r) return r;
        }
        return FLUE_OK;
    };
    int32_t pairs = nsteps / 2;
    if (use_graph && pairs > 0) {
        if (st == nullptr) return FLUE_EINVAL;  // capture needs a non-default stream
        hipGraph_t graph;
        hipGraphExec_t exec;
        FLUE_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int r = two_steps();
        hipGraph_t g2 = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g2);
        if (r) {
            if (g2) (void)hipGraphDestroy(g2);
            return r;
        }
        FLUE_TRY(e);
        graph = g2;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return hip_status(e);
        }
        for (int32_t i = 0; i < pairs && e == hipSuccess; ++i) e = hipGraphLaunch(exec, st);
        hipError_t e2 = hipStreamSynchronize(st);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        FLUE_TRY(e);
        FLUE_TRY(e2);
    } else {
        for (int32_t i = 0; i < pairs; ++i)
            if ((rc = two_steps())) return rc;
    }
    if (nsteps % 2) {
        // last odd step: U -> W2 -> W3 -> U, so its last stage writes U (no copy back)
        void* bufs[4] = {U, W2, (char*)w + flow_w3_offset(L, prec), U};
        for (int k = 0; k < 3; ++k)
            if ((rc = flow_stage_run(L, bufs[k], prec, bufs[k + 1], acc, eps, k, variant, T, st))) return rc;
    }
    return FLUE_OK;
}

// ---------------------------------------------------------------- autotuner

}  // extern "C"

namespace {

// hipEvent timing of `reps` calls of launch(v) after one warm-up call, for v = 1..nvar
template <typename Fn>
int time_variants(int nvar, int reps, hipStream_t st, Fn&& launch, int* best, double* best_ms) {
    hipEvent_t e0, e1;
    FLUE_TRY(hipEventCreate(&e0));
    if (hipError_t ec = hipEventCreate(&e1); ec != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_status(ec);
    }
    int bv = 0, rc = FLUE_OK;
    float bms = 0.f;
    hipError_t e = hipSuccess;
    for (int v = 1; v <= nvar && e == hipSuccess && rc == FLUE_OK; ++v) {
        rc = launch(v);  // warm-up
        if (rc == FLUE_OK) e = hipEventRecord(e0, st);
        for (int r = 0; r < reps && e == hipSuccess && rc == FLUE_OK; ++r) rc = launch(v);
        if (rc == FLUE_OK && e == hipSuccess) e = hipEventRecord(e1, st);
        if (rc == FLUE_OK && e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (rc == FLUE_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (rc == FLUE_OK && e == hipSuccess && std::getenv("FLUE_TUNE_LOG"))
            std::fprintf(stderr, "flue_autotune: variant %d %.4f ms per call\n", v, (double)ms / reps);
        if (rc == FLUE_OK && e == hipSuccess && (bv == 0 || ms < bms)) {
            bv = v;
            bms = ms;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != FLUE_OK) return rc;
    FLUE_TRY(e);
    *best = bv;
    *best_ms = (double)bms / reps;
    return FLUE_OK;
}

void remember(int op, const int32_t key[4], int prec, int v, double ms) {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    g_tune[tune_key(op, key, prec)] = TuneVal{v, ms};
}

}  // namespace

extern "C" {

size_t flue_autotune_work_bytes(int32_t op, const int32_t dims[4], int32_t prec) {
    if (check_dims(dims) || check_prec(prec)) return 0;
    const Lat L = make_lat(dims);
    const size_t fb = flue_field_bytes(dims, prec);
    switch (op) {
    case FLUE_OP_MAGNETIC: return flue_magnetic_work_bytes(dims, prec);
    case FLUE_OP_PLAQUETTE: return partial_bytes(plaquette_max_blocks(L));
    case FLUE_OP_B2_GRAD: return flue_b2_work_bytes(dims, prec) + fb;  // work, then a gradient
    case FLUE_OP_FLOW:  // the flow's work buffer
        return flue_flow_work_bytes(dims, prec);
    default: return 0;
    }
}

int flue_autotune(int32_t op, const int32_t dims[4], const void* U, int32_t prec, int32_t reps, int32_t* best,
                  double* best_ms, void* ws, size_t ws_bytes, void* stream) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!U) return FLUE_ENULL;
    if (check_prec(prec) || reps < 1 || op_variants(op) == 0 || op == FLUE_OP_ENSEMBLE_FLOW) return FLUE_EINVAL;
    const Lat L = make_lat(dims);
    const int is_f32 = prec == FLUE_F32;
    if (op == FLUE_OP_MAGNETIC && !fits_buffer(L.V, is_f32)) {
        // only one variant exists for fields beyond the buffer-resource range
        if (best) *best = magnetic_default_variant(is_f32);
        if (best_ms) *best_ms = -1.0;
        return FLUE_OK;
    }
    void* w;
    if ((rc = get_ws(ws, ws_bytes, flue_autotune_work_bytes(op, dims, prec), &w))) return rc;
    hipStream_t st = S(stream);
    if (op == FLUE_OP_B2_GRAD && (rc = run_su3_check(L, U, prec, 0.0, nullptr, w, st))) return rc;
    double* facc = (double*)((char*)w + flow_acc_offset(L, prec));
    auto launch = [&](int v) -> int {
        switch (op) {
        case FLUE_OP_MAGNETIC: return magnetic_run(L, U, prec, v, w, nullptr, st);
        case FLUE_OP_PLAQUETTE: return hip_status(launch_plaquette(L, U, is_f32, 0, 3, 3, (double*)w, st, v, nullptr));
        case FLUE_OP_B2_GRAD: return b2_grad_run(L, U, prec, v, w, (char*)w + flue_b2_work_bytes(dims, prec), st);
        default: {  // FLUE_OP_FLOW: one RK3 step, U -> w -> w2 -> w.  Stage 0 sets the
                    // accumulator afresh, so every timed step sees the same values (stages 1
                    // and 2 alone would scale it on every call and slow each variant more).
            void* T = (char*)facc + (size_t)L.V * 32 * 8;
            void* w2 = (char*)w + flow_w3_offset(L, prec);
            int r = flow_stage_run(L, U, prec, w, facc, 0.01, 0, v, T, st);
            if (!r) r = flow_stage_run(L, w, prec, w2, facc, 0.01, 1, v, T, st);
            if (!r) r = flow_stage_run(L, w2, prec, w, facc, 0.01, 2, v, T, st);
            return r;
        }
        }
    };
    int bv = 0;
    double bms = 0.0;
    // the rectangle-first magnetic form (the last variant) only where it applies
    const int nvar = op == FLUE_OP_MAGNETIC && !magnetic_rect_ok(L, prec) ? kMagneticRect - 1 : op_variants(op);
    if ((rc = time_variants(nvar, reps, st, launch, &bv, &bms))) return rc;
    remember(op, dims, prec, bv, bms);
    if (best) *best = bv;
    if (best_ms) *best_ms = bms;
    return FLUE_OK;
}

int flue_tuned_variant(int32_t op, const int32_t dims[4], int32_t prec, int32_t* variant) {
    int rc = check_dims(dims);
    if (rc) return rc;
    if (!variant) return FLUE_ENULL;
    if (check_prec(prec)) return FLUE_EINVAL;
    *variant = tuned_variant(op, dims, prec);
    return FLUE_OK;
}

int flue_tune_save(const char* path) {
    if (!path) return FLUE_ENULL;
    FILE* f = std::fopen(path, "w");
    if (!f) return FLUE_EINVAL;
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    std::fprintf(f, "# flue autotune: device op prec NT NX NY NZ variant ms_per_call\n");
    for (const auto& kv : g_tune)
        std::fprintf(f, "%d %d %d %d %d %d %d %d %.6f\n", kv.first.dev, kv.first.op, kv.first.prec, kv.first.d[0],
                     kv.first.d[1], kv.first.d[2], kv.first.d[3], kv.second.variant, kv.second.ms);
    std::fclose(f);
    return FLUE_OK;
}

int flue_tune_load(const char* path) {
    if (!path) return FLUE_ENULL;
    FILE* f = std::fopen(path, "r");
    if (!f) return FLUE_EINVAL;
    char line[256];
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#') continue;
        TuneKey k;
        TuneVal v;
        if (std::sscanf(line, "%d %d %d %d %d %d %d %d %lf", &k.dev, &k.op, &k.prec, &k.d[0], &k.d[1], &k.d[2],
                        &k.d[3], &v.variant, &v.ms) == 9 &&
            v.variant >= 1 && v.variant <= op_variants(k.op))
            g_tune[k] = v;
    }
    std::fclose(f);
    return FLUE_OK;
}

// ---------------------------------------------------------------- ensembles

namespace {

struct EnsPlan {
    std::vector<EnsEntry> tab;
    uint64_t field_reals = 0, acc_doubles = 0;
    uint32_t total_blocks = 0, maxV = 0;
};

int ens_plan(int32_t n, const int32_t* dims, EnsPlan& p) {
    if (n < 1 || n > 65536) return FLUE_EINVAL;
    if (!dims) return FLUE_ENULL;
    p.tab.resize(n);
    uint64_t blocks = 0;
    for (int c = 0; c < n; ++c) {
        int rc = check_dims(dims + 4 * c);
        if (rc) return rc;
        EnsEntry& e = p.tab[c];
        e.L = make_lat(dims + 4 * c);
        e.off = p.field_reals;
        e.acc_off = p.acc_doubles;
        e.blk0 = (uint32_t)blocks;
        e.nblk = (e.L.V + 255) / 256;
        p.field_reals += (uint64_t)e.L.V * kPlanes;
        p.acc_doubles += (uint64_t)e.L.V * 32;
        blocks += e.nblk;
        p.maxV = std::max(p.maxV, e.L.V);
    }
    if (blocks > 0xFFFFFFFFull) return FLUE_EDIM;
    p.total_blocks = (uint32_t)blocks;
    return FLUE_OK;
}

// autotuner key of an ensemble: its size, logical blocks and largest volume
void ens_key(const EnsPlan& p, int32_t n, int32_t key[4]) {
    key[0] = n;
    key[1] = (int32_t)p.total_blocks;
    key[2] = (int32_t)p.maxV;
    key[3] = 0;
}

// persistent grid of the ensemble kernels: 1, 2, 4 or 8 workgroups per CU (variant 1..4, and
// 5..8 with the plaquette-first stencil)
int ens_grid(const EnsPlan& p, int variant, uint32_t* grid) {
    int dev = 0, cus = 0;
    FLUE_TRY(hipGetDevice(&dev));
    FLUE_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t per_cu = 1u << ((variant - 1) % 4);
    *grid = std::max<uint32_t>(1, std::min<uint32_t>(p.total_blocks, (uint32_t)std::max(cus, 1) * per_cu));
    return FLUE_OK;
}

// the ensemble work buffer: table, second link buffer, accumulators (room for FP64), partial sums, and
// the plaquette parts of the plaquette-first stencil
struct EnsWork {
    EnsEntry* tab;
    void* W2;
    double* acc;
    double* partials;
    void* T;
};

EnsWork ens_work(void* w, int32_t n, const EnsPlan& p, size_t sz) {
    char* base = (char*)w;
    EnsWork r;
    r.tab = (EnsEntry*)base;
    base += align256(sizeof(EnsEntry) * n);
    r.W2 = base;
    base += align256((size_t)p.field_reals * sz);
    r.acc = (double*)base;
    base += align256((size_t)p.acc_doubles * 8);
    r.partials = (double*)base;
    base += align256((size_t)p.total_blocks * 16);
    r.T = base;
    return r;
}

}  // namespace

size_t flue_ensemble_field_bytes(int32_t n, const int32_t* dims, int32_t prec) {
    EnsPlan p;
    if (ens_plan(n, dims, p) || check_prec(prec)) return 0;
    return (size_t)p.field_reals * (prec == FLUE_F32 ? 4 : 8);
}

size_t flue_ensemble_work_bytes(int32_t n, const int32_t* dims, int32_t prec) {
    EnsPlan p;
    if (ens_plan(n, dims, p) || check_prec(prec)) return 0;
    // table, second link buffer, accumulators (room for FP64), per-block partial sums, plaquette parts
    const size_t sz = prec == FLUE_F32 ? 4 : 8;
    return align256(sizeof(EnsEntry) * n) + align256((size_t)p.field_reals * sz) + align256((size_t)p.acc_doubles * 8) +
           align256((size_t)p.total_blocks * 16) + align256((size_t)p.field_reals / 72 * 48 * sz);
}

int flue_ensemble_random(int32_t n, const int32_t* dims, const uint64_t* seeds, double eps, void* fields, int32_t prec,
                         void* stream) {
    EnsPlan p;
    int rc = ens_plan(n, dims, p);
    if (rc) return rc;
    if (!seeds || !fields) return FLUE_ENULL;
    if (check_prec(prec) || !std::isfinite(eps)) return FLUE_EINVAL;
    const size_t sz = prec == FLUE_F32 ? 4 : 8;
    for (int c = 0; c < n; ++c)
        FLUE_TRY(launch_random_field(p.tab[c].L, seeds[c], eps, (char*)fields + p.tab[c].off * sz, prec == FLUE_F32,
                                     S(stream)));
    return FLUE_OK;
}

int flue_ensemble_flow(int32_t n, const int32_t* dims, void* fields, int32_t prec, double eps, int32_t nsteps,
                       int32_t measure_every, double* d_plaq, double* d_B, int32_t use_graph, void* work,
                       size_t work_bytes, void* stream) {
    EnsPlan p;
    int rc = ens_plan(n, dims, p);
    if (rc) return rc;
    if (!fields) return FLUE_ENULL;
    if (check_prec(prec) || nsteps < 0 || measure_every < 0 || !std::isfinite(eps)) return FLUE_EINVAL;
    if (measure_every > 0 && (!d_plaq || !d_B)) return FLUE_ENULL;
    const int is_f32 = prec == FLUE_F32;
    const size_t sz = is_f32 ? 4 : 8;
    void* w;
    if ((rc = get_ws(work, work_bytes, flue_ensemble_work_bytes(n, dims, prec), &w))) return rc;
    const EnsWork ew = ens_work(w, n, p, sz);
    EnsEntry* d_tab = ew.tab;
    void* W2 = ew.W2;
    double* acc = ew.acc;
    double* partials = ew.partials;
    hipStream_t st = S(stream);
    FLUE_TRY(hipMemcpyAsync(d_tab, p.tab.data(), sizeof(EnsEntry) * n, hipMemcpyHostToDevice, st));
    FLUE_TRY(hipStreamSynchronize(st));  // the host table goes out of scope with this call

    int32_t key[4];
    ens_key(p, n, key);
    const int variant = variant_or(FLUE_OP_ENSEMBLE_FLOW, key, prec, kEnsembleDefault);
    uint32_t grid = 0;
    if ((rc = ens_grid(p, variant, &grid))) return rc;
    void* T = variant > 4 ? ew.T : nullptr;  // the plaquette-first variants

    const int nmeas = measure_every > 0 ? nsteps / measure_every + 1 : 0;
    static const double ca[3] = {0.0, -17.0 / 9.0, -1.0};
    static const double cb[3] = {0.25, 8.0 / 9.0, 0.75};
    void* bufs[2] = {fields, W2};
    auto measure = [&](const void* fld, int m) -> hipError_t {
        hipError_t e = launch_ens_sum(d_tab, n, p.total_blocks, p.maxV, fld, is_f32, 0, partials, grid, st);
        if (e == hipSuccess) e = launch_ens_finalize(d_tab, n, partials, d_plaq + (size_t)m * n, 1, 0, st);
        if (e == hipSuccess) e = launch_ens_sum(d_tab, n, p.total_blocks, p.maxV, fld, is_f32, 1, partials, grid, st);
        if (e == hipSuccess) e = launch_ens_finalize(d_tab, n, partials, d_B + (size_t)m * n, 1, 1, st);
        return e;
    };
    // the whole run: measurement at step 0 and after every measure_every steps
    auto run = [&]() -> hipError_t {
        hipError_t e = hipSuccess;
        if (nmeas > 0) e = measure(fields, 0);
        for (int32_t j = 0; j < nsteps && e == hipSuccess; ++j) {
            for (int k = 0; k < 3 && e == hipSuccess; ++k) {
                const int q = 3 * j + k;
                e = launch_ens_flow_stage(d_tab, n, p.total_blocks, p.maxV, bufs[q & 1], is_f32, bufs[(q + 1) & 1], acc,
                                          eps, ca[k], cb[k], k == 2, grid, T, st);
            }
            if (e == hipSuccess && nmeas > 0 && (j + 1) % measure_every == 0)
                e = measure(bufs[(j + 1) & 1], (j + 1) / measure_every);
        }
        if (e == hipSuccess && (nsteps & 1)) e = launch_copy((size_t)p.field_reals, is_f32, W2, fields, st);
        return e;
    };
    if (use_graph) {
        if (st == nullptr) return FLUE_EINVAL;  // capture needs a non-default stream
        FLUE_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = run();
        hipGraph_t graph = nullptr;
        hipError_t e2 = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return hip_status(e != hipSuccess ? e : e2);
        }
        hipGraphExec_t exec;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e == hipSuccess) {
            e = hipGraphLaunch(exec, st);
            e2 = hipStreamSynchronize(st);
            (void)hipGraphExecDestroy(exec);
        }
        (void)hipGraphDestroy(graph);
        FLUE_TRY(e);
        FLUE_TRY(e2);
    } else {
        FLUE_TRY(run());
    }
    return FLUE_OK;
}

int flue_ensemble_autotune(int32_t n, const int32_t* dims, const void* fields, int32_t prec, int32_t reps,
                           int32_t* best, double* best_ms, void* work, size_t work_bytes, void* stream) {
    EnsPlan p;
    int rc = ens_plan(n, dims, p);
    if (rc) return rc;
    if (!fields) return FLUE_ENULL;
    if (check_prec(prec) || reps < 1) return FLUE_EINVAL;
    const int is_f32 = prec == FLUE_F32;
    const size_t sz = is_f32 ? 4 : 8;
    void* w;
    if ((rc = get_ws(work, work_bytes, flue_ensemble_work_bytes(n, dims, prec), &w))) return rc;
    const EnsWork ew = ens_work(w, n, p, sz);
    hipStream_t st = S(stream);
    FLUE_TRY(hipMemcpyAsync(ew.tab, p.tab.data(), sizeof(EnsEntry) * n, hipMemcpyHostToDevice, st));
    FLUE_TRY(hipStreamSynchronize(st));
    // the first RK3 stage of every field into the second buffer (the fields are not changed;
    // it only writes the accumulator, so repeated calls do not drift, as flue_autotune)
    auto launch = [&](int v) -> int {
        uint32_t grid = 0;
        int r = ens_grid(p, v, &grid);
        if (r) return r;
        return hip_status(launch_ens_flow_stage(ew.tab, n, p.total_blocks, p.maxV, fields, is_f32, ew.W2, ew.acc, 0.01,
                                                0.0, 0.25, 0, grid, v > 4 ? ew.T : nullptr, st));
    };
    int bv = 0;
    double bms = 0.0;
    if ((rc = time_variants(kEnsembleVariants, reps, st, launch, &bv, &bms))) return rc;
    int32_t key[4];
    ens_key(p, n, key);
    remember(FLUE_OP_ENSEMBLE_FLOW, key, prec, bv, bms);
    if (best) *best = bv;
    if (best_ms) *best_ms = bms;
    return FLUE_OK;
}

// ---------------------------------------------------------------- host drop-ins

int genericpath_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
                  const int32_t coordBase[4], const int32_t* npath, const int32_t* path, flue_c128 out[9]) {
    if (!npath || !coordBase || !out) return FLUE_ENULL;
    int rc = check_path(path, *npath);
    if (rc) return rc;
    HostField h;
    if ((rc = load_host_field(NT, NX, NY, NZ, data, h))) return rc;
    DevBuf d;
    FLUE_TRY(hipMalloc(&d.p, 18 * sizeof(double)));
    if ((rc = flue_generic_path(h.dims, h.soa.p, FLUE_F64, coordBase, path, *npath, (double*)d.p, nullptr)))
        return rc;
    FLUE_TRY(hipMemcpy(out, d.p, 18 * sizeof(double), hipMemcpyDeviceToHost));
    return FLUE_OK;
}

static int plaq_host(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ,
                     const flue_c128* data, const int32_t* muStart, const int32_t* muEnd, const int32_t* nuEnd,
                     double* sumTrP, int32_t* nP, double* time) {
    const double t0 = now_s();
    if (!muStart || !muEnd || !nuEnd || !sumTrP || !nP || !time) return FLUE_ENULL;
    int rc = check_plaq_args(*muStart, *muEnd, *nuEnd);
    if (rc) return rc;
    HostField h;
    if ((rc = load_host_field(NT, NX, NY, NZ, data, h))) return rc;
    if ((rc = plaquette_count(h.dims, *muStart, *muEnd, *nuEnd, nP))) return rc;
    DevBuf d;
    FLUE_TRY(hipMalloc(&d.p, sizeof(double)));
    if ((rc = flue_plaquette(h.dims, h.soa.p, FLUE_F64, *muStart, *muEnd, *nuEnd, (double*)d.p, nullptr, 0, nullptr)))
        return rc;
    if ((rc = fetch_scalar((double*)d.p, sumTrP))) return rc;
    *time = now_s() - t0;
    return FLUE_OK;
}

int genplaquette_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
                   const int32_t* muStart, const int32_t* muEnd, const int32_t* nuEnd, double* sumTrP, int32_t* nP,
                   double* time) {
    return plaq_host(NT, NX, NY, NZ, data, muStart, muEnd, nuEnd, sumTrP, nP, time);
}

int plaquette_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
                const int32_t* muStart, const int32_t* muEnd, const int32_t* nuEnd, double* sumTrP, int32_t* nP,
                double* time) {
    return plaq_host(NT, NX, NY, NZ, data, muStart, muEnd, nuEnd, sumTrP, nP, time);
}

int polyakov_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
               double* sumTrP, int32_t* nP, double* time) {
    const double t0 = now_s();
    if (!sumTrP || !nP || !time) return FLUE_ENULL;
    if (NT && *NT < 3 && *NT >= 1) return FLUE_ENT;
    HostField h;
    int rc;
    if ((rc = load_host_field(NT, NX, NY, NZ, data, h))) return rc;
    DevBuf d;
    FLUE_TRY(hipMalloc(&d.p, sizeof(double)));
    if ((rc = flue_polyakov(h.dims, h.soa.p, FLUE_F64, (double*)d.p, nullptr, 0, nullptr))) return rc;
    if ((rc = fetch_scalar((double*)d.p, sumTrP))) return rc;
    *nP = (int32_t)h.L.Vs;
    *time = now_s() - t0;
    return FLUE_OK;
}

static int fmunu_host(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ,
                      const flue_c128* data, const int32_t* mu, const int32_t* nu, flue_c128* out, int variant) {
    if (!mu || !nu || !out) return FLUE_ENULL;
    if (*mu < 1 || *mu > 4 || *nu < 1 || *nu > 4) return FLUE_EINVAL;
    HostField h;
    int rc;
    if ((rc = load_host_field(NT, NX, NY, NZ, data, h))) return rc;
    const size_t n = (size_t)h.L.V * 9;
    DevBuf d;
    FLUE_TRY(hipMalloc(&d.p, n * sizeof(flue_c128)));
    const int64_t V = h.L.V;
    const int64_t st[6] = {1, h.dims[0], (int64_t)h.dims[0] * h.dims[1], (int64_t)h.dims[0] * h.dims[1] * h.dims[2],
                           V, 3 * V};
    if ((rc = flue_fmunu(h.dims, h.soa.p, FLUE_F64, *mu, *nu, variant, d.p, FLUE_OUT_STRIDED, st, nullptr))) return rc;
    FLUE_TRY(hipMemcpy(out, d.p, n * sizeof(flue_c128), hipMemcpyDeviceToHost));
    return FLUE_OK;
}

int cloverloopmunucoord_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ,
                          const flue_c128* data, const int32_t* mu, const int32_t* nu, flue_c128* out) {
    return fmunu_host(NT, NX, NY, NZ, data, mu, nu, out, 1);
}

int loop5munucoord_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ,
                     const flue_c128* data, const int32_t* mu, const int32_t* nu, flue_c128* out) {
    return fmunu_host(NT, NX, NY, NZ, data, mu, nu, out, 0);
}

int magnetic_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
               double* B) {
    if (!B) return FLUE_ENULL;
    HostField h;
    int rc;
    if ((rc = load_host_field(NT, NX, NY, NZ, data, h))) return rc;
    DevBuf d;
    FLUE_TRY(hipMalloc(&d.p, sizeof(double)));
    if ((rc = flue_magnetic(h.dims, h.soa.p, FLUE_F64, (double*)d.p, nullptr, 0, nullptr))) return rc;
    return fetch_scalar((double*)d.p, B);
}

static int rect_host(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ,
                     const flue_c128* data, int32_t tau, const int32_t* mu, const int32_t* nu, const int32_t* nMu,
                     const int32_t* nNu, double* loopVal) {
    if (!mu || !nu || !nMu || !nNu || !loopVal) return FLUE_ENULL;
    if (*mu < 1 || *mu > 4 || *nu < 1 || *nu > 4 || *nMu < 1 || *nNu < 1) return FLUE_EINVAL;
    if (2 * (*nMu + *nNu) > 64) return FLUE_EPATH;
    int32_t path[64];
    int n = 0;
    for (int i = 0; i < *nMu; ++i) path[n++] = *mu;
    for (int i = 0; i < *nNu; ++i) path[n++] = *nu;
    for (int i = 0; i < *nMu; ++i) path[n++] = -*mu;
    for (int i = 0; i < *nNu; ++i) path[n++] = -*nu;
    HostField h;
    int rc;
    if ((rc = load_host_field(NT, NX, NY, NZ, data, h))) return rc;
    if (tau < 0 || tau > h.dims[0]) return FLUE_EINVAL;
    DevBuf d;
    FLUE_TRY(hipMalloc(&d.p, sizeof(double)));
    if ((rc = flue_path_trace(h.dims, h.soa.p, FLUE_F64, path, n, tau, (double*)d.p, nullptr, 0, nullptr))) return rc;
    double s;
    if ((rc = fetch_scalar((double*)d.p, &s))) return rc;
    *loopVal = s / (double)(tau > 0 ? h.L.Vs : h.L.V);
    return FLUE_OK;
}

int wmunu_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
            const int32_t* mu, const int32_t* nu, const int32_t* nMu, const int32_t* nNu, double* loopVal) {
    return rect_host(NT, NX, NY, NZ, data, 0, mu, nu, nMu, nNu, loopVal);
}

int wmunutau_c(const int32_t* NT, const int32_t* NX, const int32_t* NY, const int32_t* NZ, const flue_c128* data,
               const int32_t* tau, const int32_t* mu, const int32_t* nu, const int32_t* nMu, const int32_t* nNu,
               double* loopVal) {
    if (!tau) return FLUE_ENULL;
    if (*tau < 1) return FLUE_EINVAL;
    return rect_host(NT, NX, NY, NZ, data, *tau, mu, nu, nMu, nNu, loopVal);
}

int calc_mom_space_scalard_c(const int32_t* NS, const int32_t* NT, const flue_c128* U, flue_c128* D,
                             const int32_t* mu_start, const double* xi) {
    if (!NS || !NT || !U || !D || !mu_start || !xi) return FLUE_ENULL;
    if (*NT < 0 || *NS < 0 || *NT == INT32_MAX || *NS == INT32_MAX) return FLUE_EDIM;
    if (*mu_start < 0) return FLUE_EINVAL;
    // U(0:NT, 0:NS, 0:NS, 0:NS, 0:3, 0:2, 0:2) and D(0:NT/2, 0:NS/2, 0:NS/2, 0:NS/2)
    const int32_t Et = *NT + 1, Es = *NS + 1;
    HostField h;
    int rc;
    if ((rc = load_host_field(&Et, &Es, &Es, &Es, U, h))) return rc;
    const int32_t dext[4] = {*NT / 2 + 1, *NS / 2 + 1, *NS / 2 + 1, *NS / 2 + 1};
    const int64_t dst[4] = {1, dext[0], (int64_t)dext[0] * dext[1], (int64_t)dext[0] * dext[1] * dext[2]};
    const size_t n = (size_t)dst[3] * dext[3];
    DevBuf d;
    FLUE_TRY(hipMalloc(&d.p, n * sizeof(flue_c128)));
    if ((rc = flue_gluon_prop_mom(h.dims, h.soa.p, FLUE_F64, nullptr, *mu_start, *xi, dext, d.p, dst, nullptr)))
        return rc;
    FLUE_TRY(hipMemcpy(D, d.p, n * sizeof(flue_c128), hipMemcpyDeviceToHost));
    return FLUE_OK;
}

int scalargluonprop_c(const int32_t* NS, const int32_t* NT, const flue_c128* U, const double* xi, double* D) {
    if (!NS || !NT || !U || !xi || !D) return FLUE_ENULL;
    if (*NT > *NS || *NS < 1) return FLUE_EDIM;
    const uint64_t ns = (uint64_t)*NS;
    if (ns * ns * ns * ns > kMaxSites) return FLUE_EDIM;
    HostField h;
    int rc;
    if ((rc = load_host_field(NT, NS, NS, NS, U, h))) return rc;
    const size_t n = (size_t)h.dims[1] * h.dims[1] * h.dims[1] * h.dims[1];
    DevBuf d;
    FLUE_TRY(hipMalloc(&d.p, n * sizeof(double)));
    if ((rc = flue_scalar_gluon_prop(h.dims, h.soa.p, FLUE_F64, (double*)d.p, nullptr))) return rc;
    FLUE_TRY(hipMemcpy(D, d.p, n * sizeof(double), hipMemcpyDeviceToHost));
    return FLUE_OK;
}

}  // extern "C"
