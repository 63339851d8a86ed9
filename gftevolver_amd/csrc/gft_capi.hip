// gft_capi.hip -- the C ABI of libgft.so (declared in include/gft.h).
//
// One translation unit: the kernel files are included here so every launch
// site sees its kernel.  Each context owns its device buffers and one HIP
// stream created at the default priority.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gft.h"
#include "gft_math.hpp"
#include "gft_reduce.hpp"
#include "gft_twoslot.hip"
#include "gft_momentum.hip"
#include "gft_mlp.hip"
#include "gft_grf.hip"

// The finishing reduction declared in gft_reduce.hpp.
namespace gft {
__global__ void __launch_bounds__(kFinThreads) finalize_partials(const double* __restrict__ partial, long long nblocks,
                                                                 int stride, int nsum, int nmax,
                                                                 double* __restrict__ out) {
    __shared__ double s_hi[kFinWidth * kFinThreads], s_lo[kFinWidth * kFinThreads];
    const int width = nsum + nmax;
    Kahan k[kFinWidth];
    double mx[kFinWidth];
#pragma unroll
    for (int q = 0; q < kFinWidth; q++) { k[q].init(); mx[q] = 0.0; }
    for (long long b = threadIdx.x; b < nblocks; b += kFinThreads) {
        const double* row = partial + b * stride;
#pragma unroll
        for (int q = 0; q < kFinWidth; q++) {
            if (q < nsum) k[q].add(row[q]);
            else if (q < width) mx[q] = fmax(mx[q], row[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < kFinWidth; q++) {
        s_hi[q * kFinThreads + threadIdx.x] = (q < nsum) ? k[q].s : mx[q];
        s_lo[q * kFinThreads + threadIdx.x] = (q < nsum) ? k[q].c : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < width) {
        const int q = threadIdx.x;
        Kahan t;
        t.init();
        double m = 0.0;
        for (int i = 0; i < kFinThreads; i++) {
            if (q < nsum) { t.add(s_hi[q * kFinThreads + i]); t.add(s_lo[q * kFinThreads + i]); }
            else m = fmax(m, s_hi[q * kFinThreads + i]);
        }
        out[q] = (q < nsum) ? t.value() : m;
    }
}
}  // namespace gft

using namespace gft;

namespace {

thread_local std::string g_create_error;

constexpr int kBlock = 256;
constexpr int kEventPool = 256;

bool is_two_slot(int form) { return form == GFT_FORM_TWO_SLOT_F64 || form == GFT_FORM_TWO_SLOT_F32; }
bool is_f64(int form) { return form == GFT_FORM_TWO_SLOT_F64 || form == GFT_FORM_MOMENTUM_F64; }
// Gauss generators of the gauge group: gaugeNum 4, 5 or 7; none without links.
int gauge_num(const gft_desc& d) {
    if (d.model == GFT_MODEL_GLOBAL) return 0;
    return (d.gauge_num == 5 || d.gauge_num == 7) ? d.gauge_num : 4;
}
// Link reals per site of the two-slot arrays: 3 x [y, SU(2) as c0..c3 or w1..w3,
// Higgs-family part (gaugeNum 5: y'; 7: c0'..c3' or w1'..w3')].
long long link_reals(const gft_desc& d) {
    if (d.model == GFT_MODEL_GLOBAL) return 0;
    const int base = d.gen_rep ? 4 : 5, g = gauge_num(d);
    return 3 * (base + (g == 5 ? 1 : g == 7 ? base - 1 : 0));
}

}  // namespace

struct gft_ctx {
    gft_desc desc{};
    Coef coef{};
    int form = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    long long N = 0;
    // two-slot state (form TWO_SLOT_*): scalar[2][N][Ns], vector[2][N][15]
    void* S = nullptr;
    void* V = nullptr;
    // momentum state (form MOMENTUM_*)
    MomentumState mom{};
    double* partial = nullptr;     // per-block reduction partials
    size_t partial_cap = 0;        // doubles
    double* obs_dev = nullptr;     // finalised observables (32 doubles)
    double* obs_host = nullptr;    // pinned mirror
    unsigned* err_dev = nullptr;
    unsigned time_iter = 0;
    bool imported = false;
    // device-side snapshot (gft_snapshot / gft_rollback): a copy of the state
    // arrays and the host-side fields that go with them
    void* snap = nullptr;
    bool snap_valid = false;
    unsigned snap_time_iter = 0;
    int snap_cur = 0, snap_prev_valid = 0;
    double snap_ref[2][4] = {};
    std::string err;
    // step-kernel timing
    std::vector<hipEvent_t> ev_start, ev_stop;
    int ev_used = 0;
    bool seg_open = false;   // a timing segment's start event is queued
    double step_ms_total = 0.0;
    unsigned step_count = 0;
};

#define CK(expr)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                        \
            return GFT_EHIP;                                                                      \
        }                                                                                         \
    } while (0)

static int fail(gft_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg; else g_create_error = msg;
    return code;
}

// ---- allocation-free scratch ---------------------------------------------------
// Reduction partials of gft_observe_noalloc: a file-scope device array that the
// code object carries, one copy per device, held until the process ends.  Host
// code never uses the symbol itself (in host code it is only a stand-in):
// gft_create resolves the device address once per device with that device
// current (scratch_resolve) and every launch takes that pointer.  Every
// gft_observe_noalloc call on one device shares the array: calls that overlap on
// one device (other streams or threads) race on it.
constexpr size_t kScratchDoubles = size_t(1) << 20;   // 8 MiB
constexpr int kScratchMaxDevices = 64;
static __device__ __align__(256) double scratch[kScratchDoubles];
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kScratchMaxDevices] = {};

// Device address of `scratch` on device dev (the current device), resolved on
// the first call for that device.
static int scratch_resolve(int dev, std::string& msg) {
    if (dev < 0 || dev >= kScratchMaxDevices) { msg = "scratch: device index out of range"; return GFT_EINVAL; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return GFT_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        msg = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return GFT_EHIP;
    }
    g_scratch_ptr[dev] = static_cast<double*>(p);
    return GFT_OK;
}

static Coef make_coef(const gft_desc& d) {
    Coef c{};
    const bool global = d.model == GFT_MODEL_GLOBAL;
    c.ns = global ? d.num_scalar : d.model == GFT_MODEL_2HDM ? 8 : 4;
    c.nd = c.ns / 4;
    c.so_n = global ? d.potential != 2 : d.model == GFT_MODEL_SM;
    c.nx = d.nx; c.ny = d.ny; c.nz = d.nz;
    c.nsites = (long long)d.nx * d.ny * d.nz;
    c.dt = d.dt;
    const double sp[3] = {d.dx, d.dy, d.dz};
    for (int a = 0; a < 3; a++) {
        c.lap[a][0] = 1.0 / (sp[a] * sp[a]);
        c.lap[a][1] = -2.0 / (sp[a] * sp[a]);
        c.lap[a][2] = 1.0 / (sp[a] * sp[a]);
        c.est[a] = 1.0 / sp[a];
        c.inv_sp[a] = 1.0 / (sp[a] * sp[a]);
    }
    c.inv_sqr_dt = 1.0 / (d.dt * d.dt);
    const double gs = d.g * d.g, gps = d.gp * d.gp;
    c.inv_g_sqr = gs > d.div_tol ? 1.0 / gs : 0.0;
    c.inv_gp_sqr = gps > d.div_tol ? 1.0 / gps : 0.0;
    c.gsq[0] = (double)(float)gps;
    c.gsq[1] = c.gsq[2] = c.gsq[3] = (double)(float)gs;
    const double gpps = d.gpp * d.gpp;
    c.gsq[4] = c.gsq[5] = c.gsq[6] = (double)(float)gpps;
    c.inv_gpp_sqr = gpps > d.div_tol ? 1.0 / gpps : 0.0;
    c.lambda = d.lambda; c.eta_sqr = d.eta_sqr;
    c.m1 = d.m1; c.m2 = d.m2; c.l1 = d.l1; c.l2 = d.l2; c.l3 = d.l3;
    c.m12r = d.m12r; c.l4p5 = d.l4p5; c.l6r = d.l6r; c.l7r = d.l7r;
    c.m12i = d.m12i; c.l4m5 = d.l4m5; c.l5i = d.l5i; c.l6i = d.l6i; c.l7i = d.l7i;
    c.break_son = d.break_to_son; c.break_uhalfn = d.break_to_uhalfn;
    c.pot_norm = d.pot_norm; c.pot_offset = d.pot_offset;
    c.dvol = d.dx * d.dy * d.dz;
    c.fixed = d.fixed_faces;
    c.tol = d.div_tol;
    c.G = gauge_num(d);
    return c;
}

static int ensure_partial(gft_ctx* ctx, size_t doubles) {
    if (ctx->partial && ctx->partial_cap >= doubles) return GFT_OK;
    if (ctx->partial) hipFree(ctx->partial);
    // The capacity goes with the buffer: a failed growth leaves no buffer and
    // no capacity, so the next call allocates again instead of writing through null.
    ctx->partial = nullptr;
    ctx->partial_cap = 0;
    CK(hipMalloc(&ctx->partial, doubles * sizeof(double)));
    ctx->partial_cap = doubles;
    return GFT_OK;
}

extern "C" {

int gft_device_info(char* name, size_t name_len, int* cu_count) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return GFT_ENODEV;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return GFT_ENODEV;
    if (name && name_len) std::snprintf(name, name_len, "%s", p.gcnArchName);
    if (cu_count) *cu_count = p.multiProcessorCount;
    return GFT_OK;
}

const char* gft_last_error(const gft_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int gft_create(const gft_desc* desc, gft_ctx** out) {
    if (!desc || !out) return fail(nullptr, GFT_EINVAL, "gft_create: null argument");
    *out = nullptr;
    const gft_desc& d = *desc;
    if (d.model != GFT_MODEL_SM && d.model != GFT_MODEL_2HDM && d.model != GFT_MODEL_GLOBAL)
        return fail(nullptr, GFT_EINVAL, "MODEL: The chosen model (" + std::to_string(d.model) + ") is invalid.");
    if (d.form < GFT_FORM_TWO_SLOT_F64 || d.form > GFT_FORM_MOMENTUM_F32_PLAIN)
        return fail(nullptr, GFT_EINVAL, "gft_create: invalid form " + std::to_string(d.form));
    if (d.model == GFT_MODEL_GLOBAL) {
        if (!is_two_slot(d.form))
            return fail(nullptr, GFT_EINVAL, "gft_create: the global O(N) model runs in the two-slot forms");
        if (d.num_scalar < 1 || d.num_scalar > 8)
            return fail(nullptr, GFT_EINVAL, "gft_create: the global O(N) model takes 1 to 8 scalar components");
        if (d.potential < 0 || d.potential > 2 || (d.potential == 2 && d.num_scalar != 8))
            return fail(nullptr, GFT_EINVAL, "gft_create: potential is 1 (SO_N) or 2 (Double_SO_N, 8 components)");
    }
    if (d.nx < 2 || d.ny < 2 || d.nz < 2)
        return fail(nullptr, GFT_EINVAL, "LATTICE:: periodic lattices need at least 2 sites per direction");
    if ((long long)d.nx * d.ny * d.nz > (1LL << 28))
        return fail(nullptr, GFT_EINVAL, "gft_create: at most 2^28 sites per device (32-bit site offsets)");
    if (!(d.dx > 0 && d.dy > 0 && d.dz > 0 && d.dt > 0))
        return fail(nullptr, GFT_EINVAL, "LATTICE:: spacings and time step must be positive");
    if (d.fixed_faces < 0 || d.fixed_faces > 63)
        return fail(nullptr, GFT_EINVAL, "LATTICE:: fixed_faces holds six face bits");
    for (int a = 0; a < 3; a++) {
        const int pair = (d.fixed_faces >> (2 * a)) & 3;
        if (pair == 1 || pair == 2)   // Lattice.cpp:245-260
            return fail(nullptr, GFT_EINVAL, "LATTICE:: Periodic boundary conditions must be assigned to both faces "
                                             "of an axis (axis " + std::to_string(a) + ").");
    }
    if (d.gen_rep != 0 && d.gen_rep != 1)
        return fail(nullptr, GFT_EINVAL, "gft_create: gen_rep is 0 (quaternion links) or 1 (generator form)");
    if (d.gauge_num != 0 && d.gauge_num != 4 && d.gauge_num != 5 && d.gauge_num != 7)
        return fail(nullptr, GFT_EINVAL, "gft_create: gauge_num is 4 (SU(2) x U(1)), 5 or 7 (Higgs-family U(1) or SU(2))");
    const bool family = gauge_num(d) == 5 || gauge_num(d) == 7;
    if (family && d.model != GFT_MODEL_2HDM)
        return fail(nullptr, GFT_EINVAL, "gft_create: a Higgs-family gauge field needs two doublets (GFT_MODEL_2HDM)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(nullptr, GFT_ENODEV, "gft_create: no HIP device");
    gft_ctx* ctx = new gft_ctx();
    ctx->desc = d;
    ctx->form = d.form;
    ctx->coef = make_coef(d);
    ctx->N = ctx->coef.nsites;
    hipGetDevice(&ctx->device);
    auto bail = [&](int rc) { gft_destroy(ctx); return rc; };
    {
        std::string msg;
        if (int rc = scratch_resolve(ctx->device, msg)) { g_create_error = msg; return bail(rc); }
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        g_create_error = "hipStreamCreate failed";
        return bail(GFT_EHIP);
    }
    const size_t esz = is_f64(d.form) ? sizeof(double) : sizeof(float);
    const long long N = ctx->N;
    if (is_two_slot(d.form)) {
        if (hipMalloc(&ctx->S, 2 * N * ctx->coef.ns * esz) != hipSuccess
            || (link_reals(d) && hipMalloc(&ctx->V, 2 * N * link_reals(d) * esz) != hipSuccess)) {
            g_create_error = "gft_create: device allocation failed";
            return bail(GFT_ENOMEM);
        }
    } else {
        int rc = momentum_alloc(ctx->mom, ctx->coef, is_f64(d.form), d.form == GFT_FORM_MOMENTUM_F32);
        if (rc) { g_create_error = "gft_create: device allocation failed"; return bail(rc); }
    }
    if (hipMalloc(&ctx->obs_dev, 64 * sizeof(double)) != hipSuccess
        || hipHostMalloc(&ctx->obs_host, 64 * sizeof(double)) != hipSuccess
        || hipMalloc(&ctx->err_dev, sizeof(unsigned)) != hipSuccess) {
        g_create_error = "gft_create: device allocation failed";
        return bail(GFT_ENOMEM);
    }
    hipMemsetAsync(ctx->err_dev, 0, sizeof(unsigned), ctx->stream);
    ctx->ev_start.resize(kEventPool);
    ctx->ev_stop.resize(kEventPool);
    for (int i = 0; i < kEventPool; i++) {
        if (hipEventCreate(&ctx->ev_start[i]) != hipSuccess || hipEventCreate(&ctx->ev_stop[i]) != hipSuccess) {
            g_create_error = "hipEventCreate failed";
            return bail(GFT_EHIP);
        }
    }
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) { g_create_error = "stream sync failed"; return bail(GFT_EHIP); }
    *out = ctx;
    return GFT_OK;
}

void gft_destroy(gft_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->S) hipFree(ctx->S);
    if (ctx->V) hipFree(ctx->V);
    momentum_free(ctx->mom);
    if (ctx->snap) hipFree(ctx->snap);
    if (ctx->partial) hipFree(ctx->partial);
    if (ctx->obs_dev) hipFree(ctx->obs_dev);
    if (ctx->obs_host) hipHostFree(ctx->obs_host);
    if (ctx->err_dev) hipFree(ctx->err_dev);
    for (auto e : ctx->ev_start) if (e) hipEventDestroy(e);
    for (auto e : ctx->ev_stop) if (e) hipEventDestroy(e);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

int gft_sync(gft_ctx* ctx) {
    if (!ctx) return GFT_EINVAL;
    CK(hipStreamSynchronize(ctx->stream));
    return GFT_OK;
}

}  // extern "C"

// ---- two-slot import/export --------------------------------------------------

// Generator-form links [y, w1, w2, w3] to the quaternion form [y, c0, c1, c2, c3]
// and back, in FP64.  The momentum forms keep quaternions; these run at their
// import and export only.  The forward map is the reference's
// (WilsonLoops/Source/StandardModel.cpp:78-111).  The inverse takes
// |w| = atan2(|c|, c0), equal in exact arithmetic to the reference's
// asin(|c|) with its pi - asin reflection for c0 < 0 (StandardModel.cpp:114-157)
// but well conditioned at |w| = pi/2, where asin(|c|) turns a rounding of |c|
// into an error of eps/|c0| in w (and it also normalises a quaternion that FP32
// storage has left slightly off the unit sphere).
template <class Src>
static void gen_to_quat(const Src* w, double tol, double* q) {
    q[0] = (double)w[0];
    double wm = 0.0;
    for (int k = 1; k < 4; k++) wm += (double)w[k] * (double)w[k];
    wm = std::sqrt(wm);
    q[1] = std::cos(wm);
    q[2] = q[3] = q[4] = 0.0;
    if (wm > tol) {
        for (int k = 1; k < 4; k++) q[1 + k] = std::sin(wm) * (double)w[k] / wm;
    } else {
        q[2] = std::sin(wm);
    }
}
template <class Dst>
static void quat_to_gen(const double* q, double tol, Dst* w) {
    w[0] = (Dst)q[0];
    double cm = 0.0;
    for (int k = 2; k < 5; k++) cm += q[k] * q[k];
    cm = std::sqrt(cm);
    const double ph = std::atan2(cm, q[1]);
    if (cm > tol) {
        for (int k = 0; k < 3; k++) w[1 + k] = (Dst)(ph * q[2 + k] / cm);
    } else {
        w[1] = (Dst)ph;
        w[2] = w[3] = (Dst)0;
    }
}

template <class Src>
static int import_two_slot(gft_ctx* ctx, const Src* scalar, const Src* vec, unsigned time_iter) {
    const long long N = ctx->N;
    const size_t ns = 2 * N * ctx->coef.ns, nv = 2 * N * link_reals(ctx->desc);
    if (!is_two_slot(ctx->form) && ctx->desc.gen_rep) {
        // Momentum forms: the links enter as FP64 quaternions, the scalars exactly.
        const int G = gauge_num(ctx->desc), vg = G == 5 ? 5 : G == 7 ? 7 : 4, vq = G == 5 ? 6 : G == 7 ? 9 : 5;
        std::vector<double> s(scalar, scalar + ns), v(2 * N * 3 * vq);
        for (long long l = 0; l < 2 * N * 3; l++) {
            gen_to_quat(vec + vg * l, ctx->coef.tol, v.data() + vq * l);
            if (G == 5) v[vq * l + 5] = (double)vec[vg * l + 4];
            if (G == 7) {   // the family SU(2) from its generators w' (positions 4..6)
                const Src w[4] = {Src(0), vec[vg * l + 4], vec[vg * l + 5], vec[vg * l + 6]};
                double q[5];
                gen_to_quat(w, ctx->coef.tol, q);
                for (int k = 0; k < 4; k++) v[vq * l + 5 + k] = q[1 + k];
            }
        }
        int rc = momentum_import(ctx->mom, ctx->coef, s.data(), v.data(), time_iter, ctx->stream, ctx->err);
        if (rc) return rc;
        CK(hipStreamSynchronize(ctx->stream));
        ctx->time_iter = time_iter;
        ctx->imported = true;
        return GFT_OK;
    }
    if (is_two_slot(ctx->form)) {
        if (is_f64(ctx->form) == (sizeof(Src) == 8)) {
            CK(hipMemcpyAsync(ctx->S, scalar, ns * sizeof(Src), hipMemcpyHostToDevice, ctx->stream));
            if (nv) CK(hipMemcpyAsync(ctx->V, vec, nv * sizeof(Src), hipMemcpyHostToDevice, ctx->stream));
        } else if (is_f64(ctx->form)) {
            std::vector<double> s(scalar, scalar + ns), v(vec, vec + nv);
            CK(hipMemcpyAsync(ctx->S, s.data(), ns * 8, hipMemcpyHostToDevice, ctx->stream));
            if (nv) CK(hipMemcpyAsync(ctx->V, v.data(), nv * 8, hipMemcpyHostToDevice, ctx->stream));
            CK(hipStreamSynchronize(ctx->stream));
        } else {
            std::vector<float> s(ns), v(nv);
            for (size_t i = 0; i < ns; i++) s[i] = (float)scalar[i];
            for (size_t i = 0; i < nv; i++) v[i] = (float)vec[i];
            CK(hipMemcpyAsync(ctx->S, s.data(), ns * 4, hipMemcpyHostToDevice, ctx->stream));
            if (nv) CK(hipMemcpyAsync(ctx->V, v.data(), nv * 4, hipMemcpyHostToDevice, ctx->stream));
            CK(hipStreamSynchronize(ctx->stream));
        }
    } else {
        int rc = momentum_import(ctx->mom, ctx->coef, scalar, vec, time_iter, ctx->stream, ctx->err);
        if (rc) return rc;
    }
    CK(hipStreamSynchronize(ctx->stream));
    ctx->time_iter = time_iter;
    ctx->imported = true;
    return GFT_OK;
}

template <class Dst>
static int export_two_slot(gft_ctx* ctx, Dst* scalar, Dst* vec, unsigned* time_iter) {
    const long long N = ctx->N;
    const size_t ns = 2 * N * ctx->coef.ns, nv = 2 * N * link_reals(ctx->desc);
    if (!is_two_slot(ctx->form) && ctx->desc.gen_rep) {
        const int G = gauge_num(ctx->desc), vg = G == 5 ? 5 : G == 7 ? 7 : 4, vq = G == 5 ? 6 : G == 7 ? 9 : 5;
        std::vector<double> s(ns), v(2 * N * 3 * vq);
        int rc = momentum_export(ctx->mom, ctx->coef, s.data(), v.data(), ctx->time_iter, ctx->stream, ctx->err);
        if (rc) return rc;
        for (size_t i = 0; i < ns; i++) scalar[i] = (Dst)s[i];
        for (long long l = 0; l < 2 * N * 3; l++) {
            quat_to_gen(v.data() + vq * l, ctx->coef.tol, vec + vg * l);
            if (G == 5) vec[vg * l + 4] = (Dst)v[vq * l + 5];
            if (G == 7) {
                const double q[5] = {0.0, v[vq * l + 5], v[vq * l + 6], v[vq * l + 7], v[vq * l + 8]};
                Dst w[4];
                quat_to_gen(q, ctx->coef.tol, w);
                for (int k = 0; k < 3; k++) vec[vg * l + 4 + k] = w[1 + k];
            }
        }
        if (time_iter) *time_iter = ctx->time_iter;
        return GFT_OK;
    }
    if (is_two_slot(ctx->form)) {
        if (is_f64(ctx->form) == (sizeof(Dst) == 8)) {
            CK(hipMemcpyAsync(scalar, ctx->S, ns * sizeof(Dst), hipMemcpyDeviceToHost, ctx->stream));
            if (nv) CK(hipMemcpyAsync(vec, ctx->V, nv * sizeof(Dst), hipMemcpyDeviceToHost, ctx->stream));
            CK(hipStreamSynchronize(ctx->stream));
        } else if (is_f64(ctx->form)) {
            std::vector<double> s(ns), v(nv);
            CK(hipMemcpyAsync(s.data(), ctx->S, ns * 8, hipMemcpyDeviceToHost, ctx->stream));
            if (nv) CK(hipMemcpyAsync(v.data(), ctx->V, nv * 8, hipMemcpyDeviceToHost, ctx->stream));
            CK(hipStreamSynchronize(ctx->stream));
            for (size_t i = 0; i < ns; i++) scalar[i] = (Dst)s[i];
            for (size_t i = 0; i < nv; i++) vec[i] = (Dst)v[i];
        } else {
            std::vector<float> s(ns), v(nv);
            CK(hipMemcpyAsync(s.data(), ctx->S, ns * 4, hipMemcpyDeviceToHost, ctx->stream));
            if (nv) CK(hipMemcpyAsync(v.data(), ctx->V, nv * 4, hipMemcpyDeviceToHost, ctx->stream));
            CK(hipStreamSynchronize(ctx->stream));
            for (size_t i = 0; i < ns; i++) scalar[i] = (Dst)s[i];
            for (size_t i = 0; i < nv; i++) vec[i] = (Dst)v[i];
        }
    } else {
        int rc = momentum_export(ctx->mom, ctx->coef, scalar, vec, ctx->time_iter, ctx->stream, ctx->err);
        if (rc) return rc;
    }
    if (time_iter) *time_iter = ctx->time_iter;
    return GFT_OK;
}

extern "C" {

int gft_import_two_slot_f32(gft_ctx* ctx, const float* s, const float* v, unsigned it) {
    if (!ctx || !s || (!v && link_reals(ctx->desc))) return GFT_EINVAL;
    return import_two_slot(ctx, s, v, it);
}
int gft_import_two_slot_f64(gft_ctx* ctx, const double* s, const double* v, unsigned it) {
    if (!ctx || !s || (!v && link_reals(ctx->desc))) return GFT_EINVAL;
    return import_two_slot(ctx, s, v, it);
}
int gft_export_two_slot_f32(gft_ctx* ctx, float* s, float* v, unsigned* it) {
    if (!ctx || !s || (!v && link_reals(ctx->desc))) return GFT_EINVAL;
    return export_two_slot(ctx, s, v, it);
}
int gft_export_two_slot_f64(gft_ctx* ctx, double* s, double* v, unsigned* it) {
    if (!ctx || !s || (!v && link_reals(ctx->desc))) return GFT_EINVAL;
    return export_two_slot(ctx, s, v, it);
}

}  // extern "C"

// ---- evolution ----------------------------------------------------------------

// Step timing: one pair of HIP events brackets each run of consecutive step
// launches (a segment), closed before any other kernel is queued; no event
// sits between two steps of a segment, so the steps follow each other as they
// would untimed.  avg = sum of segment times / steps.
static int timed_begin(gft_ctx* ctx) {
    if (ctx->seg_open) return GFT_OK;
    if (ctx->ev_used == kEventPool) {
        CK(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < kEventPool; i++) {
            float ms = 0.f;
            CK(hipEventElapsedTime(&ms, ctx->ev_start[i], ctx->ev_stop[i]));
            ctx->step_ms_total += ms;
        }
        ctx->ev_used = 0;
    }
    CK(hipEventRecord(ctx->ev_start[ctx->ev_used], ctx->stream));
    ctx->seg_open = true;
    return GFT_OK;
}

static int timed_end(gft_ctx* ctx) {
    if (!ctx->seg_open) return GFT_OK;
    CK(hipEventRecord(ctx->ev_stop[ctx->ev_used], ctx->stream));
    ctx->ev_used++;
    ctx->seg_open = false;
    return GFT_OK;
}

static int flush_timing(gft_ctx* ctx) {
    CK(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < ctx->ev_used; i++) {
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, ctx->ev_start[i], ctx->ev_stop[i]));
        ctx->step_ms_total += ms;
    }
    ctx->ev_used = 0;
    return GFT_OK;
}

// Reads the finalised observables into a row.
static int collect_row(gft_ctx* ctx, bool e, bool g, gft_obs* row) {
    CK(hipMemcpyAsync(ctx->obs_host, ctx->obs_dev, 32 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    std::memset(row, 0, sizeof(*row));
    row->time_iter = ctx->time_iter;
    const int G = gauge_num(ctx->desc);
    row->num_eq = G;
    const double dv = ctx->coef.dvol;
    if (e) {
        row->has_energy = 1;
        double tot = 0.0;
        for (int m = 0; m < 5; m++) {
            row->energy[m + 1] = ctx->obs_host[m] * dv;
            tot += ctx->obs_host[m];
        }
        row->energy[0] = tot * dv;
    }
    if (g) {
        row->has_gauss = 1;
        for (int m = 0; m < G; m++) {
            row->gauss_int[m] = ctx->obs_host[8 + m] * dv;
            row->gauss_scale[m] = ctx->obs_host[8 + G + m] * dv;
            row->gauss_max[m] = ctx->obs_host[8 + 2 * G + m];
        }
    }
    return GFT_OK;
}

static long long obs_blocks(const gft_ctx* ctx) { return (ctx->N + kBlock - 1) / kBlock; }

// Finishes the Gauss partials (per block: G integrated |G^a|, G scales, G
// maxima) into obs_dev[8 ..): a row wider than the finisher is done in two
// column slices.
static void finish_gauss(gft_ctx* ctx, const double* partial, long long nblocks) {
    const int G = gauge_num(ctx->desc), W = 3 * G;
    if (W <= kFinWidth) {
        finalize_partials<<<1, kFinThreads, 0, ctx->stream>>>(partial, nblocks, W, 2 * G, G, ctx->obs_dev + 8);
        return;
    }
    const int s1 = std::min(2 * G, kFinWidth);
    finalize_partials<<<1, kFinThreads, 0, ctx->stream>>>(partial, nblocks, W, s1, 0, ctx->obs_dev + 8);
    finalize_partials<<<1, kFinThreads, 0, ctx->stream>>>(partial + s1, nblocks, W, 2 * G - s1, G,
                                                          ctx->obs_dev + 8 + s1);
}

// Two-slot kernels by storage type, link form (quaternion or generator) and
// gauge group (gaugeNum 4, 5, 7).
template <class T, bool GEN, int G>
static void twoslot_step_as(gft_ctx* ctx, long long nblocks, unsigned now, bool me, double damping_step) {
    // The doublet count as a template parameter where that measured faster
    // (quaternion links with gaugeNum 4 or 5; FP32 17-21% at 128^3-256^3), at
    // run time otherwise -- also for the FP64 SM kernel, whose compile-time
    // form is 3% faster at 128^3 but 10% slower at 256^3 (11.03 against
    // 9.95 ms).
    if constexpr (GEN || G == 7) {
        twoslot_step<T, GEN, G, 0><<<nblocks, kBlock, 0, ctx->stream>>>(ctx->coef, (T*)ctx->S, (T*)ctx->V, now, me,
                                                                         damping_step, ctx->partial, ctx->err_dev);
    } else if constexpr (G == 4) {
        if (ctx->coef.nd == 1)
            twoslot_step<T, GEN, G, sizeof(T) == 8 ? 0 : 1><<<nblocks, kBlock, 0, ctx->stream>>>(
                ctx->coef, (T*)ctx->S, (T*)ctx->V, now, me, damping_step, ctx->partial, ctx->err_dev);
        else
            twoslot_step<T, GEN, G, 2><<<nblocks, kBlock, 0, ctx->stream>>>(
                ctx->coef, (T*)ctx->S, (T*)ctx->V, now, me, damping_step, ctx->partial, ctx->err_dev);
    } else {
        twoslot_step<T, GEN, G, 2><<<nblocks, kBlock, 0, ctx->stream>>>(ctx->coef, (T*)ctx->S, (T*)ctx->V, now, me,
                                                                         damping_step, ctx->partial, ctx->err_dev);
    }
}
template <class T, bool GEN, int G>
static void twoslot_gauss_as(gft_ctx* ctx, long long nblocks, unsigned now) {
    twoslot_gauss<T, GEN, G><<<nblocks, kBlock, 0, ctx->stream>>>(ctx->coef, (const T*)ctx->S, (const T*)ctx->V, now,
                                                                   ctx->partial);
}
template <class T, bool GEN>
static void twoslot_dispatch(gft_ctx* ctx, long long nblocks, unsigned now, bool gauss, bool me, double damping_step) {
    switch (gauge_num(ctx->desc)) {
    case 5: gauss ? twoslot_gauss_as<T, GEN, 5>(ctx, nblocks, now)
                  : twoslot_step_as<T, GEN, 5>(ctx, nblocks, now, me, damping_step); break;
    case 7: gauss ? twoslot_gauss_as<T, GEN, 7>(ctx, nblocks, now)
                  : twoslot_step_as<T, GEN, 7>(ctx, nblocks, now, me, damping_step); break;
    default: gauss ? twoslot_gauss_as<T, GEN, 4>(ctx, nblocks, now)
                   : twoslot_step_as<T, GEN, 4>(ctx, nblocks, now, me, damping_step); break;
    }
}
template <class T, int NS>
static void global_step_as(gft_ctx* ctx, long long nblocks, unsigned now, bool me, double damping_step) {
    // nblocks logical blocks over 8 XCD ranges (the kernel's block map).
    twoslot_step_global<T, NS><<<8 * ((nblocks + 7) / 8), kBlock, 0, ctx->stream>>>(
        ctx->coef, (T*)ctx->S, now, me, damping_step, ctx->partial, static_cast<unsigned>(nblocks));
}
// The global model's kernel for its component count (1..8, checked by gft_create).
template <class T>
static void global_step(gft_ctx* ctx, long long nblocks, unsigned now, bool me, double damping_step) {
    switch (ctx->coef.ns) {
    case 1: global_step_as<T, 1>(ctx, nblocks, now, me, damping_step); break;
    case 2: global_step_as<T, 2>(ctx, nblocks, now, me, damping_step); break;
    case 3: global_step_as<T, 3>(ctx, nblocks, now, me, damping_step); break;
    case 4: global_step_as<T, 4>(ctx, nblocks, now, me, damping_step); break;
    case 5: global_step_as<T, 5>(ctx, nblocks, now, me, damping_step); break;
    case 6: global_step_as<T, 6>(ctx, nblocks, now, me, damping_step); break;
    case 7: global_step_as<T, 7>(ctx, nblocks, now, me, damping_step); break;
    default: global_step_as<T, 8>(ctx, nblocks, now, me, damping_step); break;
    }
}
static void twoslot_launch(gft_ctx* ctx, long long nblocks, unsigned now, bool gauss, bool me, double damping_step) {
    const bool f64 = ctx->form == GFT_FORM_TWO_SLOT_F64, gen = ctx->desc.gen_rep != 0;
    if (ctx->desc.model == GFT_MODEL_GLOBAL) {   // no links, no Gauss law
        if (gauss) return;
        (f64 ? global_step<double> : global_step<float>)(ctx, nblocks, now, me, damping_step);
        return;
    }
    if (f64) (gen ? twoslot_dispatch<double, true> : twoslot_dispatch<double, false>)(ctx, nblocks, now, gauss, me,
                                                                                       damping_step);
    else (gen ? twoslot_dispatch<float, true> : twoslot_dispatch<float, false>)(ctx, nblocks, now, gauss, me,
                                                                                 damping_step);
}
static void twoslot_launch_step(gft_ctx* ctx, long long nblocks, unsigned now, bool me, double damping_step) {
    twoslot_launch(ctx, nblocks, now, false, me, damping_step);
}
static void twoslot_launch_gauss(gft_ctx* ctx, long long nblocks, unsigned now) {
    twoslot_launch(ctx, nblocks, now, true, false, 0.0);
}

extern "C" int gft_evolve(gft_ctx* ctx, unsigned nsteps, unsigned energy_freq, unsigned gauss_freq, gft_obs* rows,
                          unsigned max_rows, unsigned* nrows) {
    if (!ctx) return GFT_EINVAL;
    if ((energy_freq || gauss_freq) && !rows && max_rows) return fail(ctx, GFT_EINVAL, "gft_evolve: rows is NULL");
    const long long nblocks = obs_blocks(ctx);
    int rc = ensure_partial(ctx, (size_t)nblocks * 3 * GFT_MAX_EQ);
    if (rc) return rc;
    const bool two = is_two_slot(ctx->form);
    unsigned row = 0;
    ctx->step_ms_total = 0.0;
    ctx->step_count = 0;
    ctx->ev_used = 0;
    ctx->seg_open = false;
    for (unsigned k = 0; k < nsteps; k++) {
        const unsigned it = ctx->time_iter;
        const unsigned now = (it + 1) % 2;
        const bool me = energy_freq && it % energy_freq == 0;
        const bool mg = gauss_freq && it % gauss_freq == 0 && gauge_num(ctx->desc) > 0;
        const float damp = it < ctx->desc.nt_damped ? ctx->desc.damping : 0.f;
        const double damping_step = 0.5 * (double)damp * ctx->desc.dt;
        if (me && !two) {
            // The momentum form measures the state before the update in its own pass.
            if ((rc = timed_end(ctx))) return rc;
            if ((rc = momentum_observe(ctx->mom, ctx->coef, MomObs::Energy, ctx->partial, nblocks, ctx->stream, ctx->err)))
                return rc;
            finalize_partials<<<1, kFinThreads, 0, ctx->stream>>>(ctx->partial, nblocks, 5, 5, 0, ctx->obs_dev);
        }
        if ((rc = timed_begin(ctx))) return rc;
        if (two) {
            twoslot_launch_step(ctx, nblocks, now, me, damping_step);
            CK(hipGetLastError());
        } else {
            if ((rc = momentum_step(ctx->mom, ctx->coef, damping_step, ctx->stream, ctx->err_dev, ctx->err))) return rc;
        }
        ctx->step_count++;
        if ((me && two) || mg)
            if ((rc = timed_end(ctx))) return rc;
        if (me && two) finalize_partials<<<1, kFinThreads, 0, ctx->stream>>>(ctx->partial, nblocks, 5, 5, 0, ctx->obs_dev);
        if (mg) {
            if (two)
                twoslot_launch_gauss(ctx, nblocks, now);
            else if ((rc = momentum_observe(ctx->mom, ctx->coef, MomObs::Gauss, ctx->partial, nblocks, ctx->stream,
                                            ctx->err)))
                return rc;
            finish_gauss(ctx, ctx->partial, nblocks);
            CK(hipGetLastError());
        }
        if ((me || mg) && row < max_rows && rows) {
            if ((rc = collect_row(ctx, me, mg, &rows[row]))) return rc;
            row++;
        }
        ctx->time_iter++;
    }
    if ((rc = timed_end(ctx))) return rc;
    if ((rc = flush_timing(ctx))) return rc;
    if (nrows) *nrows = row;
    unsigned bad = 0;
    CK(hipMemcpy(&bad, ctx->err_dev, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (bad) {
        CK(hipMemset(ctx->err_dev, 0, sizeof(unsigned)));
        return fail(ctx, GFT_EDOMAIN, "Panic in evolve: |R| > 1 in a link update (asin/sqrt out of domain)");
    }
    return GFT_OK;
}

extern "C" int gft_set_launch(gft_ctx* ctx, int bz, int by, int order, int min_waves, int stage) {
    if (!ctx) return GFT_EINVAL;
    if (stage < 0 || stage > 2) return fail(ctx, GFT_EINVAL, "gft_set_launch: stage must be 0, 1 or 2");
    if (min_waves != 0 && min_waves != 3 && min_waves != 4)
        return fail(ctx, GFT_EINVAL, "gft_set_launch: min_waves must be 0, 3 or 4");
    if (order < 0 || order > 4) return fail(ctx, GFT_EINVAL, "gft_set_launch: order must be 0 .. 4");
    if (bz < 1 || by < 1 || bz * by != 256 || (bz & (bz - 1)) || bz < 16)
        return fail(ctx, GFT_EINVAL, "gft_set_launch: need bz*by = 256 with bz a power of two >= 16");
    ctx->mom.bz = bz;
    ctx->mom.by = by;
    ctx->mom.order = order;
    ctx->mom.min_waves = min_waves;
    ctx->mom.stage = stage;
    return GFT_OK;
}

extern "C" int gft_last_step_ms(gft_ctx* ctx, double* avg_ms, double* total_ms, unsigned* nsteps) {
    if (!ctx || ctx->step_count == 0) return GFT_EINVAL;
    if (avg_ms) *avg_ms = ctx->step_ms_total / ctx->step_count;
    if (total_ms) *total_ms = ctx->step_ms_total;
    if (nsteps) *nsteps = ctx->step_count;
    return GFT_OK;
}

// Energy and (after a step) Gauss-law reductions of the current state with the
// per-block partials in `partial` (nblocks x 12 doubles at least).
static int observe_into(gft_ctx* ctx, double* partial, long long nblocks, gft_obs* out) {
    int rc;
    if ((rc = momentum_observe(ctx->mom, ctx->coef, MomObs::Energy, partial, nblocks, ctx->stream, ctx->err)))
        return rc;
    finalize_partials<<<1, kFinThreads, 0, ctx->stream>>>(partial, nblocks, 5, 5, 0, ctx->obs_dev);
    const bool g = ctx->mom.prev_valid;
    if (g) {
        if ((rc = momentum_observe(ctx->mom, ctx->coef, MomObs::Gauss, partial, nblocks, ctx->stream, ctx->err)))
            return rc;
        finish_gauss(ctx, partial, nblocks);
    }
    CK(hipGetLastError());
    return collect_row(ctx, true, g, out);
}

extern "C" int gft_observe(gft_ctx* ctx, gft_obs* out) {
    if (!ctx || !out) return GFT_EINVAL;
    if (is_two_slot(ctx->form))
        return fail(ctx, GFT_EINVAL, "gft_observe: the two-slot forms measure inside gft_evolve (reference cadence)");
    const long long nblocks = obs_blocks(ctx);
    int rc = ensure_partial(ctx, (size_t)nblocks * 12);
    if (rc) return rc;
    return observe_into(ctx, ctx->partial, nblocks, out);
}

extern "C" int gft_observe_noalloc(gft_ctx* ctx, gft_obs* out) {
    if (!ctx || !out) return GFT_EINVAL;
    if (is_two_slot(ctx->form))
        return fail(ctx, GFT_EINVAL, "gft_observe_noalloc: the two-slot forms measure inside gft_evolve");
    const long long nblocks = obs_blocks(ctx);
    if ((size_t)nblocks * 12 > kScratchDoubles)
        return fail(ctx, GFT_EINVAL, "gft_observe_noalloc: the lattice has " + std::to_string(ctx->N) +
                                         " sites; the scratch array holds the partials of at most " +
                                         std::to_string(kScratchDoubles / 12 * kBlock) + " sites");
    double* partial = (ctx->device >= 0 && ctx->device < kScratchMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
    if (!partial) return fail(ctx, GFT_EHIP, "gft_observe_noalloc: scratch address of the device not resolved");
    return observe_into(ctx, partial, nblocks, out);
}

extern "C" int gft_wilson_loops(gft_ctx* ctx, gft_loops* out) {
    if (!ctx || !out) return GFT_EINVAL;
    if (is_two_slot(ctx->form)) return fail(ctx, GFT_EINVAL, "gft_wilson_loops: available for the momentum forms");
    const long long nblocks = obs_blocks(ctx);
    int rc = ensure_partial(ctx, (size_t)nblocks * 12);
    if (rc) return rc;
    if ((rc = momentum_observe(ctx->mom, ctx->coef, MomObs::Loops, ctx->partial, nblocks, ctx->stream, ctx->err)))
        return rc;
    finalize_partials<<<1, kFinThreads, 0, ctx->stream>>>(ctx->partial, nblocks, 5, 5, 0, ctx->obs_dev + 32);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(ctx->obs_host, ctx->obs_dev, 40 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    const double n = (double)ctx->N;
    out->w11_su2 = ctx->obs_host[32] / (3 * n);
    out->w11_u1 = ctx->obs_host[33] / (3 * n);
    out->w22_su2 = ctx->obs_host[34] / (3 * n);
    out->w22_u1 = ctx->obs_host[35] / (3 * n);
    out->phi_sq = ctx->obs_host[36] / n;
    return GFT_OK;
}

// Streaming copy of n float4 elements (grid-stride); traffic = 32 n bytes.
// One 16-byte element per thread: the simplest form reaches the practical HBM
// bandwidth (a grid-stride loop over fewer blocks stays ~25% below it).
__global__ void __launch_bounds__(256) gft_copy_kernel(const float4* __restrict__ src, float4* __restrict__ dst,
                                                       long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

extern "C" int gft_bench_copy(size_t bytes, int reps, double* avg_ms) {
    if (bytes < 16 || reps < 1 || !avg_ms) return GFT_EINVAL;
    const long long n = (long long)(bytes / 16);
    float4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GFT_OK;
    if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) rc = GFT_ENOMEM;
    if (!rc && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess
                || hipEventCreate(&e1) != hipSuccess))
        rc = GFT_EHIP;
    if (!rc && hipMemsetAsync(a, 0, n * 16, st) != hipSuccess) rc = GFT_EHIP;
    if (!rc && n > (long long)INT32_MAX * 256) rc = GFT_EINVAL;
    if (!rc) {
        const int grid = (int)((n + 255) / 256);
        gft_copy_kernel<<<grid, 256, 0, st>>>(a, b, n);
        hipEventRecord(e0, st);
        for (int r = 0; r < reps; r++) gft_copy_kernel<<<grid, 256, 0, st>>>(a, b, n);
        hipEventRecord(e1, st);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = GFT_EHIP;
        *avg_ms = ms / reps;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    if (a) hipFree(a);
    if (b) hipFree(b);
    return rc;
}

// Streams the arrays of one FP32 SM momentum-form step, own site only, with
// the step kernel's access kinds (cached loads of q, y, phi; streaming loads
// of P, E and the residual word; streaming stores), its tiles and block order, and
// trivial arithmetic: the HBM floor of the step's data set without its stencil.
template <bool CMP>
__global__ void __launch_bounds__(256) gft_stream_kernel(StepArgs<float> a) {
    const int4 tile = a.tiles[blockIdx.x];
    const int zs = tile.w * a.bz + (static_cast<int>(threadIdx.x) & (a.bz - 1));
    const int ys = tile.z * a.by + (static_cast<int>(threadIdx.x) >> a.bz_shift);
    if (zs >= a.nz || ys >= a.ny) return;
    const unsigned s = static_cast<unsigned>((tile.y * a.ny + ys) * a.nz + zs);
    float4 q[3], f = at(a.phi[0], s), p = ld_nt(ptr_at(a.P[0], s)), e[3];
    float y[3];
    uint2 lw = make_uint2(0u, 0u);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        q[k] = at(a.q[k], s);
        y[k] = at(a.y[k], s);
        e[k] = ld_nt(ptr_at(a.E[k], s));
    }
    if (CMP) lw = ld_nt(ptr_at(a.lo, s));
    const float t = f.x + p.x + y[0] + y[1] + y[2];
    p.x += t;
    f.x += t;
    st_nt(p, ptr_at(a.P[0], s));
    st_nt(f, ptr_at(a.phi_out[0], s));
    if (CMP) st_nt(make_uint2(lw.x + 1u, lw.y), ptr_at(a.lo, s));
#pragma unroll
    for (int k = 0; k < 3; k++) {
        e[k].x += t;
        q[k].x += t;
        st_nt(e[k], ptr_at(a.E[k], s));
        st_nt(y[k] + t, ptr_at(a.y_out[k], s));
        st_nt(q[k], ptr_at(a.q_out[k], s));
    }
}

extern "C" int gft_bench_streams(gft_ctx* ctx, int reps, double* avg_ms) {
    if (!ctx || reps < 1 || !avg_ms) return GFT_EINVAL;
    if ((ctx->form != GFT_FORM_MOMENTUM_F32 && ctx->form != GFT_FORM_MOMENTUM_F32_PLAIN) || ctx->coef.nd != 1)
        return fail(ctx, GFT_EINVAL, "gft_bench_streams: FP32 momentum-form SM contexts only");
    int rc = GFT_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = GFT_EHIP;
    std::string msg;
    if (!rc && ensure_tiles(ctx->mom, ctx->coef, ctx->stream, msg) != GFT_OK) rc = fail(ctx, GFT_EHIP, msg);
    if (!rc) {
        const StepArgs<float> a = make_step_args<float>(ctx->mom, ctx->mom.ff, ctx->coef, 0.0, ctx->err_dev);
        const long long nb = mom_nblocks(ctx->mom, ctx->coef);
        auto launch = [&]() {
            if (ctx->mom.cmp) gft_stream_kernel<true><<<nb, 256, 0, ctx->stream>>>(a);
            else gft_stream_kernel<false><<<nb, 256, 0, ctx->stream>>>(a);
        };
        launch();
        hipEventRecord(e0, ctx->stream);
        for (int r = 0; r < reps; r++) launch();
        hipEventRecord(e1, ctx->stream);
        float ms = 0.f;
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(ctx, GFT_EHIP, "gft_bench_streams: launch failed");
        *avg_ms = ms / reps;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    ctx->imported = false;   // the state now holds the kernel's scratch values
    return rc;
}

// ---- checkpoint, restore, snapshot and rollback ---------------------------------
// The state of a context is a set of device arrays (two-slot forms: scalar and
// vector; momentum forms: both ping-pong buffers, P, E and the residual words,
// one allocation) plus host-side fields: the step counter, which ping-pong
// buffer is current, whether the previous step's buffer is valid (Gauss law) and
// the reference doublets.  A checkpoint holds all of it, so a restored context
// steps on bitwise as the original would have.

namespace {

constexpr unsigned long long kCkptMagic = 0x31544b5043544647ull;   // "GFTCPKT1"

// Fixed-size header in front of the raw state bytes.
struct CkptHeader {
    unsigned long long magic;
    int form, model, nx, ny, nz, ns, gauge_num, gen_rep, fixed_faces;
    int cur, prev_valid;
    unsigned time_iter;
    double ref[2][4];
    unsigned long long payload;   // bytes of state after the header
};

// The device arrays of the state: up to two (pointer, bytes) ranges.
int state_ranges(const gft_ctx* ctx, void* ptr[2], size_t bytes[2]) {
    const long long N = ctx->N;
    const size_t esz = is_f64(ctx->form) ? 8 : 4;
    if (is_two_slot(ctx->form)) {
        ptr[0] = ctx->S;
        bytes[0] = (size_t)(2 * N * ctx->coef.ns) * esz;
        ptr[1] = ctx->V;
        bytes[1] = ctx->V ? (size_t)(2 * N * link_reals(ctx->desc)) * esz : 0;
    } else {
        ptr[0] = ctx->mom.base;
        bytes[0] = ctx->mom.bytes;
        ptr[1] = nullptr;
        bytes[1] = 0;
    }
    return GFT_OK;
}

size_t state_bytes(const gft_ctx* ctx) {
    void* p[2];
    size_t b[2];
    state_ranges(ctx, p, b);
    return b[0] + b[1];
}

CkptHeader make_header(const gft_ctx* ctx) {
    CkptHeader h{};
    h.magic = kCkptMagic;
    h.form = ctx->form;
    h.model = ctx->desc.model;
    h.nx = ctx->desc.nx; h.ny = ctx->desc.ny; h.nz = ctx->desc.nz;
    h.ns = ctx->coef.ns;
    h.gauge_num = gauge_num(ctx->desc);
    h.gen_rep = ctx->desc.gen_rep;
    h.fixed_faces = ctx->desc.fixed_faces;
    h.cur = ctx->mom.cur;
    h.prev_valid = ctx->mom.prev_valid;
    h.time_iter = ctx->time_iter;
    for (int k = 0; k < 2; k++)
        for (int m = 0; m < 4; m++) h.ref[k][m] = ctx->mom.ref[k][m];
    h.payload = state_bytes(ctx);
    return h;
}

}  // namespace

extern "C" {

int gft_checkpoint_size(const gft_ctx* ctx, size_t* bytes) {
    if (!ctx || !bytes) return GFT_EINVAL;
    *bytes = sizeof(CkptHeader) + state_bytes(ctx);
    return GFT_OK;
}

int gft_checkpoint(gft_ctx* ctx, void* buf, size_t bytes) {
    if (!ctx || !buf) return GFT_EINVAL;
    if (!ctx->imported) return fail(ctx, GFT_EINVAL, "gft_checkpoint: the context holds no state (import or generate one)");
    const CkptHeader h = make_header(ctx);
    if (bytes < sizeof(CkptHeader) + h.payload)
        return fail(ctx, GFT_EINVAL, "gft_checkpoint: buffer smaller than gft_checkpoint_size");
    std::memcpy(buf, &h, sizeof(h));
    void* p[2];
    size_t b[2];
    state_ranges(ctx, p, b);
    char* dst = static_cast<char*>(buf) + sizeof(h);
    for (int r = 0; r < 2; r++) {
        if (!b[r]) continue;
        CK(hipMemcpyAsync(dst, p[r], b[r], hipMemcpyDeviceToHost, ctx->stream));
        dst += b[r];
    }
    CK(hipStreamSynchronize(ctx->stream));
    return GFT_OK;
}

int gft_restore(gft_ctx* ctx, const void* buf, size_t bytes) {
    if (!ctx || !buf) return GFT_EINVAL;
    if (bytes < sizeof(CkptHeader)) return fail(ctx, GFT_EINVAL, "gft_restore: buffer too short for a checkpoint");
    CkptHeader h;
    std::memcpy(&h, buf, sizeof(h));
    const CkptHeader mine = make_header(ctx);
    if (h.magic != kCkptMagic) return fail(ctx, GFT_EINVAL, "gft_restore: not a gft checkpoint");
    if (h.form != mine.form || h.model != mine.model || h.nx != mine.nx || h.ny != mine.ny || h.nz != mine.nz ||
        h.ns != mine.ns || h.gauge_num != mine.gauge_num || h.gen_rep != mine.gen_rep ||
        h.fixed_faces != mine.fixed_faces || h.payload != mine.payload)
        return fail(ctx, GFT_EINVAL, "gft_restore: the checkpoint was taken from a context of another form, model, "
                                     "lattice or boundary");
    if (bytes < sizeof(CkptHeader) + h.payload) return fail(ctx, GFT_EINVAL, "gft_restore: truncated checkpoint");
    void* p[2];
    size_t b[2];
    state_ranges(ctx, p, b);
    const char* src = static_cast<const char*>(buf) + sizeof(h);
    for (int r = 0; r < 2; r++) {
        if (!b[r]) continue;
        CK(hipMemcpyAsync(p[r], src, b[r], hipMemcpyHostToDevice, ctx->stream));
        src += b[r];
    }
    CK(hipStreamSynchronize(ctx->stream));
    ctx->time_iter = h.time_iter;
    ctx->mom.cur = h.cur;
    ctx->mom.prev_valid = h.prev_valid;
    for (int k = 0; k < 2; k++)
        for (int m = 0; m < 4; m++) ctx->mom.ref[k][m] = h.ref[k][m];
    ctx->mom.epoch++;
    ctx->imported = true;
    return GFT_OK;
}

int gft_snapshot(gft_ctx* ctx) {
    if (!ctx) return GFT_EINVAL;
    if (!ctx->imported) return fail(ctx, GFT_EINVAL, "gft_snapshot: the context holds no state (import or generate one)");
    void* p[2];
    size_t b[2];
    state_ranges(ctx, p, b);
    if (!ctx->snap && hipMalloc(&ctx->snap, b[0] + b[1]) != hipSuccess) {
        ctx->snap = nullptr;
        return fail(ctx, GFT_ENOMEM, "gft_snapshot: device allocation failed");
    }
    char* dst = static_cast<char*>(ctx->snap);
    for (int r = 0; r < 2; r++) {
        if (!b[r]) continue;
        CK(hipMemcpyAsync(dst, p[r], b[r], hipMemcpyDeviceToDevice, ctx->stream));
        dst += b[r];
    }
    CK(hipStreamSynchronize(ctx->stream));
    ctx->snap_valid = true;
    ctx->snap_time_iter = ctx->time_iter;
    ctx->snap_cur = ctx->mom.cur;
    ctx->snap_prev_valid = ctx->mom.prev_valid;
    for (int k = 0; k < 2; k++)
        for (int m = 0; m < 4; m++) ctx->snap_ref[k][m] = ctx->mom.ref[k][m];
    return GFT_OK;
}

int gft_rollback(gft_ctx* ctx, unsigned* time_iter) {
    if (!ctx) return GFT_EINVAL;
    if (!ctx->snap_valid) return fail(ctx, GFT_EINVAL, "gft_rollback: no snapshot (gft_snapshot first)");
    void* p[2];
    size_t b[2];
    state_ranges(ctx, p, b);
    const char* src = static_cast<const char*>(ctx->snap);
    for (int r = 0; r < 2; r++) {
        if (!b[r]) continue;
        CK(hipMemcpyAsync(p[r], src, b[r], hipMemcpyDeviceToDevice, ctx->stream));
        src += b[r];
    }
    CK(hipStreamSynchronize(ctx->stream));
    ctx->time_iter = ctx->snap_time_iter;
    ctx->mom.cur = ctx->snap_cur;
    ctx->mom.prev_valid = ctx->snap_prev_valid;
    for (int k = 0; k < 2; k++)
        for (int m = 0; m < 4; m++) ctx->mom.ref[k][m] = ctx->snap_ref[k][m];
    ctx->mom.epoch++;
    ctx->imported = true;
    if (time_iter) *time_iter = ctx->time_iter;
    return GFT_OK;
}

}  // extern "C"

// ---- batched ensembles -----------------------------------------------------------

struct gft_ensemble {
    std::vector<gft_ctx*> members;
    hipStream_t stream = nullptr;
    EnsMember* args[2] = {nullptr, nullptr};   // step arguments with buffer 0 / 1 current
    int4* work = nullptr;
    long long nwork = 0;
    long long grid = 0;      // blocks of the ensemble kernel: one per tile, 8 x the longest XCD range
    size_t lds = 0;          // dynamic LDS per block of the staged kernel (0: unstaged)
    int stage = 0;           // staging mode of the ensemble kernel
    unsigned* err_dev = nullptr;
    bool dirty = true;
    std::vector<int4> launch_key;   // (bz, by, stage, state epoch) of every member at the last prepare
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_parity = -1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double last_ms = 0.0;
    unsigned last_steps = 0;
    std::string err;
};

static int efail(gft_ensemble* e, int code, const std::string& msg) {
    if (e) e->err = msg; else g_create_error = msg;
    return code;
}

#define EK(expr)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return efail(ens, GFT_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void ens_drop_graph(gft_ensemble* ens) {
    if (ens->exec) (void)hipGraphExecDestroy(ens->exec);
    if (ens->graph) (void)hipGraphDestroy(ens->graph);
    ens->exec = nullptr;
    ens->graph = nullptr;
    ens->graph_parity = -1;
}

// Uploads the per-member step arguments (both parities) and the work list.
static int ens_prepare(gft_ensemble* ens) {
    // A member's launch shape may also change through its own context (gft_set_launch),
    // and its reference doublets (part of the step arguments) with every import.
    std::vector<int4> key;
    for (auto c : ens->members) key.push_back(make_int4(c->mom.bz, c->mom.by, c->mom.stage, (int)c->mom.epoch));
    const bool same = key.size() == ens->launch_key.size() &&
                      std::equal(key.begin(), key.end(), ens->launch_key.begin(), [](const int4& p, const int4& q) {
                          return p.x == q.x && p.y == q.y && p.z == q.z && p.w == q.w;
                      });
    if (!same) ens->dirty = true;
    if (!ens->dirty) return GFT_OK;
    ens->launch_key = key;
    ens_drop_graph(ens);
    const size_t nm = ens->members.size();
    std::vector<EnsMember> host[2];
    std::vector<int4> work;
    // The staging mode every member can run (the lowest over the members).
    int stage = 2;
    for (size_t i = 0; i < nm; i++) {
        gft_ctx* c = ens->members[i];
        stage = std::min(stage, step_stage_mode(c->mom.stage, c->mom.bz, c->mom.by, 1));
    }
    size_t lds = 0;
    for (size_t i = 0; i < nm; i++)
        lds = std::max(lds, step_lds_bytes(ens->members[i]->mom.bz, ens->members[i]->mom.by, 1, stage));
    for (size_t i = 0; i < nm; i++) {
        gft_ctx* c = ens->members[i];
        const int saved = c->mom.cur;
        for (int p = 0; p < 2; p++) {
            c->mom.cur = p;
            EnsMember em{};
            em.a = make_step_args<float>(c->mom, c->mom.ff, c->coef, 0.0, ens->err_dev);
            em.a.tiles = nullptr;
            em.c = c->coef;
            host[p].push_back(em);
        }
        c->mom.cur = saved;
        const std::vector<int4> t = build_tiles_banded((int)i, c->coef.nx, c->coef.ny, c->coef.nz, c->mom.bz, c->mom.by);
        work.insert(work.end(), t.begin(), t.end());
    }
    for (int p = 0; p < 2; p++) {
        if (!ens->args[p]) EK(hipMalloc(&ens->args[p], nm * sizeof(EnsMember)));
        EK(hipMemcpyAsync(ens->args[p], host[p].data(), nm * sizeof(EnsMember), hipMemcpyHostToDevice, ens->stream));
    }
    if (ens->work) (void)hipFree(ens->work);
    ens->work = nullptr;
    EK(hipMalloc(&ens->work, work.size() * sizeof(int4)));
    EK(hipMemcpyAsync(ens->work, work.data(), work.size() * sizeof(int4), hipMemcpyHostToDevice, ens->stream));
    EK(hipStreamSynchronize(ens->stream));
    ens->nwork = (long long)work.size();
    ens->lds = lds;
    ens->stage = stage;
    // XCD k owns work items [k n / 8, (k + 1) n / 8), at most ceil(n / 8) of them.
    ens->grid = 8 * ((ens->nwork + 7) / 8);
    ens->dirty = false;
    return GFT_OK;
}

extern "C" {

int gft_ensemble_create(const gft_desc* descs, int nmembers, gft_ensemble** out) {
    if (!descs || nmembers < 1 || !out) return efail(nullptr, GFT_EINVAL, "gft_ensemble_create: bad arguments");
    *out = nullptr;
    for (int i = 0; i < nmembers; i++)
        if (descs[i].form != GFT_FORM_MOMENTUM_F32 || descs[i].model != GFT_MODEL_SM)
            return efail(nullptr, GFT_EINVAL, "gft_ensemble_create: members must be FP32 momentum-form SM lattices");
        else if (descs[i].damping != 0.f && descs[i].nt_damped > 0)
            return efail(nullptr, GFT_EINVAL, "gft_ensemble_create: damped members are not supported");
    gft_ensemble* ens = new gft_ensemble();
    auto bail = [&](int rc) { gft_ensemble_destroy(ens); return rc; };
    for (int i = 0; i < nmembers; i++) {
        gft_ctx* c = nullptr;
        int rc = gft_create(&descs[i], &c);
        if (rc) return bail(rc);
        ens->members.push_back(c);
    }
    if (hipStreamCreateWithFlags(&ens->stream, hipStreamNonBlocking) != hipSuccess
        || hipMalloc(&ens->err_dev, sizeof(unsigned)) != hipSuccess || hipEventCreate(&ens->e0) != hipSuccess
        || hipEventCreate(&ens->e1) != hipSuccess) {
        g_create_error = "gft_ensemble_create: HIP resources";
        return bail(GFT_EHIP);
    }
    if (hipMemset(ens->err_dev, 0, sizeof(unsigned)) != hipSuccess) return bail(GFT_EHIP);
    *out = ens;
    return GFT_OK;
}

void gft_ensemble_destroy(gft_ensemble* ens) {
    if (!ens) return;
    if (ens->stream) (void)hipStreamSynchronize(ens->stream);
    ens_drop_graph(ens);
    for (auto c : ens->members) gft_destroy(c);
    for (int p = 0; p < 2; p++) if (ens->args[p]) (void)hipFree(ens->args[p]);
    if (ens->work) (void)hipFree(ens->work);
    if (ens->err_dev) (void)hipFree(ens->err_dev);
    if (ens->e0) (void)hipEventDestroy(ens->e0);
    if (ens->e1) (void)hipEventDestroy(ens->e1);
    if (ens->stream) (void)hipStreamDestroy(ens->stream);
    delete ens;
}

const char* gft_ensemble_last_error(const gft_ensemble* ens) { return ens ? ens->err.c_str() : g_create_error.c_str(); }

int gft_ensemble_member(gft_ensemble* ens, int member, gft_ctx** ctx) {
    if (!ens || !ctx || member < 0 || member >= (int)ens->members.size())
        return efail(ens, GFT_EINVAL, "gft_ensemble_member: no such member");
    *ctx = ens->members[member];
    return GFT_OK;
}

int gft_ensemble_set_launch(gft_ensemble* ens, int member, int bz, int by) {
    gft_ctx* c = nullptr;
    int rc = gft_ensemble_member(ens, member, &c);
    if (rc) return rc;
    rc = gft_set_launch(c, bz, by, c->mom.order, c->mom.min_waves, c->mom.stage);
    if (rc) return efail(ens, rc, c->err);
    ens->dirty = true;
    return GFT_OK;
}

int gft_ensemble_evolve(gft_ensemble* ens, unsigned nsteps, int use_graph) {
    if (!ens) return GFT_EINVAL;
    const int parity = ens->members[0]->mom.cur;
    for (auto c : ens->members)
        if (c->mom.cur != parity || c->time_iter != ens->members[0]->time_iter)
            return efail(ens, GFT_EINVAL, "gft_ensemble_evolve: members are not at the same step (import them all first)");
    for (auto c : ens->members) {
        if (!c->imported) return efail(ens, GFT_EINVAL, "gft_ensemble_evolve: import every member first");
        if (gft_sync(c) != GFT_OK) return efail(ens, GFT_EHIP, c->err);
    }
    int rc = ens_prepare(ens);
    if (rc) return rc;
    const dim3 grid((unsigned)ens->grid), block(256);
    auto launch = [&](int p) {
        if (ens->stage == 2) ens_step<4, 2><<<grid, block, ens->lds, ens->stream>>>(ens->args[p], ens->work, ens->nwork);
        else if (ens->stage == 1) ens_step<4, 1><<<grid, block, ens->lds, ens->stream>>>(ens->args[p], ens->work, ens->nwork);
        else ens_step<4, 0><<<grid, block, 0, ens->stream>>>(ens->args[p], ens->work, ens->nwork);
    };
    if (use_graph && nsteps >= 2 && ens->graph_parity != parity) {
        ens_drop_graph(ens);
        EK(hipStreamBeginCapture(ens->stream, hipStreamCaptureModeThreadLocal));
        launch(parity);
        launch(!parity);
        EK(hipStreamEndCapture(ens->stream, &ens->graph));
        EK(hipGraphInstantiate(&ens->exec, ens->graph, nullptr, nullptr, 0));
        ens->graph_parity = parity;
    }
    EK(hipEventRecord(ens->e0, ens->stream));
    unsigned k = 0;
    if (use_graph && nsteps >= 2)
        for (; k + 2 <= nsteps; k += 2) EK(hipGraphLaunch(ens->exec, ens->stream));
    for (; k < nsteps; k++) {
        launch((parity + k) % 2);
        EK(hipGetLastError());
    }
    EK(hipEventRecord(ens->e1, ens->stream));
    EK(hipEventSynchronize(ens->e1));
    float ms = 0.f;
    EK(hipEventElapsedTime(&ms, ens->e0, ens->e1));
    ens->last_ms = ms;
    ens->last_steps = nsteps;
    for (auto c : ens->members) {
        c->mom.cur = (parity + nsteps) % 2;
        if (nsteps) c->mom.prev_valid = 1;
        c->time_iter += nsteps;
    }
    unsigned bad = 0;
    EK(hipMemcpy(&bad, ens->err_dev, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (bad) {
        EK(hipMemset(ens->err_dev, 0, sizeof(unsigned)));
        return efail(ens, GFT_EDOMAIN, "Panic in evolve: |R| > 1 in a link update (asin/sqrt out of domain)");
    }
    return GFT_OK;
}

int gft_ensemble_last_step_ms(gft_ensemble* ens, double* avg_ms) {
    if (!ens || !avg_ms || !ens->last_steps) return GFT_EINVAL;
    *avg_ms = ens->last_ms / ens->last_steps;
    return GFT_OK;
}

}  // extern "C"

// ---- Gaussian-random-field initial states on the GPU ----------------------------

namespace gft {

__global__ void __launch_bounds__(256) grf_sumsq4(const double* __restrict__ f, long long n, double* __restrict__ partial) {
    __shared__ double lds[4 * 1];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double v[1] = {0.0};
    if (s < n)
        for (int c = 0; c < 4; c++) v[0] += f[c * n + s] * f[c * n + s];
    block_reduce_store<1, 0>(v, nullptr, lds, partial, blockIdx.x);
}

// phi_k = (0, 0, vev_k, 0) + scale_k * comps, with scale_k = rms / sqrt(<sum_c f_c^2>)
// from sq[k] (sum over sites), stored as phi_k - ref_k (ref_k = (0, 0, vevr_k, 0),
// the vev rounded to FP32 for the FP32 forms, zero for FP64); P = 0, E = 0,
// links identity.
template <class R>
__global__ void __launch_bounds__(256) grf_assemble_mom(MomFields<R> f, const double* __restrict__ comps, long long n,
                                                        int nd, double rms, double vev0, double vev1, double vevr0,
                                                        double vevr1, const double* __restrict__ sq) {
    using V4 = typename V4T<R>::type;
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    for (int k = 0; k < nd; k++) {
        const double scale = sq ? rms / sqrt(sq[k] / (double)n) : rms;   // no sq: comps are scaled by rms as they are
        double v[4];
        for (int m = 0; m < 4; m++)
            v[m] = (m == 2 ? (k ? vev1 - vevr1 : vev0 - vevr0) : 0.0) + scale * comps[(4 * k + m) * n + s];
        V4 hi;
        hi.x = (R)v[0]; hi.y = (R)v[1]; hi.z = (R)v[2]; hi.w = (R)v[3];
        f.phi[0][k][s] = hi;
        V4 z;
        z.x = z.y = z.z = z.w = R(0);
        f.P[k][s] = z;
    }
    for (int i = 0; i < 3; i++) {
        V4 e, q;
        e.x = e.y = e.z = e.w = R(0);
        q.x = R(1); q.y = q.z = q.w = R(0);
        f.E[i][s] = e;
        f.q[0][i][s] = q;
        f.y[0][i][s] = R(0);
    }
    if (f.lo) f.lo[s] = make_uint2(0u, 0u);   // P = 0, E = 0: no residuals
    if (f.lo_P2) f.lo_P2[s] = 0;
    // Higgs-family links at the identity with no electric field.
    for (int i = 0; i < 3; i++) {
        if (f.fy[0][i]) {
            f.fy[0][i][s] = R(0);
            f.fe[i][s] = R(0);
        }
        if (f.fq[0][i]) {
            V4 e, q;
            e.x = e.y = e.z = e.w = R(0);
            q.x = R(1); q.y = q.z = q.w = R(0);
            f.fq[0][i][s] = q;
            f.fE[i][s] = e;
        }
    }
    if (f.flo5) f.flo5[s] = 0u;
    if (f.flo7) f.flo7[s] = make_uint2(0u, 0u);
}

template <class T>
__global__ void __launch_bounds__(256) grf_assemble_two_slot(T* __restrict__ S, T* __restrict__ V,
                                                             const double* __restrict__ comps, long long n, int nd,
                                                             double rms, double vev0, double vev1,
                                                             const double* __restrict__ sq, int vd) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const int ns = 4 * nd;
    for (int k = 0; k < nd; k++) {
        const double scale = sq ? rms / sqrt(sq[k] / (double)n) : rms;
        for (int m = 0; m < 4; m++) {
            const T v = (T)((m == 2 ? (k ? vev1 : vev0) : 0.0) + scale * comps[(4 * k + m) * n + s]);
            S[s * ns + 4 * k + m] = v;
            S[(n + s) * ns + 4 * k + m] = v;
        }
    }
    // Identity links: c0 = 1 in the quaternion form, w = 0 in the generator form
    // (Lattice.cpp:369-392).
    for (int slot = 0; slot < 2; slot++)
        for (int i = 0; i < 3; i++)
            for (int c = 0; c < vd; c++) V[(slot * n + s) * 3 * vd + vd * i + c] = (T)(vd == 5 && c == 1 ? 1.0 : 0.0);
}

}  // namespace gft

extern "C" int gft_init_grf(gft_ctx* ctx, const gft_grf_desc* g) {
    if (!ctx || !g) return GFT_EINVAL;
    if (ctx->desc.model == GFT_MODEL_GLOBAL)
        return fail(ctx, GFT_EINVAL, "gft_init_grf: doublet fields only (not the global O(N) model)");
    const Coef& c = ctx->coef;
    FftAxis ax[3];
    if (!grf_plan(c.nz, ax[0]) || !grf_plan(c.ny, ax[1]) || !grf_plan(c.nx, ax[2]))
        return fail(ctx, GFT_EINVAL, "gft_init_grf: every side must be 2^a 3^b and at most 256");
    if (!(g->rms >= 0.0)) return fail(ctx, GFT_EINVAL, "gft_init_grf: rms must be non-negative");
    const long long N = ctx->N;
    ax[0].es = 1; ax[0].inner = 1; ax[0].outer = c.nz; ax[0].inner_stride = 0;                 // z lines
    ax[1].es = c.nz; ax[1].inner = c.nz; ax[1].outer = (long long)c.ny * c.nz; ax[1].inner_stride = 1;   // y lines
    ax[2].es = (long long)c.ny * c.nz; ax[2].inner = (long long)c.ny * c.nz; ax[2].outer = 0; ax[2].inner_stride = 1;  // x lines
    const long long lines[3] = {(long long)c.nx * c.ny, (long long)c.nx * c.nz, (long long)c.ny * c.nz};
    const gft_desc& dd = ctx->desc;
    const double kmax = g->kmax > 0 ? g->kmax : 3.14159265358979323846 / (2.0 * dd.dx);
    const int nc = 4 * c.nd;
    const long long nb = (N + 255) / 256;
    int rc = ensure_partial(ctx, (size_t)nb * 2);
    if (rc) return rc;
    double2* w = nullptr;
    double* comps = nullptr;
    auto cleanup = [&]() {
        if (w) (void)hipFree(w);
        if (comps) (void)hipFree(comps);
    };
    if (hipMalloc(&w, N * sizeof(double2)) != hipSuccess || hipMalloc(&comps, nc * N * sizeof(double)) != hipSuccess) {
        cleanup();
        return fail(ctx, GFT_ENOMEM, "gft_init_grf: work arrays");
    }
    hipStream_t st = ctx->stream;
    for (int comp = 0; comp < nc; comp++) {
        grf_noise<<<nb, 256, 0, st>>>(w, N, g->seed, (unsigned)comp);
        for (int a = 0; a < 3; a++) grf_fft_lines<<<(unsigned)lines[a], 256, 0, st>>>(w, ax[a], -1);
        grf_filter<<<nb, 256, 0, st>>>(w, c.nx, c.ny, c.nz, dd.dx, dd.dy, dd.dz, g->slope, kmax);
        for (int a = 0; a < 3; a++) grf_fft_lines<<<(unsigned)lines[a], 256, 0, st>>>(w, ax[a], +1);
        grf_real_stats<<<nb, 256, 0, st>>>(w, comps + comp * N, N, ctx->partial);
        finalize_partials<<<1, kFinThreads, 0, st>>>(ctx->partial, nb, 2, 2, 0, ctx->obs_dev);
        grf_normalise<<<nb, 256, 0, st>>>(comps + comp * N, N, ctx->obs_dev, ctx->partial);
    }
    for (int k = 0; k < c.nd; k++) {
        grf_sumsq4<<<nb, 256, 0, st>>>(comps + 4 * k * N, N, ctx->partial);
        finalize_partials<<<1, kFinThreads, 0, st>>>(ctx->partial, nb, 1, 1, 0, ctx->obs_dev + 2 + k);
    }
    const double vev0 = g->vev[0], vev1 = g->vev[1];
    if (is_two_slot(ctx->form)) {
        if (is_f64(ctx->form))
            grf_assemble_two_slot<double><<<nb, 256, 0, st>>>((double*)ctx->S, (double*)ctx->V, comps, N, c.nd, g->rms,
                                                                vev0, vev1, ctx->obs_dev + 2,
                                                                (int)link_reals(ctx->desc) / 3);
        else
            grf_assemble_two_slot<float><<<nb, 256, 0, st>>>((float*)ctx->S, (float*)ctx->V, comps, N, c.nd, g->rms,
                                                               vev0, vev1, ctx->obs_dev + 2,
                                                               (int)link_reals(ctx->desc) / 3);
    } else {
        ctx->mom.cur = 0;
        ctx->mom.prev_valid = 0;
        const double vr0 = ctx->mom.f64 ? 0.0 : (double)(float)vev0, vr1 = ctx->mom.f64 ? 0.0 : (double)(float)vev1;
        for (int k = 0; k < 2; k++)
            for (int m4 = 0; m4 < 4; m4++) ctx->mom.ref[k][m4] = (m4 == 2 && k < c.nd) ? (k ? vr1 : vr0) : 0.0;
        ctx->mom.epoch++;
        if (ctx->mom.f64)
            grf_assemble_mom<double><<<nb, 256, 0, st>>>(ctx->mom.fd, comps, N, c.nd, g->rms, vev0, vev1, vr0, vr1,
                                                         ctx->obs_dev + 2);
        else
            grf_assemble_mom<float><<<nb, 256, 0, st>>>(ctx->mom.ff, comps, N, c.nd, g->rms, vev0, vev1, vr0, vr1,
                                                        ctx->obs_dev + 2);
    }
    const bool ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    cleanup();
    if (!ok) return fail(ctx, GFT_EHIP, "gft_init_grf: kernel failure");
    ctx->time_iter = 0;
    ctx->imported = true;
    return GFT_OK;
}

namespace gft {

// comps[c * n + s] = standard normal of (site s, component c): the GRF's white
// noise, unfiltered.
__global__ void __launch_bounds__(256) gauss_comps(double* __restrict__ comps, long long n, int nc,
                                                   unsigned long long seed) {
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    for (int c = 0; c < nc; c++) comps[c * n + s] = grf_normal(seed, (unsigned)s, (unsigned)c);
}

}  // namespace gft

extern "C" int gft_init_gaussian(gft_ctx* ctx, const gft_gauss_desc* g) {
    if (!ctx || !g) return GFT_EINVAL;
    if (ctx->desc.model == GFT_MODEL_GLOBAL)
        return fail(ctx, GFT_EINVAL, "gft_init_gaussian: doublet fields only (not the global O(N) model)");
    if (!(g->sigma >= 0.0)) return fail(ctx, GFT_EINVAL, "gft_init_gaussian: sigma must be non-negative");
    const Coef& c = ctx->coef;
    const long long N = ctx->N;
    const int nc = 4 * c.nd;
    const long long nb = (N + 255) / 256;
    double* comps = nullptr;
    if (hipMalloc(&comps, nc * N * sizeof(double)) != hipSuccess)
        return fail(ctx, GFT_ENOMEM, "gft_init_gaussian: work array");
    hipStream_t st = ctx->stream;
    gauss_comps<<<nb, 256, 0, st>>>(comps, N, nc, g->seed);
    const double vev0 = g->vev[0], vev1 = g->vev[1];
    if (is_two_slot(ctx->form)) {
        if (is_f64(ctx->form))
            grf_assemble_two_slot<double><<<nb, 256, 0, st>>>((double*)ctx->S, (double*)ctx->V, comps, N, c.nd,
                                                                g->sigma, vev0, vev1, nullptr,
                                                                (int)link_reals(ctx->desc) / 3);
        else
            grf_assemble_two_slot<float><<<nb, 256, 0, st>>>((float*)ctx->S, (float*)ctx->V, comps, N, c.nd,
                                                               g->sigma, vev0, vev1, nullptr,
                                                               (int)link_reals(ctx->desc) / 3);
    } else {
        ctx->mom.cur = 0;
        ctx->mom.prev_valid = 0;
        const double vr0 = ctx->mom.f64 ? 0.0 : (double)(float)vev0, vr1 = ctx->mom.f64 ? 0.0 : (double)(float)vev1;
        for (int k = 0; k < 2; k++)
            for (int m4 = 0; m4 < 4; m4++) ctx->mom.ref[k][m4] = (m4 == 2 && k < c.nd) ? (k ? vr1 : vr0) : 0.0;
        ctx->mom.epoch++;
        if (ctx->mom.f64)
            grf_assemble_mom<double><<<nb, 256, 0, st>>>(ctx->mom.fd, comps, N, c.nd, g->sigma, vev0, vev1, vr0, vr1,
                                                         nullptr);
        else
            grf_assemble_mom<float><<<nb, 256, 0, st>>>(ctx->mom.ff, comps, N, c.nd, g->sigma, vev0, vev1, vr0, vr1,
                                                        nullptr);
    }
    const bool ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(st) == hipSuccess;
    (void)hipFree(comps);
    if (!ok) return fail(ctx, GFT_EHIP, "gft_init_gaussian: kernel failure");
    ctx->time_iter = 0;
    ctx->imported = true;
    return GFT_OK;
}
