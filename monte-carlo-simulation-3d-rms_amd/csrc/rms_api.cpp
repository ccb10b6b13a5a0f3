// rms_api.cpp -- the C ABI of librms_amd.so (include/rms_amd.h).
//
// Host side of the engine: validation, the derived parameters of
// hpc_code/lib/main.cu:429-458 and :484-487, PGSE lobe tables
// (main_PGSE.cu:341-345), device buffers, launch sequencing (optionally as a
// captured hipGraph) and the copies back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/rms_amd.h"
#include "rms_internal.h"
#include "rms_types.h"

using namespace rms;

struct GeomDev {
    rms_geometry meta;     // labels / label_class pointers cleared
    uint64_t *d_cells = nullptr;  // cell words (straight-move outcomes per crossing pattern)
    uint8_t *d_cls = nullptr;     // class per voxel
    int64_t nvox = 0;
    bool has_walkable = false;
};

struct rms_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int ncu = 0;
    std::string err;
    std::vector<GeomDev> geoms;
};

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t n = 0;
};

struct JobHost {
    bool exchange;      // walkers can cross between different non-null classes
    int32_t geom_id;
    int64_t nwalk, nbatch, batch0, wbase;
    int32_t nbvec, ngrad, nval, nbound, ntriple;
    std::vector<int32_t> gperm;
};

}  // namespace

struct rms_plan {
    rms_ctx *ctx = nullptr;
    rms_run_params prm{};
    std::vector<JobHost> jobs;
    int64_t nbatch = 0, ws = 0;
    int sig = 0;
    int grid = 0, block = 256;
    int xcd_split = 0;
    std::vector<void *> allocs;
    JobDesc *d_jobs = nullptr;
    int32_t *d_batch_job = nullptr;
    ReduceDesc *d_red = nullptr;
    std::vector<ReduceDesc> red;   // host copy (device pointers)
    int64_t max_nq = 0;
    unsigned long long *d_counters = nullptr;
    StateBuf st{};
    double *d_snap = nullptr;
    FinalOut fin{};
    bool fin_alloc = false;
    std::vector<JobDesc> hjobs;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int maxbound = 0;
    // adjoint pass buffers, allocated by the first rms_plan_adjoint call
    bool adj_ready = false;
    int adj_np = 1, adj_grid = 0;
    JobDesc *d_jobs_adj = nullptr;
    ReduceDesc *d_red_adj = nullptr;
    std::vector<ReduceDesc> red_adj;
    std::vector<double *> cot_rec, cot_sig;
    double *d_tsnap = nullptr;
};

static const char *kVersion = "rms_amd 0.1 (gfx950)";

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) {                                                                    \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                        \
            return e_ == hipErrorOutOfMemory ? RMS_ERR_OUT_OF_MEMORY : RMS_ERR_HIP;               \
        }                                                                                          \
    } while (0)

static rms_status fail(rms_ctx *ctx, rms_status s, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    return s;
}

// C++ exceptions (host allocation failures in the std containers) must not
// cross the C ABI: every entry point that allocates on the host runs its body
// through guard(), which turns them into status codes.
template <typename F>
static rms_status guard(rms_ctx *ctx, F &&body) noexcept
{
    const char *msg;
    rms_status s;
    try {
        return body();
    } catch (const std::bad_alloc &) {
        msg = "host memory allocation failed";
        s = RMS_ERR_OUT_OF_MEMORY;
    } catch (...) {
        msg = "unexpected host exception";
        s = RMS_ERR_STATE;
    }
    if (ctx) {
        try {
            ctx->err = msg;
        } catch (...) {
        }
    }
    return s;
}

extern "C" const char *rms_version(void) { return kVersion; }

extern "C" int32_t rms_nval(int32_t nclass, int32_t nbvec) { return 1 + nclass + 6 + 15 + nbvec; }

static rms_status create_impl(const rms_device_opts *opts, rms_ctx **out)
{
    if (!out) return RMS_ERR_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return RMS_ERR_NO_DEVICE;
    int dev = opts ? opts->device : 0;
    if (dev < 0 || dev >= ndev) return RMS_ERR_INVALID_ARG;
    rms_ctx *c = new rms_ctx();
    c->device = dev;
    if (hipSetDevice(dev) != hipSuccess) {
        delete c;
        return RMS_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        delete c;
        return RMS_ERR_HIP;
    }
    c->ncu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return RMS_ERR_HIP;
    }
    *out = c;
    return RMS_OK;
}

extern "C" rms_status rms_create(const rms_device_opts *opts, rms_ctx **out)
{
    return guard(nullptr, [&] { return create_impl(opts, out); });
}

extern "C" rms_status rms_destroy(rms_ctx *ctx)
{
    if (!ctx) return RMS_ERR_INVALID_ARG;
    (void)hipSetDevice(ctx->device);
    for (auto &g : ctx->geoms) {
        if (g.d_cells) (void)hipFree(g.d_cells);
        if (g.d_cls) (void)hipFree(g.d_cls);
    }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RMS_OK;
}

extern "C" const char *rms_last_error(const rms_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

// main.cu:429-458 (+ the kappa extension); identical double arithmetic to the
// host restatement used for parity.
static bool derive(const rms_geometry &g, double dt, double step[NC], double inv_step[NC], double T2s[NC],
                   double P[NC * NC], int Pcat[NC * NC], std::string &err)
{
    if (g.nclass < 1 || g.nclass > NC) {
        err = "nclass must be in 1..4";
        return false;
    }
    for (int i = 0; i < NC; i++) {
        step[i] = inv_step[i] = 0.0;
        T2s[i] = INFINITY;
    }
    for (int i = 0; i < NC * NC; i++) {
        P[i] = 0.0;
        Pcat[i] = 0;
    }
    for (int i = 0; i < g.nclass; i++) {
        T2s[i] = g.T2[i] / dt;
        step[i] = std::sqrt(6.0 * dt * g.D[i]) / g.res;
        inv_step[i] = (step[i] > 0.0) ? 1.0 / step[i] : 0.0;
    }
    for (int i = 0; i < g.nclass; i++) {
        for (int j = 0; j < g.nclass; j++) {
            double p;
            if ((i == 0) || (j == 0)) {
                p = 0.0;
            } else if (i == j) {
                p = 1.0;
            } else {
                p = std::fmin(1.0, std::sqrt(g.D[j] / g.D[i]));
                double kap = g.kappa[i * NC + j];
                if (kap >= 0.0) {
                    double smin = std::fmin(step[i], step[j]);
                    double pi0 = 4.0 * kap * dt / (g.res * smin);
                    p = p * (pi0 / (1.0 + pi0));
                }
            }
            P[i * NC + j] = p;
            Pcat[i * NC + j] = (p > 0.999999) ? 1 : ((p < 0.000001) ? 0 : 2);
        }
    }
    return true;
}

static rms_status derive_impl(const rms_geometry *geom, double dt, double *step, double *T2s, double *P)
{
    if (!geom || !(dt > 0.0)) return RMS_ERR_INVALID_ARG;
    double st[NC], ist[NC], t2[NC], pp[NC * NC];
    int pc[NC * NC];
    std::string err;
    if (!derive(*geom, dt, st, ist, t2, pp, pc, err)) return RMS_ERR_INVALID_ARG;
    if (step) std::memcpy(step, st, sizeof(st));
    if (T2s) std::memcpy(T2s, t2, sizeof(t2));
    if (P) std::memcpy(P, pp, sizeof(pp));
    return RMS_OK;
}

extern "C" rms_status rms_derive(const rms_geometry *geom, double dt, double *step, double *T2s, double *P)
{
    return guard(nullptr, [&] { return derive_impl(geom, dt, step, T2s, P); });
}

// Constants of the tangent (adjoint) pass, from the same quantities as
// derive(): parameter p < 3 is D of class p+1, p = 3 is kappa.  For the
// probabilistic pairs P = min(1, sqrt(Dj/Di)) * pi0/(1+pi0) with
// pi0 = 4 kappa dt / (res min(s_i, s_j)); d log P and d log(1-P) = -P/(1-P) d log P.
static void derive_tangent(const rms_geometry &g, double dt, const double step[NC], const double inv_step[NC],
                           const double T2s[NC], const double P[NC * NC], const int Pcat[NC * NC], JobDesc &D)
{
    for (int c = 0; c < NC; c++) {
        D.dstep_d[c] = D.dinv_d[c] = D.rT[c] = D.invT2[c] = 0.0;
        D.dstep_f[c] = D.dinv_f[c] = 0.0f;
    }
    for (int i = 0; i < NC * NC * 4; i++) D.dlogP[i] = D.dlog1mP[i] = 0.0;
    for (int c = 0; c < g.nclass; c++) {
        const double ds = (g.D[c] > 0.0) ? step[c] / (2.0 * g.D[c]) : 0.0;
        const double di = (g.D[c] > 0.0) ? -inv_step[c] / (2.0 * g.D[c]) : 0.0;
        D.dstep_d[c] = ds;
        D.dinv_d[c] = di;
        D.dstep_f[c] = (float)ds;
        D.dinv_f[c] = (float)di;
        D.rT[c] = 1.0 / T2s[c];
        D.invT2[c] = 1.0 / g.T2[c];
    }
    for (int i = 1; i < g.nclass; i++) {
        for (int j = 1; j < g.nclass; j++) {
            const int ij = i * NC + j;
            if (Pcat[ij] != 2) continue;
            double dl[3] = {0.0, 0.0, 0.0};
            double dk = 0.0;
            const double ratio = std::sqrt(g.D[j] / g.D[i]);
            if (ratio < 1.0) {
                dl[j - 1] = dl[j - 1] + 0.5 / g.D[j];
                dl[i - 1] = dl[i - 1] - 0.5 / g.D[i];
            }
            const double kap = g.kappa[ij];
            if (kap >= 0.0) {
                const double smin = std::fmin(step[i], step[j]);
                const double pi0 = 4.0 * kap * dt / (g.res * smin);
                const int m = (step[i] <= step[j]) ? i : j;
                dl[m - 1] = dl[m - 1] - 0.5 / (g.D[m] * (1.0 + pi0));
                dk = (kap > 0.0) ? 1.0 / (kap * (1.0 + pi0)) : 0.0;
            }
            const double r = -P[ij] / (1.0 - P[ij]);
            for (int q = 0; q < 3; q++) {
                D.dlogP[ij * 4 + q] = dl[q];
                D.dlog1mP[ij * 4 + q] = r * dl[q];
            }
            D.dlogP[ij * 4 + 3] = dk;
            D.dlog1mP[ij * 4 + 3] = r * dk;
        }
    }
}

static rms_status add_geometry_impl(rms_ctx *ctx, const rms_geometry *geom, int32_t *geom_id)
{
    if (!ctx || !geom || !geom_id) return fail(ctx, RMS_ERR_INVALID_ARG, "null argument");
    const rms_geometry &g = *geom;
    if (g.nx < 1 || g.ny < 1 || g.nz < 1) return fail(ctx, RMS_ERR_INVALID_ARG, "lattice dimensions must be >= 1");
    int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    if (nvox > 2147483647LL)
        return fail(ctx, RMS_ERR_INVALID_ARG, "lattice has >= 2^31 voxels (the reference int index overflows too)");
    if (g.label_bytes != 1 && g.label_bytes != 2) return fail(ctx, RMS_ERR_INVALID_ARG, "label_bytes must be 1 or 2");
    if (!g.labels) return fail(ctx, RMS_ERR_INVALID_ARG, "labels is null");
    if (g.nclass < 1 || g.nclass > NC) return fail(ctx, RMS_ERR_INVALID_ARG, "nclass must be in 1..4");
    if (!(g.res > 0.0)) return fail(ctx, RMS_ERR_INVALID_ARG, "voxel size must be positive");
    if (g.periodic & ~7) return fail(ctx, RMS_ERR_INVALID_ARG, "periodic: bits 0..2 (x, y, z) only");
    if (g.label_class) {
        if (g.nlabel < 1) return fail(ctx, RMS_ERR_INVALID_ARG, "nlabel must be >= 1 with a class map");
        for (int i = 0; i < g.nlabel; i++)
            if (g.label_class[i] >= g.nclass) return fail(ctx, RMS_ERR_INVALID_ARG, "class map value >= nclass");
    }
    // scan labels: range and presence of a walkable voxel
    uint32_t maxlab = 0;
    bool walk = false;
    for (int64_t i = 0; i < nvox; i++) {
        uint32_t lab = g.label_bytes == 1 ? ((const uint8_t *)g.labels)[i] : ((const uint16_t *)g.labels)[i];
        if (lab > maxlab) maxlab = lab;
        uint32_t c = g.label_class ? (lab < (uint32_t)g.nlabel ? g.label_class[lab] : 0u) : lab;
        if (c != 0) walk = true;
    }
    if (g.label_class ? maxlab >= (uint32_t)g.nlabel : maxlab >= (uint32_t)g.nclass)
        return fail(ctx, RMS_ERR_INVALID_ARG, "label value out of range of the class table");
    if (!walk)
        return fail(ctx, RMS_ERR_INVALID_ARG,
                    "lattice has no non-null voxel (the reference initialisation loop would never end)");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    GeomDev gd;
    gd.meta = g;
    gd.meta.labels = nullptr;
    gd.meta.label_class = nullptr;
    gd.nvox = nvox;
    gd.has_walkable = walk;
    // the labels and class map travel once; the kernels read only what is
    // built from them: a class per voxel (uint8) and a cell word (uint64)
    const size_t bytes = (size_t)nvox * g.label_bytes;
    void *d_lab = nullptr;
    uint8_t *d_map = nullptr;
    hipError_t e = hipMalloc(&gd.d_cells, sizeof(uint64_t) * (size_t)nvox);
    if (e == hipSuccess) e = hipMalloc(&gd.d_cls, (size_t)nvox);
    if (e == hipSuccess) e = hipMalloc(&d_lab, bytes);
    if (e == hipSuccess) e = hipMemcpy(d_lab, g.labels, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && g.label_class) {
        e = hipMalloc((void **)&d_map, (size_t)g.nlabel);
        if (e == hipSuccess) e = hipMemcpy(d_map, g.label_class, (size_t)g.nlabel, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = launch_build_cells(d_lab, g.label_bytes, d_map, g.nx, g.ny, g.nz, gd.d_cells, gd.d_cls, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (d_lab) (void)hipFree(d_lab);
    if (d_map) (void)hipFree(d_map);
    if (e != hipSuccess) {
        if (gd.d_cells) (void)hipFree(gd.d_cells);
        if (gd.d_cls) (void)hipFree(gd.d_cls);
        ctx->err = std::string("geometry upload: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? RMS_ERR_OUT_OF_MEMORY : RMS_ERR_HIP;
    }
    try {
        ctx->geoms.push_back(gd);
    } catch (...) {
        (void)hipFree(gd.d_cells);
        (void)hipFree(gd.d_cls);
        throw;
    }
    *geom_id = (int32_t)ctx->geoms.size() - 1;
    return RMS_OK;
}

extern "C" rms_status rms_add_geometry(rms_ctx *ctx, const rms_geometry *geom, int32_t *geom_id)
{
    return guard(ctx, [&] { return add_geometry_impl(ctx, geom, geom_id); });
}

// ---------------------------------------------------------------- plans
template <typename T>
static rms_status palloc(rms_plan *p, T **ptr, size_t bytes)
{
    *ptr = nullptr;
    if (bytes == 0) return RMS_OK;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        p->ctx->err = std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? RMS_ERR_OUT_OF_MEMORY : RMS_ERR_HIP;
    }
    p->allocs.push_back(q);
    *ptr = (T *)q;
    return RMS_OK;
}

template <typename T>
static rms_status pupload(rms_plan *p, T **ptr, const T *src, size_t n)
{
    rms_status s = palloc(p, ptr, n * sizeof(T));
    if (s != RMS_OK || n == 0) return s;
    hipError_t e = hipMemcpy(*ptr, src, n * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        p->ctx->err = std::string("hipMemcpy: ") + hipGetErrorString(e);
        return RMS_ERR_HIP;
    }
    return RMS_OK;
}

// The PGSE sums keep sum(g) of the steps since the last flush in int32; with
// voxel indices below max(nx, ny, nz) a flush every 2^k steps, k the largest
// value <= 16 with 2^k max(nx, ny, nz) < 2^31, keeps it in range.
static int32_t pgse_flush_log2(int32_t nx, int32_t ny, int32_t nz)
{
    const int64_t m = std::max<int64_t>(1, std::max<int64_t>(nx, std::max<int64_t>(ny, nz)));
    int32_t k = 16;
    while (k > 0 && (m << k) > 2147483647LL) k--;
    return k;
}

// Threshold on the high half xhi of an FP32 step's first exchange draw: the
// draw's uniform is at least u01_f32(xhi << 16) = (256 xhi + 1) 2^-24, so
// xhi >= the smallest xhi with that bound >= P reflects whatever the low half.
// -1: the face reflects without a draw (P = 0); 65536: never decided by xhi
// alone (P = 1 crosses, and with more than two non-null classes the relation
// word does not name the other class).
static int32_t exchange_threshold(int32_t nclass, int cat, float P)
{
    if (nclass != 3 || cat == 1) return 65536;
    if (cat == 0) return -1;
    int32_t lo = 0, hi = 65536;  // smallest xhi in [0, 65536) with the bound >= P, else 65536
    while (lo < hi) {
        const int32_t mid = (lo + hi) / 2;
        const float umin = (2.0f * (float)(mid * 128) + 1.0f) * 5.9604644775390625e-08f;
        if (umin >= P)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

static int64_t first_false_lt(double dt, int64_t TN, double thr)
{
    int64_t lo = 0, hi = TN;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if ((double)(mid + 1) * dt < thr)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

static int64_t first_true_ge(double dt, int64_t TN, double thr)
{
    int64_t lo = 0, hi = TN;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if ((double)(mid + 1) * dt >= thr)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

static void plan_free(rms_plan *p)
{
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (auto &e : p->ev)
        if (e) (void)hipEventDestroy(e);
    for (void *q : p->allocs) (void)hipFree(q);
    delete p;
}

extern "C" rms_status rms_plan_destroy(rms_plan *plan)
{
    if (!plan) return RMS_ERR_INVALID_ARG;
    plan_free(plan);
    return RMS_OK;
}

namespace {
// owns a plan under construction; released into *out on success
struct PlanHolder {
    rms_plan *p = nullptr;
    ~PlanHolder() { plan_free(p); }
};
}  // namespace

static rms_status plan_create_impl(rms_ctx *ctx, const rms_job *jobs, int32_t njobs, const rms_run_params *prm,
                                   rms_plan **out)
{
    if (!ctx || !jobs || njobs < 1 || !prm || !out) return fail(ctx, RMS_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    const rms_run_params &P = *prm;
    if (!(P.dt > 0.0)) return fail(ctx, RMS_ERR_INVALID_ARG, "dt must be positive");
    if (P.timepoints < 1) return fail(ctx, RMS_ERR_INVALID_ARG, "timepoints must be >= 1");
    if (P.TN < P.timepoints)
        return fail(ctx, RMS_ERR_INVALID_ARG,
                    "TN must be >= timepoints (the reference divides by TN/timepoints, main.cu:68,333)");
    if (P.TN > 2147483647LL) return fail(ctx, RMS_ERR_INVALID_ARG, "TN must be < 2^31");
    if (P.precision != RMS_PREC_F32 && P.precision != RMS_PREC_F64)
        return fail(ctx, RMS_ERR_INVALID_ARG, "precision must be RMS_PREC_F32 or RMS_PREC_F64");
    const int R = P.replicas < 1 ? 1 : P.replicas;
    const int64_t T = P.timepoints;
    const int64_t Tstep = P.TN / P.timepoints;

    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, RMS_ERR_HIP, "hipSetDevice failed");
    PlanHolder holder;
    rms_plan *p = holder.p = new rms_plan();
    p->ctx = ctx;
    p->prm = P;
    p->prm.replicas = R;
    rms_status st = RMS_OK;
#define TRY(x)                       \
    do {                             \
        st = (x);                    \
        if (st != RMS_OK) return st; \
    } while (0)
#define BAD(msg) TRY(fail(ctx, RMS_ERR_INVALID_ARG, msg))

    int sig = 0;
    int64_t nbatch = 0;
    int maxbound = 0;
    p->hjobs.resize(njobs);
    p->jobs.resize(njobs);
    std::vector<std::vector<double>> qscale(njobs);
    for (int j = 0; j < njobs; j++) {
        const rms_job &J = jobs[j];
        if (J.geom_id < 0 || J.geom_id >= (int32_t)ctx->geoms.size()) BAD("job " + std::to_string(j) + ": bad geom_id");
        if (J.nwalk < 1) BAD("job " + std::to_string(j) + ": nwalk must be >= 1");
        if (J.nbvec < 0 || (J.nbvec > 0 && (!J.bval || !J.bvec))) BAD("job: bad b-table");
        if (J.ngrad < 0 || (J.ngrad > 0 && !J.grad)) BAD("job: bad gradient table");
        if (J.nbvec > 0) sig |= 1;
        if (J.ngrad > 0) sig |= 2;
    }
    p->sig = sig;
    for (int j = 0; j < njobs; j++) {
        const rms_job &J = jobs[j];
        const GeomDev &G = ctx->geoms[J.geom_id];
        JobDesc &D = p->hjobs[j];
        std::memset(&D, 0, sizeof(D));
        double step[NC], inv_step[NC], T2s[NC], Pp[NC * NC];
        int Pcat[NC * NC];
        std::string err;
        if (!derive(G.meta, P.dt, step, inv_step, T2s, Pp, Pcat, err)) BAD(err);
        for (int c = 1; c < G.meta.nclass; c++) {
            if (!(G.meta.D[c] > 0.0)) BAD("non-null classes need a positive diffusivity");
            if (!(step[c] < 1.0))
                BAD("step size sqrt(6*D*dt) must be below one voxel (main.cu:210-214 assumes it)");
            if (!(G.meta.T2[c] > 0.0)) BAD("T2 must be positive");
        }
        D.cells = G.d_cells;
        D.cls = G.d_cls;
        D.nx = G.meta.nx;
        D.ny = G.meta.ny;
        D.nz = G.meta.nz;
        D.pg_log2 = pgse_flush_log2(G.meta.nx, G.meta.ny, G.meta.nz);
        D.periodic = G.meta.periodic;
        D.nclass = G.meta.nclass;
        D.nbvec = J.nbvec;
        D.ngrad = J.ngrad;
        D.nwalk = J.nwalk;
        D.key0 = (uint32_t)J.seed;
        D.key1 = (uint32_t)(J.seed >> 32);
        D.res = G.meta.res;
        uint32_t pcat = 0;
        for (int i = 0; i < NC * NC; i++) pcat |= (uint32_t)(Pcat[i] & 3) << (2 * i);
        D.pcat = pcat;
        for (int c = 0; c < NC; c++) {
            D.step_f[c] = (float)step[c];
            D.inv_step_f[c] = (float)inv_step[c];
            D.step_d[c] = step[c];
            D.inv_step_d[c] = inv_step[c];
            D.T2s[c] = T2s[c];
        }
        for (int i = 0; i < NC * NC; i++) {
            D.P_f[i] = (float)Pp[i];
            D.P_d[i] = Pp[i];
        }
        D.xh12 = exchange_threshold(G.meta.nclass, Pcat[1 * NC + 2], D.P_f[1 * NC + 2]);
        D.xh21 = exchange_threshold(G.meta.nclass, Pcat[2 * NC + 1], D.P_f[2 * NC + 1]);
        derive_tangent(G.meta, P.dt, step, inv_step, T2s, Pp, Pcat, D);
        JobHost &H = p->jobs[j];
        H.exchange = false;
        for (int a = 1; a < G.meta.nclass; a++)
            for (int b = 1; b < G.meta.nclass; b++)
                if (a != b && Pcat[a * NC + b] != 0) H.exchange = true;
        // the walk kernel with faces between the two non-null classes on the
        // fast path (reflect at P = 0, draw at 0 < P < 1); it is also the only
        // variant whose march compiles exchange draws (category 2)
        bool draws = false;
        for (int i = 0; i < NC * NC; i++) draws = draws || Pcat[i] == 2;
        if (G.meta.nclass == 3 || draws) p->sig |= 4;
        // the walk kernel whose march wraps periodic lattice edges
        if (G.meta.periodic != 0) p->sig |= 8;
        H.geom_id = J.geom_id;
        H.nwalk = J.nwalk;
        H.nbatch = (J.nwalk + WAVE - 1) / WAVE;
        H.batch0 = nbatch;
        H.wbase = nbatch * WAVE;
        H.nbvec = J.nbvec;
        H.ngrad = J.ngrad;
        H.nval = 1 + G.meta.nclass + 21 + ((sig & 1) ? J.nbvec : 0);
        H.nbound = 0;
        H.ntriple = 0;
        D.batch0 = H.batch0;
        D.wbase = H.wbase;
        nbatch += H.nbatch;
        // narrow-pulse q scale table, sqrt(bval/TD) (main.cu:375)
        if (J.nbvec > 0) {
            for (int b = 0; b < J.nbvec; b++)
                if (!(J.bval[b] >= 0.0)) BAD("b-values must be >= 0");
            qscale[j].resize((size_t)T * J.nbvec);
            for (int64_t r = 0; r < T; r++) {
                double TD = (double)(r * Tstep + 1) * P.dt;  // main.cu:486
                for (int b = 0; b < J.nbvec; b++) qscale[j][(size_t)r * J.nbvec + b] = std::sqrt(J.bval[b] / TD);
            }
        }
    }
    p->nbatch = nbatch;
    p->ws = nbatch * WAVE;

    // PGSE tables (window sums), per job
    for (int j = 0; j < njobs; j++) {
        const rms_job &J = jobs[j];
        JobDesc &D = p->hjobs[j];
        JobHost &H = p->jobs[j];
        if (J.ngrad <= 0) continue;
        std::vector<std::array<int64_t, 3>> abe(J.ngrad);
        std::vector<int64_t> bl;
        bl.push_back(P.TN);
        for (int g = 0; g < J.ngrad; g++) {
            const double *gr = J.grad + 6 * g;
            int64_t a = first_false_lt(P.dt, P.TN, gr[1]);
            int64_t b = first_true_ge(P.dt, P.TN, gr[0]);
            int64_t e = first_false_lt(P.dt, P.TN, gr[0] + gr[1]);
            if (b < a) b = a;
            if (e < b) e = b;
            abe[g] = {a, b, e};
            bl.push_back(a);
            bl.push_back(b);
            bl.push_back(e);
        }
        std::sort(bl.begin(), bl.end());
        bl.erase(std::unique(bl.begin(), bl.end()), bl.end());
        std::vector<int64_t> bounds(1, 0);
        for (int64_t v : bl)
            if (v != 0) bounds.push_back(v);
        if ((int)bounds.size() - 1 > MAX_BOUNDS) BAD("too many distinct PGSE lobe boundaries (max 64)");
        H.nbound = (int)bounds.size() - 1;
        maxbound = std::max(maxbound, H.nbound);
        auto kidx = [&](int64_t v) {
            return (int32_t)(std::lower_bound(bounds.begin(), bounds.end(), v) - bounds.begin());
        };
        std::vector<std::tuple<int32_t, int32_t, int32_t, int32_t>> order;
        for (int g = 0; g < J.ngrad; g++) order.emplace_back(kidx(abe[g][0]), kidx(abe[g][1]), kidx(abe[g][2]), g);
        std::stable_sort(order.begin(), order.end());
        std::vector<double> gs((size_t)J.ngrad * 6);
        std::vector<int32_t> perm(J.ngrad), trip;
        for (int g = 0; g < J.ngrad; g++) {
            int src = std::get<3>(order[g]);
            perm[g] = src;
            std::memcpy(&gs[(size_t)g * 6], J.grad + 6 * src, 6 * sizeof(double));
            bool newt = g == 0 || std::get<0>(order[g]) != std::get<0>(order[g - 1]) ||
                        std::get<1>(order[g]) != std::get<1>(order[g - 1]) ||
                        std::get<2>(order[g]) != std::get<2>(order[g - 1]);
            if (newt) {
                trip.push_back(std::get<0>(order[g]));
                trip.push_back(std::get<1>(order[g]));
                trip.push_back(std::get<2>(order[g]));
                trip.push_back(g);
            }
        }
        H.ntriple = (int)trip.size() / 4;
        trip.push_back(0);
        trip.push_back(0);
        trip.push_back(0);
        trip.push_back(J.ngrad);  // sentinel
        H.gperm = perm;
        D.ntriple = H.ntriple;
        D.nbound = H.nbound;
        double *dg;
        int32_t *dp, *dt3;
        int64_t *db;
        TRY(pupload(p, &dg, gs.data(), gs.size()));
        TRY(pupload(p, &dp, perm.data(), perm.size()));
        TRY(pupload(p, &dt3, trip.data(), trip.size()));
        TRY(pupload(p, &db, bounds.data(), bounds.size()));
        D.grad = dg;
        D.gperm = dp;
        D.triple = dt3;
        D.bounds = db;
    }
    for (int j = 0; j < njobs; j++) {
        JobDesc &D = p->hjobs[j];
        if (jobs[j].nbvec > 0) {
            double *dq, *dbv;
            TRY(pupload(p, &dq, qscale[j].data(), qscale[j].size()));
            TRY(pupload(p, &dbv, jobs[j].bvec, (size_t)jobs[j].nbvec * 3));
            D.qscale = dq;
            D.bvec = dbv;
        }
    }
    // outputs: per-batch partials, replicas, totals
    TRY(palloc(p, &p->d_counters, sizeof(unsigned long long) * 5 * njobs));
    for (int j = 0; j < njobs; j++) {
        JobDesc &D = p->hjobs[j];
        JobHost &H = p->jobs[j];
        D.counters = p->d_counters + 5 * j;
        double *part, *rep, *tot;
        int64_t nq = T * H.nval;
        TRY(palloc(p, &part, sizeof(double) * H.nbatch * nq));
        TRY(palloc(p, &rep, sizeof(double) * R * nq));
        TRY(palloc(p, &tot, sizeof(double) * nq));
        D.part = part;
        p->red.push_back(ReduceDesc{part, rep, tot, H.nbatch, nq});
        p->max_nq = std::max(p->max_nq, nq);
        if (H.ngrad > 0) {
            double *ps, *sr, *stt;
            TRY(palloc(p, &ps, sizeof(double) * H.nbatch * H.ngrad));
            TRY(palloc(p, &sr, sizeof(double) * R * H.ngrad));
            TRY(palloc(p, &stt, sizeof(double) * H.ngrad));
            D.psig = ps;
            p->red.push_back(ReduceDesc{ps, sr, stt, H.nbatch, H.ngrad});
            p->max_nq = std::max(p->max_nq, (int64_t)H.ngrad);
        }
    }
    p->maxbound = maxbound;
    if (sig & 2) {
        TRY(palloc(p, &p->d_snap, sizeof(double) * (size_t)(maxbound + 1) * 3 * p->ws));
        for (int j = 0; j < njobs; j++) p->hjobs[j].snap = p->d_snap;
    }
    // chunked launches keep walker state in HBM between launches
    const bool chunked = P.chunk_steps > 0 && P.chunk_steps < P.TN;
    p->st.ws = p->ws;
    if (chunked) {
        const int64_t ws = p->ws;
        TRY(palloc(p, &p->st.g, sizeof(int32_t) * 3 * ws));
        TRY(palloc(p, &p->st.gi, sizeof(int32_t) * 3 * ws));
        TRY(palloc(p, &p->st.c, sizeof(int32_t) * ws));
        if (P.precision == RMS_PREC_F32) {
            TRY(palloc(p, &p->st.of, sizeof(float) * 6 * ws));
            TRY(palloc(p, &p->st.tcf, sizeof(float) * ws));
        } else {
            TRY(palloc(p, &p->st.od, sizeof(double) * 6 * ws));
            TRY(palloc(p, &p->st.tcd, sizeof(double) * ws));
        }
        TRY(palloc(p, &p->st.tsum, sizeof(double) * NC * ws));
        TRY(palloc(p, &p->st.cnt, sizeof(uint32_t) * 5 * ws));
        if (sig & 2) {
            TRY(palloc(p, &p->st.ps, sizeof(double) * 3 * ws));
            TRY(palloc(p, &p->st.sg, sizeof(int32_t) * 3 * ws));
        }
    }
    // batch -> job table
    std::vector<int32_t> bj(nbatch);
    for (int j = 0; j < njobs; j++)
        for (int64_t b = 0; b < p->jobs[j].nbatch; b++) bj[p->jobs[j].batch0 + b] = j;
    TRY(pupload(p, &p->d_batch_job, bj.data(), bj.size()));
    TRY(pupload(p, &p->d_jobs, p->hjobs.data(), p->hjobs.size()));
    TRY(pupload(p, &p->d_red, p->red.data(), p->red.size()));

    // launch shape: persistent grid at the resident limit
    p->block = P.block_threads > 0 ? P.block_threads : 256;
    if (p->block % WAVE != 0 || p->block > 256) BAD("block_threads must be a multiple of 64 and <= 256");
    int bpc = 0;
    if (walk_occupancy(P.precision == RMS_PREC_F64, p->sig, p->block, &bpc) != hipSuccess || bpc < 1) bpc = 1;
    const int wpb = p->block / WAVE;
    int64_t need = (nbatch + wpb - 1) / wpb;
    int64_t resident = (int64_t)bpc * ctx->ncu;
    int64_t grid = P.grid_blocks > 0 ? P.grid_blocks : std::min(need, resident);
    if (grid < 1) grid = 1;
    int xs = P.xcd_split;
    bool split = xs > 0 || (xs == 0 && njobs > 1);
    if (split && grid >= 8) {
        grid = (grid / 8) * 8;
        p->xcd_split = 1;
    } else {
        p->xcd_split = 0;
    }
    p->grid = (int)grid;
    for (auto &e : p->ev) {
        if (hipEventCreate(&e) != hipSuccess) TRY(fail(ctx, RMS_ERR_HIP, "hipEventCreate failed"));
    }
    *out = p;
    holder.p = nullptr;
    return RMS_OK;
#undef TRY
#undef BAD
}

extern "C" rms_status rms_plan_create(rms_ctx *ctx, const rms_job *jobs, int32_t njobs, const rms_run_params *prm,
                                      rms_plan **out)
{
    return guard(ctx, [&] { return plan_create_impl(ctx, jobs, njobs, prm, out); });
}

static LaunchArgs make_args(rms_plan *p, int64_t s0, int64_t s1)
{
    LaunchArgs a{};
    a.jobs = p->d_jobs;
    a.batch_job = p->d_batch_job;
    a.nbatch = p->nbatch;
    a.TN = p->prm.TN;
    a.step0 = s0;
    a.step1 = s1;
    a.Tstep = p->prm.TN / p->prm.timepoints;
    a.timepoints = p->prm.timepoints;
    a.dt = p->prm.dt;
    a.st = p->st;
    a.fin = p->fin;
    a.xcd_split = p->xcd_split;
    return a;
}

// Enqueue the whole run: zero counters, walk chunks, ordered reductions.
static rms_status enqueue(rms_plan *p, bool with_events)
{
    rms_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, launch_fill_zero(p->d_counters, sizeof(unsigned long long) * 5 * (int64_t)p->jobs.size(), s));
    if (with_events) HIPCHK(ctx, hipEventRecord(p->ev[0], s));
    const int64_t TN = p->prm.TN;
    const int64_t ch = (p->prm.chunk_steps > 0 && p->prm.chunk_steps < TN) ? p->prm.chunk_steps : TN;
    for (int64_t s0 = 0; s0 < TN; s0 += ch) {
        int64_t s1 = std::min(TN, s0 + ch);
        LaunchArgs a = make_args(p, s0, s1);
        HIPCHK(ctx, launch_walk(a, p->prm.precision == RMS_PREC_F64, p->sig, p->grid, p->block, s));
    }
    if (with_events) HIPCHK(ctx, hipEventRecord(p->ev[1], s));
    HIPCHK(ctx, launch_reduce(p->d_red, (int32_t)p->red.size(), p->max_nq, p->prm.replicas, s));
    if (with_events) HIPCHK(ctx, hipEventRecord(p->ev[2], s));
    return RMS_OK;
}

static rms_status plan_execute_impl(rms_plan *p, rms_job_out *outs, rms_timing *timing)
{
    rms_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int nj = (int)p->jobs.size();
    // final-state outputs requested?
    bool want_fin = false;
    if (outs)
        for (int j = 0; j < nj; j++) want_fin = want_fin || outs[j].fin_g;
    if (want_fin && !p->fin_alloc) {
        const int64_t ws = p->ws;
        rms_status st;
        if ((st = palloc(p, &p->fin.g, sizeof(int32_t) * 3 * ws)) != RMS_OK) return st;
        if ((st = palloc(p, &p->fin.o, sizeof(double) * 3 * ws)) != RMS_OK) return st;
        if ((st = palloc(p, &p->fin.gi, sizeof(int32_t) * 3 * ws)) != RMS_OK) return st;
        if ((st = palloc(p, &p->fin.oi, sizeof(double) * 3 * ws)) != RMS_OK) return st;
        if ((st = palloc(p, &p->fin.c, sizeof(int32_t) * ws)) != RMS_OK) return st;
        if ((st = palloc(p, &p->fin.cnt, sizeof(uint32_t) * 5 * ws)) != RMS_OK) return st;
        p->fin_alloc = true;
        if (p->gexec) {  // the captured graph does not write the final state
            (void)hipGraphExecDestroy(p->gexec);
            (void)hipGraphDestroy(p->graph);
            p->gexec = nullptr;
            p->graph = nullptr;
        }
    }
    hipStream_t s = ctx->stream;
    if (p->prm.use_graph) {
        if (!p->gexec) {
            HIPCHK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            rms_status st = enqueue(p, false);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s, &g);
            if (st != RMS_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            HIPCHK(ctx, e);
            p->graph = g;
            HIPCHK(ctx, hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
        }
        HIPCHK(ctx, hipEventRecord(p->ev[0], s));
        HIPCHK(ctx, hipGraphLaunch(p->gexec, s));
        HIPCHK(ctx, hipEventRecord(p->ev[2], s));
    } else {
        rms_status st = enqueue(p, true);
        if (st != RMS_OK) return st;
    }
    HIPCHK(ctx, hipStreamSynchronize(s));
    if (timing) {
        std::memset(timing, 0, sizeof(*timing));
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev[0], p->ev[2]));
        timing->total_ms = ms;
        if (!p->prm.use_graph) {
            HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
            timing->walk_ms = ms;
            HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev[1], p->ev[2]));
            timing->reduce_ms = ms;
        } else {
            timing->walk_ms = timing->total_ms;
            timing->reduce_ms = 0.0;
        }
        timing->grid = p->grid;
        timing->block = p->block;
        const int64_t TN = p->prm.TN;
        const int64_t ch = (p->prm.chunk_steps > 0 && p->prm.chunk_steps < TN) ? p->prm.chunk_steps : TN;
        timing->launches = (int32_t)((TN + ch - 1) / ch);
    }
    if (!outs) return RMS_OK;
    const int R = p->prm.replicas;
    std::vector<unsigned long long> cnt(5 * nj);
    HIPCHK(ctx, hipMemcpy(cnt.data(), p->d_counters, sizeof(unsigned long long) * 5 * nj, hipMemcpyDeviceToHost));
    size_t ri = 0;
    for (int j = 0; j < nj; j++) {
        const JobHost &H = p->jobs[j];
        rms_job_out &O = outs[j];
        const ReduceDesc &rd = p->red[ri++];
        if (O.rec) HIPCHK(ctx, hipMemcpy(O.rec, rd.tot, sizeof(double) * rd.nq, hipMemcpyDeviceToHost));
        if (O.rec_rep) HIPCHK(ctx, hipMemcpy(O.rec_rep, rd.rep, sizeof(double) * rd.nq * R, hipMemcpyDeviceToHost));
        if (H.ngrad > 0) {
            const ReduceDesc &rs = p->red[ri++];
            if (O.sig) HIPCHK(ctx, hipMemcpy(O.sig, rs.tot, sizeof(double) * rs.nq, hipMemcpyDeviceToHost));
            if (O.sig_rep)
                HIPCHK(ctx, hipMemcpy(O.sig_rep, rs.rep, sizeof(double) * rs.nq * R, hipMemcpyDeviceToHost));
        }
        for (int k = 0; k < 5; k++) O.counters[k] = cnt[5 * j + k];
        if (O.fin_g && p->fin_alloc) {
            const int64_t ws = p->ws, n = H.nwalk, b = H.wbase;
            for (int k = 0; k < 3; k++) {
                HIPCHK(ctx, hipMemcpy(O.fin_g + k * n, p->fin.g + k * ws + b, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
                HIPCHK(ctx, hipMemcpy(O.fin_o + k * n, p->fin.o + k * ws + b, sizeof(double) * n, hipMemcpyDeviceToHost));
                HIPCHK(ctx, hipMemcpy(O.fin_gi + k * n, p->fin.gi + k * ws + b, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
                HIPCHK(ctx, hipMemcpy(O.fin_oi + k * n, p->fin.oi + k * ws + b, sizeof(double) * n, hipMemcpyDeviceToHost));
            }
            HIPCHK(ctx, hipMemcpy(O.fin_c, p->fin.c + b, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
            for (int k = 0; k < 5; k++)
                HIPCHK(ctx, hipMemcpy(O.fin_cnt + k * n, p->fin.cnt + k * ws + b, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        }
    }
    return RMS_OK;
}

extern "C" rms_status rms_plan_execute(rms_plan *p, rms_job_out *outs, rms_timing *timing)
{
    if (!p) return RMS_ERR_INVALID_ARG;
    return guard(p->ctx, [&] { return plan_execute_impl(p, outs, timing); });
}

extern "C" rms_status rms_run(rms_ctx *ctx, const rms_job *jobs, int32_t njobs, const rms_run_params *params,
                              rms_job_out *outs, rms_timing *timing)
{
    rms_plan *p = nullptr;
    rms_status s = rms_plan_create(ctx, jobs, njobs, params, &p);
    if (s != RMS_OK) return s;
    s = rms_plan_execute(p, outs, timing);
    rms_plan_destroy(p);
    return s;
}

// Buffers of the adjoint pass: cotangents, per-batch gradient partials with
// their ordered reductions, and the PGSE tangent prefix-sum snapshots.
static rms_status adjoint_setup(rms_plan *p)
{
    rms_ctx *ctx = p->ctx;
    const int nj = (int)p->jobs.size();
    const int R = p->prm.replicas;
    const int64_t T = p->prm.timepoints;
    int np = 1;
    for (int j = 0; j < nj; j++) np = std::max(np, (int)p->hjobs[j].nclass - 1);
    p->adj_np = std::min(np, 3);
    p->cot_rec.assign(nj, nullptr);
    p->cot_sig.assign(nj, nullptr);
    p->red_adj.clear();  // a failed earlier setup may have left entries
    rms_status st;
    for (int j = 0; j < nj; j++) {
        const JobHost &H = p->jobs[j];
        double *pg, *rep, *tot;
        if ((st = palloc(p, &p->cot_rec[j], sizeof(double) * T * H.nval)) != RMS_OK) return st;
        if (H.ngrad > 0 && (st = palloc(p, &p->cot_sig[j], sizeof(double) * H.ngrad)) != RMS_OK) return st;
        if ((st = palloc(p, &pg, sizeof(double) * H.nbatch * NPARAM)) != RMS_OK) return st;
        if ((st = palloc(p, &rep, sizeof(double) * R * NPARAM)) != RMS_OK) return st;
        if ((st = palloc(p, &tot, sizeof(double) * NPARAM)) != RMS_OK) return st;
        p->red_adj.push_back(ReduceDesc{pg, rep, tot, H.nbatch, NPARAM});
    }
    if (p->sig & 2) {
        const size_t n = (size_t)(p->maxbound + 1) * p->adj_np * 3 * p->ws;
        if ((st = palloc(p, &p->d_tsnap, sizeof(double) * n)) != RMS_OK) return st;
    }
    if ((st = palloc(p, &p->d_jobs_adj, sizeof(JobDesc) * nj)) != RMS_OK) return st;
    if ((st = pupload(p, &p->d_red_adj, p->red_adj.data(), p->red_adj.size())) != RMS_OK) return st;
    int bpc = 0;
    const bool f64 = p->prm.precision == RMS_PREC_F64;
    if (adjoint_occupancy(f64, p->sig, p->adj_np, p->block, &bpc) != hipSuccess || bpc < 1) bpc = 1;
    const int wpb = p->block / WAVE;
    const int64_t need = (p->nbatch + wpb - 1) / wpb;
    int64_t grid = p->prm.grid_blocks > 0 ? p->prm.grid_blocks : std::min(need, (int64_t)bpc * ctx->ncu);
    if (grid < 1) grid = 1;
    if (p->xcd_split) grid = std::max<int64_t>(8, (grid / 8) * 8);
    p->adj_grid = (int)grid;
    p->adj_ready = true;
    return RMS_OK;
}

static rms_status plan_adjoint_impl(rms_plan *p, const rms_job_adj *adj, rms_timing *timing)
{
    rms_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int nj = (int)p->jobs.size();
    for (int j = 0; j < nj; j++)
        if (!adj[j].grad) return fail(ctx, RMS_ERR_INVALID_ARG, "adjoint: grad output is required for every job");
    if (!p->adj_ready) {
        rms_status st = adjoint_setup(p);
        if (st != RMS_OK) return st;
    }
    hipStream_t s = ctx->stream;
    const int64_t T = p->prm.timepoints;
    std::vector<JobDesc> hj = p->hjobs;
    for (int j = 0; j < nj; j++) {
        const JobHost &H = p->jobs[j];
        JobDesc &D = hj[j];
        D.cot_rec = nullptr;
        D.cot_sig = nullptr;
        if (adj[j].cot_rec) {
            HIPCHK(ctx, hipMemcpyAsync(p->cot_rec[j], adj[j].cot_rec, sizeof(double) * T * H.nval,
                                       hipMemcpyHostToDevice, s));
            D.cot_rec = p->cot_rec[j];
        }
        if (adj[j].cot_sig && H.ngrad > 0) {
            HIPCHK(ctx, hipMemcpyAsync(p->cot_sig[j], adj[j].cot_sig, sizeof(double) * H.ngrad,
                                       hipMemcpyHostToDevice, s));
            D.cot_sig = p->cot_sig[j];
        }
        D.pgrad = const_cast<double *>(p->red_adj[j].part);
        D.tsnap = p->d_tsnap;
        HIPCHK(ctx, launch_fill_zero(D.pgrad, sizeof(double) * H.nbatch * NPARAM, s));
    }
    HIPCHK(ctx, hipMemcpyAsync(p->d_jobs_adj, hj.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipEventRecord(p->ev[0], s));
    LaunchArgs a = make_args(p, 0, p->prm.TN);
    a.jobs = p->d_jobs_adj;
    a.st.g = nullptr;  // single launch: no walker state is saved
    a.fin = FinalOut{};
    HIPCHK(ctx, launch_adjoint(a, p->prm.precision == RMS_PREC_F64, p->sig, p->adj_np, p->adj_grid, p->block, s));
    HIPCHK(ctx, hipEventRecord(p->ev[1], s));
    HIPCHK(ctx, launch_reduce(p->d_red_adj, (int32_t)p->red_adj.size(), NPARAM, p->prm.replicas, s));
    HIPCHK(ctx, hipEventRecord(p->ev[2], s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    if (timing) {
        std::memset(timing, 0, sizeof(*timing));
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev[0], p->ev[2]));
        timing->total_ms = ms;
        HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
        timing->walk_ms = ms;
        HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev[1], p->ev[2]));
        timing->reduce_ms = ms;
        timing->grid = p->adj_grid;
        timing->block = p->block;
        timing->launches = 1;
    }
    const int R = p->prm.replicas;
    for (int j = 0; j < nj; j++) {
        const ReduceDesc &rd = p->red_adj[j];
        HIPCHK(ctx, hipMemcpy(adj[j].grad, rd.tot, sizeof(double) * NPARAM, hipMemcpyDeviceToHost));
        if (adj[j].grad_rep)
            HIPCHK(ctx, hipMemcpy(adj[j].grad_rep, rd.rep, sizeof(double) * NPARAM * R, hipMemcpyDeviceToHost));
        if (p->jobs[j].exchange) {
            // a crossing between non-null classes changes the class (and the
            // step length) discontinuously in D: no pathwise D derivative
            for (int c = 1; c < NC; c++) {
                adj[j].grad[c] = NAN;
                if (adj[j].grad_rep)
                    for (int r = 0; r < R; r++) adj[j].grad_rep[(size_t)r * NPARAM + c] = NAN;
            }
        }
    }
    return RMS_OK;
}

extern "C" rms_status rms_plan_adjoint(rms_plan *p, const rms_job_adj *adj, rms_timing *timing)
{
    if (!p || !adj) return RMS_ERR_INVALID_ARG;
    return guard(p->ctx, [&] { return plan_adjoint_impl(p, adj, timing); });
}
