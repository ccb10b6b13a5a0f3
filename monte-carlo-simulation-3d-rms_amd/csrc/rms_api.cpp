// rms_api.cpp -- the C ABI of librms_amd.so (include/rms_amd.h).
//
// Host side of the engine: validation, the derived parameters of
// hpc_code/lib/main.cu:429-458 and :484-487, PGSE lobe tables
// (main_PGSE.cu:341-345), device buffers, launch sequencing (optionally as a
// captured hipGraph) and the copies back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/rms_amd.h"
#include "rms_host.h"
#include "rms_internal.h"
#include "rms_types.h"

using namespace rms;

struct GeomDev {
    rms_geometry meta;     // labels / label_class pointers cleared
    uint64_t *d_cells = nullptr;  // cell words (straight-move outcomes per crossing pattern): dense or stored bricks
    uint8_t *d_cls = nullptr;     // class per voxel
    uint32_t *d_btab = nullptr;   // brick table (brick storage), else null
    int32_t nby = 0, nbz = 0;     // brick grid (brick storage)
    int32_t storage = RMS_STORE_DENSE;
    int64_t device_bytes = 0;
    bool wrapcells = false;       // cell words name the voxels across periodic sides (wrap_cells_rule)
    int64_t nvox = 0;
    bool has_walkable = false;
    uint32_t flags = 0;           // geometry_flags: boundary terms of the D derivative
};

struct rms_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int ncu = 0;
    std::string err;
    std::vector<GeomDev> geoms;
    // this device's copy of the scratch array of rms_run_noalloc, and the
    // events that time its runs (both obtained once, here at creation)
    unsigned char *scratch = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
};

namespace {

struct JobHost {
    bool exchange;      // walkers can cross between different non-null classes
    bool kappa_closed;  // a pair of non-null classes is impermeable because of its kappa
    int32_t geom_id;
    int64_t nwalk, nbatch, batch0, wbase;
    int32_t nbvec, ngrad, nval, nbound, ntriple;
    // offsets (doubles) of the job's totals and replica sums in the results block
    int64_t rec_tot, rec_rep, sig_tot, sig_rep;
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
    unsigned long long *d_counters = nullptr;  // [njobs][NCOUNT], inside the results block
    // results block: [totals of every reduction][counters][replica sums],
    // copied back by one asynchronous copy into its pinned host mirror
    double *d_res = nullptr;
    double *h_res = nullptr;
    int64_t res_head = 0, res_all = 0;  // bytes up to the replica sums / in total
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
    int adj_np = 1;
    int adj_grid[2] = {0, 0};  // workgroups of the tangent pass without / with the boundary terms
    JobDesc *d_jobs_adj = nullptr;
    ReduceDesc *d_red_adj = nullptr;
    std::vector<ReduceDesc> red_adj;
    std::vector<double *> cot_rec, cot_sig;
    double *d_tsnap = nullptr;
    double *d_adj_res = nullptr;  // [njobs][NPARAM] totals, then [njobs][R][NPARAM]
    double *h_adj_res = nullptr;
    // central differences on walker subsets: the tangent pass's per-walker
    // marks, the walker lists, and the subset launch's batch table and
    // reductions (allocated by the first adjoint call that needs them)
    uint32_t *d_wvis = nullptr, *d_wlist = nullptr;
    int32_t *d_bj_fd = nullptr;
    ReduceDesc *d_red_fd = nullptr;
    bool fd_walks = false;  // launches now are the central differences' walks
    bool lean = false;      // every job: T2 infinite, no exchange draw (lean_walk)
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
    c->err.reserve(1024);
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
    // device address of this device's scratch array (device `dev` is current)
    void *sp = nullptr;
    bool ok = scratch_address(&sp) == hipSuccess && sp != nullptr;
    c->scratch = (unsigned char *)sp;
    for (auto &e : c->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    if (!ok) {
        for (auto &e : c->ev)
            if (e) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(c->stream);
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
        if (g.d_btab) (void)hipFree(g.d_btab);
    }
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RMS_OK;
}

extern "C" const char *rms_last_error(const rms_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static rms_status derive_impl(const rms_geometry *geom, double dt, double *step, double *T2s, double *P)
{
    if (!geom || !(dt > 0.0)) return RMS_ERR_INVALID_ARG;
    Derived d;
    if (derive(*geom, dt, d)) return RMS_ERR_INVALID_ARG;
    if (step) std::memcpy(step, d.step, sizeof(d.step));
    if (T2s) std::memcpy(T2s, d.T2s, sizeof(d.T2s));
    if (P) std::memcpy(P, d.P, sizeof(d.P));
    return RMS_OK;
}

extern "C" rms_status rms_derive(const rms_geometry *geom, double dt, double *step, double *T2s, double *P)
{
    return guard(nullptr, [&] { return derive_impl(geom, dt, step, T2s, P); });
}

// Replace a geometry's dense cell words by brick storage (see rms_types.h):
// count the bricks to store per workgroup, prefix-sum the counts on the
// host, write the table and the stored bricks, free the dense words.
static hipError_t build_bricks(rms_ctx *ctx, GeomDev &gd)
{
    const rms_geometry &g = gd.meta;
    const int32_t nbx = (g.nx + 1) / 2, nby = (g.ny + 1) / 2, nbz = (g.nz + 3) / 4;
    const int64_t nbrick = (int64_t)nbx * nby * nbz;
    const int64_t nblk = (nbrick + 255) / 256;
    hipStream_t s = ctx->stream;
    uint32_t *d_cnt = nullptr;
    uint64_t *store = nullptr;
    std::vector<uint32_t> cnt((size_t)nblk);
    hipError_t e = hipMalloc(&d_cnt, sizeof(uint32_t) * (size_t)nblk);
    if (e == hipSuccess) e = launch_brick_count(gd.d_cells, gd.d_cls, g.nx, g.ny, g.nz, nby, nbz, nbrick, d_cnt, s);
    if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), d_cnt, sizeof(uint32_t) * (size_t)nblk, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    uint64_t nstored = 0;
    for (int64_t k = 0; k < nblk && e == hipSuccess; k++) {
        const uint32_t c = cnt[(size_t)k];
        cnt[(size_t)k] = (uint32_t)nstored;
        nstored += c;
    }
    if (e == hipSuccess) e = hipMemcpyAsync(d_cnt, cnt.data(), sizeof(uint32_t) * (size_t)nblk, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMalloc(&gd.d_btab, sizeof(uint32_t) * (size_t)nbrick);
    if (e == hipSuccess) e = hipMalloc(&store, sizeof(uint64_t) * 16 * (size_t)std::max<uint64_t>(nstored, 1));
    if (e == hipSuccess)
        e = launch_brick_fill(gd.d_cells, gd.d_cls, g.nx, g.ny, g.nz, nby, nbz, nbrick, d_cnt, gd.d_btab, store, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (d_cnt) (void)hipFree(d_cnt);
    if (e != hipSuccess) {
        if (store) (void)hipFree(store);
        return e;
    }
    (void)hipFree(gd.d_cells);
    gd.d_cells = store;
    gd.nby = nby;
    gd.nbz = nbz;
    gd.device_bytes = gd.nvox + 4 * nbrick + 128 * (int64_t)std::max<uint64_t>(nstored, 1);
    return hipSuccess;
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
    if (g.storage < RMS_STORE_AUTO || g.storage > RMS_STORE_BRICKS)
        return fail(ctx, RMS_ERR_INVALID_ARG, "storage must be RMS_STORE_AUTO, RMS_STORE_DENSE or RMS_STORE_BRICKS");
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
    gd.flags = geometry_flags(g);
    gd.wrapcells = wrap_cells_rule(g);
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
        e = launch_build_cells(d_lab, g.label_bytes, d_map, g.nx, g.ny, g.nz, gd.wrapcells ? g.periodic : 0,
                               gd.d_cells, gd.d_cls, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (d_lab) (void)hipFree(d_lab);
    if (d_map) (void)hipFree(d_map);
    gd.storage = g.storage;
    if (gd.storage == RMS_STORE_AUTO) {
        // dense words (the faster walk) while the device has room, bricks
        // once keeping them would leave under 1/8 of its memory free
        size_t fr = 0, tot = 0;
        const bool known = hipMemGetInfo(&fr, &tot) == hipSuccess;
        gd.storage = (known && (double)fr < 8.0 * (double)nvox + (double)tot / 8.0) ? RMS_STORE_BRICKS
                                                                                     : RMS_STORE_DENSE;
    }
    gd.device_bytes = nvox * 9;
    if (e == hipSuccess && gd.storage == RMS_STORE_BRICKS) e = build_bricks(ctx, gd);
    if (e != hipSuccess) {
        if (gd.d_cells) (void)hipFree(gd.d_cells);
        if (gd.d_cls) (void)hipFree(gd.d_cls);
        if (gd.d_btab) (void)hipFree(gd.d_btab);
        ctx->err = std::string("geometry upload: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? RMS_ERR_OUT_OF_MEMORY : RMS_ERR_HIP;
    }
    try {
        ctx->geoms.push_back(gd);
    } catch (...) {
        (void)hipFree(gd.d_cells);
        (void)hipFree(gd.d_cls);
        if (gd.d_btab) (void)hipFree(gd.d_btab);
        throw;
    }
    *geom_id = (int32_t)ctx->geoms.size() - 1;
    return RMS_OK;
}

extern "C" rms_status rms_add_geometry(rms_ctx *ctx, const rms_geometry *geom, int32_t *geom_id)
{
    return guard(ctx, [&] { return add_geometry_impl(ctx, geom, geom_id); });
}

extern "C" rms_status rms_geometry_info(rms_ctx *ctx, int32_t geom_id, int64_t *device_bytes, int32_t *storage,
                                        int32_t *wrap_cells)
{
    if (!ctx) return RMS_ERR_INVALID_ARG;
    if (geom_id < 0 || geom_id >= (int32_t)ctx->geoms.size()) return fail(ctx, RMS_ERR_INVALID_ARG, "bad geom_id");
    if (device_bytes) *device_bytes = ctx->geoms[geom_id].device_bytes;
    if (storage) *storage = ctx->geoms[geom_id].storage;
    if (wrap_cells) *wrap_cells = ctx->geoms[geom_id].wrapcells ? 1 : 0;
    return RMS_OK;
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

static void plan_free(rms_plan *p)
{
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (auto &e : p->ev)
        if (e) (void)hipEventDestroy(e);
    for (void *q : p->allocs) (void)hipFree(q);
    if (p->h_res) (void)hipHostFree(p->h_res);
    if (p->h_adj_res) (void)hipHostFree(p->h_adj_res);
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

// Stand-ins for the tables a kernel variant reads from every job of a plan,
// for the jobs that have none of their own: an empty PGSE table (bounds
// {0}, no triple, sentinel first gradient 0) and an empty b-table.
struct Sentinels {
    int64_t bounds[2] = {0, 0};
    int32_t triple[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int32_t gperm[2] = {0, 0};
    double grad[6] = {0, 0, 0, 0, 0, 0};
    double qscale[2] = {0, 0};
    double bvec[3] = {0, 0, 0};
    double psig[2] = {0, 0};
};
}  // namespace

static rms_status plan_create_impl(rms_ctx *ctx, const rms_job *jobs, int32_t njobs, const rms_run_params *prm,
                                   rms_plan **out)
{
    if (!ctx || !jobs || njobs < 1 || !prm || !out) return fail(ctx, RMS_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    const rms_run_params &P = *prm;
    if (const char *e = check_params(P)) return fail(ctx, RMS_ERR_INVALID_ARG, e);
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
    bool lean = true;
    p->hjobs.resize(njobs);
    p->jobs.resize(njobs);
    for (int j = 0; j < njobs; j++) {
        const rms_job &J = jobs[j];
        if (J.geom_id < 0 || J.geom_id >= (int32_t)ctx->geoms.size()) BAD("job " + std::to_string(j) + ": bad geom_id");
        if (J.nbvec > 0) sig |= SIG_NARROW;
        if (J.ngrad > 0) sig |= SIG_PGSE;
    }
    for (int j = 0; j < njobs; j++) {
        const rms_job &J = jobs[j];
        const GeomDev &G = ctx->geoms[J.geom_id];
        JobDesc &D = p->hjobs[j];
        JobHost &H = p->jobs[j];
        int bits = 0;
        bool exch = false;
        if (const char *e = fill_job(G.meta, J, P, D, bits, exch)) BAD("job " + std::to_string(j) + ": " + e);
        Derived d;
        (void)derive(G.meta, P.dt, d);
        sig |= bits;
        if (G.d_btab) sig |= SIG_BRICKS;
        if (G.wrapcells) sig |= SIG_WRAP;
        D.wrapcells = G.wrapcells ? 1 : 0;
        D.cells = G.d_cells;
        D.cls = G.d_cls;
        D.btab = G.d_btab;
        D.nby = G.nby;
        D.nbz = G.nbz;
        H.exchange = exch;
        H.kappa_closed = kappa_closes_pair(G.meta, d);
        lean = lean && lean_walk(G.meta, d);
        H.geom_id = J.geom_id;
        H.nwalk = J.nwalk;
        H.nbatch = (J.nwalk + WAVE - 1) / WAVE;
        H.batch0 = nbatch;
        H.wbase = nbatch * WAVE;
        H.nbvec = J.nbvec;
        H.ngrad = J.ngrad;
        H.nval = 1 + G.meta.nclass + 21 + ((sig & SIG_NARROW) ? J.nbvec : 0);
        H.nbound = 0;
        H.ntriple = 0;
        D.batch0 = H.batch0;
        D.wbase = H.wbase;
        nbatch += H.nbatch;
    }
    p->sig = sig;
    p->lean = lean;
    p->nbatch = nbatch;
    p->ws = nbatch * WAVE;

    // sentinel tables for the jobs without their own
    Sentinels sen;
    int64_t *s_bounds;
    int32_t *s_triple, *s_gperm;
    double *s_grad, *s_qscale, *s_bvec, *s_psig;
    TRY(pupload(p, &s_bounds, sen.bounds, 2));
    TRY(pupload(p, &s_triple, sen.triple, 8));
    TRY(pupload(p, &s_gperm, sen.gperm, 2));
    TRY(pupload(p, &s_grad, sen.grad, 6));
    TRY(pupload(p, &s_qscale, sen.qscale, 2));
    TRY(pupload(p, &s_bvec, sen.bvec, 3));
    TRY(pupload(p, &s_psig, sen.psig, 2));

    // PGSE tables (window sums), per job
    for (int j = 0; j < njobs; j++) {
        const rms_job &J = jobs[j];
        JobDesc &D = p->hjobs[j];
        JobHost &H = p->jobs[j];
        D.bounds = s_bounds;
        D.triple = s_triple;
        D.gperm = s_gperm;
        D.grad = s_grad;
        if (J.ngrad <= 0) continue;
        std::vector<int32_t> perm(J.ngrad), trip(4 * ((size_t)J.ngrad + 1));
        std::vector<uint64_t> keys(J.ngrad);
        std::vector<int64_t> bounds(MAX_BOUNDS + 1);
        int32_t nb = 0, nt = 0;
        if (const char *e = pgse_tables(J.grad, J.ngrad, P.dt, P.TN, perm.data(), trip.data(), J.ngrad,
                                        bounds.data(), nb, nt, keys.data()))
            BAD("job " + std::to_string(j) + ": " + e);
        std::vector<double> gs((size_t)J.ngrad * 6);
        for (int g = 0; g < J.ngrad; g++) std::memcpy(&gs[(size_t)g * 6], J.grad + 6 * (size_t)perm[g], 6 * sizeof(double));
        H.nbound = nb;
        H.ntriple = nt;
        D.nbound = nb;
        D.ntriple = nt;
        maxbound = std::max(maxbound, nb);
        double *dg;
        int32_t *dp, *dt3;
        int64_t *db;
        TRY(pupload(p, &dg, gs.data(), gs.size()));
        TRY(pupload(p, &dp, perm.data(), perm.size()));
        TRY(pupload(p, &dt3, trip.data(), 4 * ((size_t)nt + 1)));
        TRY(pupload(p, &db, bounds.data(), (size_t)nb + 1));
        D.grad = dg;
        D.gperm = dp;
        D.triple = dt3;
        D.bounds = db;
    }
    // narrow-pulse q scale table, sqrt(bval/TD) (main.cu:375), and b-vectors
    for (int j = 0; j < njobs; j++) {
        const rms_job &J = jobs[j];
        JobDesc &D = p->hjobs[j];
        D.qscale = s_qscale;
        D.bvec = s_bvec;
        if (J.nbvec > 0) {
            std::vector<double> qs((size_t)T * J.nbvec);
            for (int64_t r = 0; r < T; r++) {
                const double TD = (double)(r * Tstep + 1) * P.dt;  // main.cu:486
                for (int b = 0; b < J.nbvec; b++) qs[(size_t)r * J.nbvec + b] = std::sqrt(J.bval[b] / TD);
            }
            double *dq, *dbv;
            TRY(pupload(p, &dq, qs.data(), qs.size()));
            TRY(pupload(p, &dbv, J.bvec, (size_t)J.nbvec * 3));
            D.qscale = dq;
            D.bvec = dbv;
        }
    }
    // results block: totals of every reduction, the counters, then the
    // replica sums (one copy brings the first part back, or all of it)
    int64_t nres = 0;
    for (int j = 0; j < njobs; j++) {
        JobHost &H = p->jobs[j];
        H.rec_tot = nres;
        nres += T * H.nval;
        H.sig_tot = nres;
        nres += H.ngrad;
    }
    const int64_t cnt_off = nres;  // in doubles; NCOUNT 64-bit counters per job
    nres += (int64_t)NCOUNT * njobs;
    p->res_head = nres * 8;
    for (int j = 0; j < njobs; j++) {
        JobHost &H = p->jobs[j];
        H.rec_rep = nres;
        nres += (int64_t)R * T * H.nval;
        H.sig_rep = nres;
        nres += (int64_t)R * H.ngrad;
    }
    p->res_all = nres * 8;
    TRY(palloc(p, &p->d_res, (size_t)p->res_all));
    {
        void *h = nullptr;
        if (hipHostMalloc(&h, (size_t)p->res_all, hipHostMallocDefault) != hipSuccess)
            TRY(fail(ctx, RMS_ERR_OUT_OF_MEMORY, "hipHostMalloc of the results mirror failed"));
        p->h_res = (double *)h;
    }
    p->d_counters = (unsigned long long *)(p->d_res + cnt_off);
    for (int j = 0; j < njobs; j++) {
        JobDesc &D = p->hjobs[j];
        JobHost &H = p->jobs[j];
        D.counters = p->d_counters + NCOUNT * j;
        double *part;
        const int64_t nq = T * H.nval;
        TRY(palloc(p, &part, sizeof(double) * H.nbatch * nq));
        D.part = part;
        p->red.push_back(ReduceDesc{part, p->d_res + H.rec_rep, p->d_res + H.rec_tot, H.nbatch, nq});
        p->max_nq = std::max(p->max_nq, nq);
        D.psig = s_psig;
        if (H.ngrad > 0) {
            double *ps;
            TRY(palloc(p, &ps, sizeof(double) * H.nbatch * H.ngrad));
            D.psig = ps;
            p->red.push_back(ReduceDesc{ps, p->d_res + H.sig_rep, p->d_res + H.sig_tot, H.nbatch, H.ngrad});
            p->max_nq = std::max(p->max_nq, (int64_t)H.ngrad);
        }
    }
    p->maxbound = maxbound;
    if (sig & SIG_PGSE) {
        TRY(palloc(p, &p->d_snap, sizeof(double) * (size_t)(maxbound + 1) * 3 * p->ws));
        for (int j = 0; j < njobs; j++) p->hjobs[j].snap = p->d_snap;
    }
    // every table the chosen variant reads exists for every job
    for (int j = 0; j < njobs; j++)
        if (const char *e = check_desc(p->hjobs[j], sig)) TRY(fail(ctx, RMS_ERR_STATE, "job " + std::to_string(j) + ": " + e));
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
        if (sig & SIG_PGSE) {
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
    // the diagnostic counters: unless skipped, and always with the final
    // walker state (whose per-walker counters they are); never for the
    // central differences' walks, which report only records
    a.count = (!p->fd_walks && (p->prm.skip_counters == 0 || p->fin_alloc)) ? 1 : 0;
    return a;
}

// Enqueue the whole run: zero counters, walk chunks, ordered reductions.
static rms_status enqueue(rms_plan *p, bool with_events)
{
    rms_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, launch_fill_zero(p->d_counters, sizeof(unsigned long long) * NCOUNT * (int64_t)p->jobs.size(), s));
    if (with_events) HIPCHK(ctx, hipEventRecord(p->ev[0], s));
    const int64_t TN = p->prm.TN;
    const int64_t ch = (p->prm.chunk_steps > 0 && p->prm.chunk_steps < TN) ? p->prm.chunk_steps : TN;
    for (int64_t s0 = 0; s0 < TN; s0 += ch) {
        int64_t s1 = std::min(TN, s0 + ch);
        LaunchArgs a = make_args(p, s0, s1);
        // the lean variants only for the plan's own physics: a perturbed D
        // (the central differences) may make a pair draw
        const int sig = p->sig | ((p->lean && !p->fd_walks) ? SIG_LEAN : 0);
        HIPCHK(ctx, launch_walk(a, p->prm.precision == RMS_PREC_F64, sig, p->grid, p->block, s));
    }
    if (with_events) HIPCHK(ctx, hipEventRecord(p->ev[1], s));
    HIPCHK(ctx, launch_reduce(p->d_red, (int32_t)p->red.size(), p->max_nq, p->prm.replicas, s));
    if (with_events) HIPCHK(ctx, hipEventRecord(p->ev[2], s));
    return RMS_OK;
}

static rms_status init_failures(rms_ctx *ctx, const unsigned long long *cnt, int nj)
{
    for (int j = 0; j < nj; j++)
        if (cnt[NCOUNT * j + 5] != 0)
            return fail(ctx, RMS_ERR_STATE,
                        "job " + std::to_string(j) + ": " + std::to_string(cnt[NCOUNT * j + 5]) +
                            " walkers found no non-null start voxel in 2^26 trials (walkable fraction too small)");
    return RMS_OK;
}

static rms_status plan_execute_impl(rms_plan *p, rms_job_out *outs, rms_timing *timing)
{
    rms_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int nj = (int)p->jobs.size();
    // final-state outputs requested?
    bool want_fin = false, want_rep = false;
    if (outs)
        for (int j = 0; j < nj; j++) {
            want_fin = want_fin || outs[j].fin_g;
            want_rep = want_rep || outs[j].rec_rep || outs[j].sig_rep;
        }
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
    // one copy brings back every total and counter (and the replica sums
    // when any job asks for them)
    const int64_t nbytes = want_rep ? p->res_all : p->res_head;
    HIPCHK(ctx, hipMemcpyAsync(p->h_res, p->d_res, (size_t)nbytes, hipMemcpyDeviceToHost, s));
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
        for (const JobHost &H : p->jobs) timing->walker_steps += H.nwalk * TN;
    }
    const unsigned long long *cnt = (const unsigned long long *)(p->h_res) + (p->d_counters - (unsigned long long *)p->d_res);
    if (rms_status st = init_failures(ctx, cnt, nj)) return st;
    if (!outs) return RMS_OK;
    const int R = p->prm.replicas;
    const int64_t T = p->prm.timepoints;
    for (int j = 0; j < nj; j++) {
        const JobHost &H = p->jobs[j];
        rms_job_out &O = outs[j];
        const int64_t nq = T * H.nval;
        if (O.rec) std::memcpy(O.rec, p->h_res + H.rec_tot, sizeof(double) * nq);
        if (O.rec_rep) std::memcpy(O.rec_rep, p->h_res + H.rec_rep, sizeof(double) * nq * R);
        if (H.ngrad > 0) {
            if (O.sig) std::memcpy(O.sig, p->h_res + H.sig_tot, sizeof(double) * H.ngrad);
            if (O.sig_rep) std::memcpy(O.sig_rep, p->h_res + H.sig_rep, sizeof(double) * H.ngrad * R);
        }
        for (int k = 0; k < 5; k++) O.counters[k] = cnt[NCOUNT * j + k];
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
    const size_t nres = (size_t)nj * NPARAM * (1 + R);
    if ((st = palloc(p, &p->d_adj_res, sizeof(double) * nres)) != RMS_OK) return st;
    if (!p->h_adj_res) {
        void *h = nullptr;
        if (hipHostMalloc(&h, sizeof(double) * nres, hipHostMallocDefault) != hipSuccess)
            return fail(ctx, RMS_ERR_OUT_OF_MEMORY, "hipHostMalloc of the gradient mirror failed");
        p->h_adj_res = (double *)h;
    }
    for (int j = 0; j < nj; j++) {
        const JobHost &H = p->jobs[j];
        double *pg;
        if ((st = palloc(p, &p->cot_rec[j], sizeof(double) * T * H.nval)) != RMS_OK) return st;
        if (H.ngrad > 0 && (st = palloc(p, &p->cot_sig[j], sizeof(double) * H.ngrad)) != RMS_OK) return st;
        if ((st = palloc(p, &pg, sizeof(double) * H.nbatch * NPARAM)) != RMS_OK) return st;
        double *tot = p->d_adj_res + (size_t)j * NPARAM;
        double *rep = p->d_adj_res + (size_t)nj * NPARAM + (size_t)j * R * NPARAM;
        p->red_adj.push_back(ReduceDesc{pg, rep, tot, H.nbatch, NPARAM});
    }
    if (p->sig & SIG_PGSE) {
        const size_t n = (size_t)(p->maxbound + 1) * p->adj_np * 3 * p->ws;
        if ((st = palloc(p, &p->d_tsnap, sizeof(double) * n)) != RMS_OK) return st;
    }
    if ((st = palloc(p, &p->d_jobs_adj, sizeof(JobDesc) * nj)) != RMS_OK) return st;
    if ((st = pupload(p, &p->d_red_adj, p->red_adj.data(), p->red_adj.size())) != RMS_OK) return st;
    p->adj_ready = true;
    return RMS_OK;
}

// Workgroups of the tangent variant adj_sig: the resident limit (or the
// plan's grid_blocks), a multiple of 8 when batches are split per XCD.
static int adjoint_grid(rms_plan *p, int adj_sig)
{
    int &g = p->adj_grid[(adj_sig & SIG_JUMP) ? 1 : 0];
    if (g > 0) return g;
    int bpc = 0;
    const bool f64 = p->prm.precision == RMS_PREC_F64;
    if (adjoint_occupancy(f64, adj_sig, p->adj_np, p->block, &bpc) != hipSuccess || bpc < 1) bpc = 1;
    const int wpb = p->block / WAVE;
    const int64_t need = (p->nbatch + wpb - 1) / wpb;
    int64_t grid = p->prm.grid_blocks > 0 ? p->prm.grid_blocks : std::min(need, (int64_t)bpc * p->ctx->ncu);
    if (grid < 1) grid = 1;
    if (p->xcd_split) grid = std::max<int64_t>(8, (grid / 8) * 8);
    g = (int)grid;
    return g;
}

// Objective of one job's forward results: sum(cot_rec * rec) + sum(cot_sig * sig)
// over a record block rec [T * nval] and a signal block sig [ngrad].
static double contract(const rms_job_adj &A, const double *rec, int64_t nq, const double *sig, int32_t ngrad)
{
    double f = 0.0;
    if (A.cot_rec)
        for (int64_t q = 0; q < nq; q++) f = f + A.cot_rec[q] * rec[q];
    if (A.cot_sig && ngrad > 0)
        for (int32_t g = 0; g < ngrad; g++) f = f + A.cot_sig[g] * sig[g];
    return f;
}

// The plan's launch layout, put back by the destructor whatever happened
// while the central differences walked their own.
namespace {
struct LayoutGuard {
    rms_plan *p;
    int64_t nbatch;
    int32_t *bj;
    ReduceDesc *d_red;
    std::vector<ReduceDesc> red;
    int grid;
    explicit LayoutGuard(rms_plan *q)
        : p(q), nbatch(q->nbatch), bj(q->d_batch_job), d_red(q->d_red), red(q->red), grid(q->grid)
    {
        p->fd_walks = true;
    }
    ~LayoutGuard()
    {
        p->fd_walks = false;
        p->nbatch = nbatch;
        p->d_batch_job = bj;
        p->d_red = d_red;
        p->red = red;
        p->grid = grid;
    }
};
}  // namespace

// Launch layout of the walks for class c on walker subsets: job j (of those
// in `which` with class c) walks the walkers whose tangent-pass marks meet
// fd_walker_mask, in the order of their index, packed into whole batches
// from batch 0; jobs with none are left out (inc[j] = 0, zero[j] = 1: their
// difference is exactly zero), and a job whose mask cannot be trusted walks
// every walker.  Descriptors go to `hj`, the batch table, walker lists and
// reductions to the plan's subset buffers, and the plan points at them.  The
// plan's own reductions (red0) and grid (grid0) are the starting point every
// time: the plan may still point at an earlier class's subsets.
static rms_status fd_subset_layout(rms_plan *p, const std::vector<char> &which, int c, double h,
                                   const std::vector<uint32_t> &vis, const std::vector<ReduceDesc> &red0,
                                   int grid0, std::vector<JobDesc> &hj, std::vector<char> &inc,
                                   std::vector<char> &zero)
{
    rms_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const int nj = (int)p->jobs.size();
    hj = p->hjobs;
    inc.assign(nj, 0);
    zero.assign(nj, 0);
    std::vector<uint32_t> list;
    std::vector<int32_t> bj;
    std::vector<ReduceDesc> rd;
    int64_t nb = 0;
    size_t ri = 0;  // the plan's reductions: per job the records, then the PGSE signal
    for (int j = 0; j < nj; j++) {
        const JobHost &H = p->jobs[j];
        const size_t r0 = ri;
        ri += H.ngrad > 0 ? 2 : 1;
        if (!which[j] || p->hjobs[j].nclass <= c) continue;
        uint32_t mask = 0;
        const bool all = !fd_walker_mask(ctx->geoms[H.geom_id].meta, p->prm.dt, c, h, mask);
        const int64_t at = nb * WAVE;
        int64_t m = H.nwalk;
        if (!all) {
            list.resize((size_t)at);
            for (int64_t k = 0; k < H.nwalk; k++)
                if ((vis[(size_t)(H.wbase + k)] & mask) != 0u) list.push_back((uint32_t)k);
            m = (int64_t)list.size() - at;
        }
        if (m == 0) {
            zero[j] = 1;
            continue;
        }
        const int64_t nbj = (m + WAVE - 1) / WAVE;
        // padding lanes walk the last listed walker again (inactive: no records)
        list.resize((size_t)(at + nbj * WAVE), all ? 0u : list.back());
        JobDesc &D = hj[j];
        D.nwalk = m;
        D.batch0 = nb;
        D.wbase = at;
        D.wlist = all ? nullptr : p->d_wlist + at;
        for (int64_t b = 0; b < nbj; b++) bj.push_back(j);
        for (size_t r = r0; r < ri; r++) {
            ReduceDesc e = red0[r];
            e.nb = nbj;
            rd.push_back(e);
        }
        inc[j] = 1;
        nb += nbj;
    }
    if (nb == 0) return RMS_OK;
    HIPCHK(ctx, hipMemcpyAsync(p->d_wlist, list.data(), sizeof(uint32_t) * list.size(), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(p->d_bj_fd, bj.data(), sizeof(int32_t) * bj.size(), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(p->d_red_fd, rd.data(), sizeof(ReduceDesc) * rd.size(), hipMemcpyHostToDevice, s));
    // the host vectors die here: the copies must have read them
    HIPCHK(ctx, hipStreamSynchronize(s));
    p->nbatch = nb;
    p->d_batch_job = p->d_bj_fd;
    p->d_red = p->d_red_fd;
    p->red = rd;
    // the plan's grid, cut to the batches (a multiple of 8 when split per XCD)
    const int wpb = p->block / WAVE;
    int64_t need = (nb + wpb - 1) / wpb;
    if (p->xcd_split) need = (need + 7) / 8 * 8;
    p->grid = (int)std::min<int64_t>(grid0, need);
    return RMS_OK;
}

// D entries by central differences with common random numbers: for each
// non-null class c, the forward walk of the jobs in `which` with D_c set to
// D_c (1 + h) and D_c (1 - h) (every seed unchanged), each job's objective
// contracted per replica, and dF/dD_c = (F+ - F-) / (D+ - D-).  With
// `subsets` the walks for class c take only the walkers whose walk can
// change with D_c (fd_walker_mask on the tangent pass's marks): the others
// walk the same at both points and add nothing to F+ - F-.  Without it the
// whole plan walks.  A perturbation that leaves the plan's kernel variant (a
// new exchange pair) or the step limit gives NaN.  The plan's own
// descriptors and layout are restored before returning.
static rms_status fd_d_entries(rms_plan *p, const rms_job_adj *adj, const std::vector<char> &which, double h,
                               bool subsets, double *walk_ms, int *launches, int64_t *walked)
{
    rms_ctx *ctx = p->ctx;
    hipStream_t s = ctx->stream;
    const int nj = (int)p->jobs.size();
    const int R = p->prm.replicas;
    const int64_t T = p->prm.timepoints;
    int maxc = 0;
    for (int j = 0; j < nj; j++)
        if (which[j]) maxc = std::max(maxc, (int)p->hjobs[j].nclass - 1);
    const int64_t TN = p->prm.TN;
    const int64_t ch = (p->prm.chunk_steps > 0 && p->prm.chunk_steps < TN) ? p->prm.chunk_steps : TN;
    std::vector<double> Fp((size_t)nj * (1 + R)), Fm((size_t)nj * (1 + R));
    std::vector<double> Dp(nj), Dm(nj);
    std::vector<uint32_t> vis;
    if (subsets) {
        vis.resize((size_t)p->ws);
        HIPCHK(ctx, hipMemcpyAsync(vis.data(), p->d_wvis, sizeof(uint32_t) * vis.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipStreamSynchronize(s));
    }
    rms_status st = RMS_OK;
    {
        LayoutGuard keep(p);
        for (int c = 1; c <= maxc && st == RMS_OK; c++) {
            std::vector<char> bad(nj, 0), inc(nj, 1), zero(nj, 0);
            std::vector<JobDesc> base;
            if (subsets) {
                if ((st = fd_subset_layout(p, which, c, h, vis, keep.red, keep.grid, base, inc, zero)) != RMS_OK)
                    break;
            } else {
                base = p->hjobs;
            }
            bool any = false;
            for (int j = 0; j < nj; j++) any = any || (inc[j] && which[j] && p->hjobs[j].nclass > c);
            for (int side = 0; side < 2 && st == RMS_OK && any; side++) {
                std::vector<JobDesc> hj = base;
                for (int j = 0; j < nj; j++) {
                    if (!which[j] || p->hjobs[j].nclass <= c || !inc[j]) continue;
                    rms_geometry g = ctx->geoms[p->jobs[j].geom_id].meta;
                    const double D0 = g.D[c];
                    g.D[c] = side == 0 ? D0 * (1.0 + h) : D0 * (1.0 - h);
                    (side == 0 ? Dp : Dm)[j] = g.D[c];
                    int bits = 0;
                    bool exch = false;
                    JobDesc D = base[j];
                    const char *e = job_physics(g, p->prm.dt, D, bits, exch);
                    if (e || (bits & ~p->sig & (SIG_EXCH | SIG_PERIODIC)) != 0) {
                        bad[j] = 1;
                        continue;
                    }
                    hj[j] = D;
                }
                if (hipMemcpyAsync(p->d_jobs, hj.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, s) !=
                    hipSuccess) {
                    st = fail(ctx, RMS_ERR_HIP, "adjoint: descriptor upload failed");
                    break;
                }
                if ((st = enqueue(p, true)) != RMS_OK) break;
                if (hipMemcpyAsync(p->h_res, p->d_res, (size_t)p->res_all, hipMemcpyDeviceToHost, s) != hipSuccess ||
                    hipStreamSynchronize(s) != hipSuccess) {
                    st = fail(ctx, RMS_ERR_HIP, "adjoint: forward walk at a perturbed D failed");
                    break;
                }
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, p->ev[0], p->ev[1]) == hipSuccess) *walk_ms += ms;
                *launches += (int)((TN + ch - 1) / ch);
                for (int j = 0; j < nj; j++)
                    if (!subsets || inc[j]) *walked += hj[j].nwalk * TN;
                std::vector<double> &F = side == 0 ? Fp : Fm;
                for (int j = 0; j < nj; j++) {
                    if (!which[j] || p->hjobs[j].nclass <= c || !inc[j]) continue;
                    const JobHost &H = p->jobs[j];
                    const int64_t nq = T * H.nval;
                    F[(size_t)j * (1 + R)] =
                        contract(adj[j], p->h_res + H.rec_tot, nq, p->h_res + H.sig_tot, H.ngrad);
                    for (int r = 0; r < R; r++)
                        F[(size_t)j * (1 + R) + 1 + r] =
                            contract(adj[j], p->h_res + H.rec_rep + (int64_t)r * nq, nq,
                                     p->h_res + H.sig_rep + (int64_t)r * H.ngrad, H.ngrad);
                }
            }
            if (st != RMS_OK) break;
            for (int j = 0; j < nj; j++) {
                if (!which[j] || p->hjobs[j].nclass <= c) continue;
                const double den = Dp[j] - Dm[j];
                const double *fp = &Fp[(size_t)j * (1 + R)], *fm = &Fm[(size_t)j * (1 + R)];
                // no walker of the job can walk differently: the difference is zero
                const double g0 = bad[j] ? NAN : (zero[j] ? 0.0 : (fp[0] - fm[0]) / den);
                adj[j].grad[c] = g0;
                if (adj[j].grad_rep)
                    for (int r = 0; r < R; r++)
                        adj[j].grad_rep[(size_t)r * NPARAM + c] =
                            bad[j] ? NAN : (zero[j] ? 0.0 : (fp[1 + r] - fm[1 + r]) / den);
            }
        }
    }
    // the plan's own descriptors, whatever happened above
    if (hipMemcpyAsync(p->d_jobs, p->hjobs.data(), sizeof(JobDesc) * nj, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        if (st == RMS_OK) st = fail(ctx, RMS_ERR_HIP, "adjoint: descriptor restore failed");
    }
    return st;
}

static rms_status plan_adjoint_impl(rms_plan *p, const rms_job_adj *adj, const rms_adjoint_opts *opts,
                                    rms_timing *timing)
{
    rms_ctx *ctx = p->ctx;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const int nj = (int)p->jobs.size();
    for (int j = 0; j < nj; j++)
        if (!adj[j].grad) return fail(ctx, RMS_ERR_INVALID_ARG, "adjoint: grad output is required for every job");
    const int mode = opts ? opts->d_estimator : RMS_D_AUTO;
    if (mode < RMS_D_AUTO || mode > RMS_D_PATHWISE)
        return fail(ctx, RMS_ERR_INVALID_ARG, "adjoint: d_estimator must be one of RMS_D_AUTO .. RMS_D_PATHWISE");
    const double h = (opts && opts->fd_rel_step != 0.0) ? opts->fd_rel_step : 0.05;
    if (!(h > 0.0 && h < 0.5)) return fail(ctx, RMS_ERR_INVALID_ARG, "adjoint: fd_rel_step must be in (0, 0.5)");
    if (!p->adj_ready) {
        rms_status st = adjoint_setup(p);
        if (st != RMS_OK) return st;
    }
    // the tangent variant: boundary terms only when asked for (and only
    // without a PGSE table, where the continuations would need PGSE sums)
    const int adj_sig = mode == RMS_D_BOUNDARY ? adjoint_sig(p->sig) : p->sig;
    const int grid = adjoint_grid(p, adj_sig);
    hipStream_t s = ctx->stream;
    const int64_t T = p->prm.timepoints;
    // jobs whose D entries come from central differences: every job with
    // RMS_D_CRN_FD, those not continuous in D (exchange faces, lattice-edge
    // mirrors, voxel edges) with RMS_D_AUTO
    std::vector<char> fd(nj, 0);
    bool any_fd = false;
    for (int j = 0; j < nj; j++) {
        const bool jumps = p->jobs[j].exchange || p->ctx->geoms[p->jobs[j].geom_id].flags != 0;
        fd[j] = (mode == RMS_D_CRN_FD || (mode == RMS_D_AUTO && jumps)) ? 1 : 0;
        any_fd = any_fd || fd[j];
    }
    const bool subsets = any_fd && !(opts && opts->fd_all_walkers);
    if (subsets && !p->d_red_fd) {
        rms_status st;
        if ((st = palloc(p, &p->d_wvis, sizeof(uint32_t) * p->ws)) != RMS_OK) return st;
        if ((st = palloc(p, &p->d_wlist, sizeof(uint32_t) * p->ws)) != RMS_OK) return st;
        if ((st = palloc(p, &p->d_bj_fd, sizeof(int32_t) * p->nbatch)) != RMS_OK) return st;
        if ((st = palloc(p, &p->d_red_fd, sizeof(ReduceDesc) * p->red.size())) != RMS_OK) return st;
    }
    std::vector<JobDesc> hj = p->hjobs;
    for (int j = 0; j < nj; j++) {
        const JobHost &H = p->jobs[j];
        JobDesc &D = hj[j];
        D.wvis = (subsets && fd[j]) ? p->d_wvis + H.wbase : nullptr;
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
    a.count = 0;  // the tangent pass reports no counters
    a.st.g = nullptr;  // single launch: no walker state is saved
    a.fin = FinalOut{};
    HIPCHK(ctx, launch_adjoint(a, p->prm.precision == RMS_PREC_F64, adj_sig, p->adj_np, grid, p->block, s));
    HIPCHK(ctx, hipEventRecord(p->ev[1], s));
    HIPCHK(ctx, launch_reduce(p->d_red_adj, (int32_t)p->red_adj.size(), NPARAM, p->prm.replicas, s));
    HIPCHK(ctx, hipEventRecord(p->ev[2], s));
    const int R = p->prm.replicas;
    bool want_rep = false;
    for (int j = 0; j < nj; j++) want_rep = want_rep || adj[j].grad_rep;
    const size_t nbytes = sizeof(double) * (size_t)nj * NPARAM * (want_rep ? 1 + R : 1);
    HIPCHK(ctx, hipMemcpyAsync(p->h_adj_res, p->d_adj_res, nbytes, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    double walk_ms = 0.0, reduce_ms = 0.0, total_ms = 0.0;
    {
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev[0], p->ev[2]));
        total_ms = ms;
        HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
        walk_ms = ms;
        HIPCHK(ctx, hipEventElapsedTime(&ms, p->ev[1], p->ev[2]));
        reduce_ms = ms;
    }
    int launches = 1;
    int64_t walked = 0;
    for (int j = 0; j < nj; j++) {
        walked += p->jobs[j].nwalk * p->prm.TN;
        std::memcpy(adj[j].grad, p->h_adj_res + (size_t)j * NPARAM, sizeof(double) * NPARAM);
        if (adj[j].grad_rep)
            std::memcpy(adj[j].grad_rep, p->h_adj_res + (size_t)nj * NPARAM + (size_t)j * R * NPARAM,
                        sizeof(double) * NPARAM * R);
        auto nan_entry = [&](int c) {
            adj[j].grad[c] = NAN;
            if (adj[j].grad_rep)
                for (int r = 0; r < R; r++) adj[j].grad_rep[(size_t)r * NPARAM + c] = NAN;
        };
        // exchange faces, lattice-edge mirrors and voxel edges make the
        // walk's end discontinuous in D
        const bool jumps = p->jobs[j].exchange || p->ctx->geoms[p->jobs[j].geom_id].flags != 0;
        if (!fd[j] && jumps && !(adj_sig & SIG_JUMP)) {
            for (int c = 1; c < NC; c++) nan_entry(c);
        }
        if (p->jobs[j].kappa_closed) {
            // a pair closed by its kappa draws nothing, so the score function
            // sees no draw although P grows with kappa: no estimate
            nan_entry(8);
        }
    }
    if (any_fd) {
        const auto t0 = std::chrono::steady_clock::now();
        double fd_ms = 0.0;
        rms_status st = fd_d_entries(p, adj, fd, h, subsets, &fd_ms, &launches, &walked);
        if (st != RMS_OK) return st;
        walk_ms += fd_ms;
        total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (timing) {
        std::memset(timing, 0, sizeof(*timing));
        timing->total_ms = total_ms;
        timing->walk_ms = walk_ms;
        timing->reduce_ms = reduce_ms;
        timing->grid = grid;
        timing->block = p->block;
        timing->launches = launches;
        timing->walker_steps = walked;
    }
    return RMS_OK;
}

extern "C" rms_status rms_plan_adjoint(rms_plan *p, const rms_job_adj *adj, rms_timing *timing)
{
    if (!p || !adj) return RMS_ERR_INVALID_ARG;
    return guard(p->ctx, [&] { return plan_adjoint_impl(p, adj, nullptr, timing); });
}

extern "C" rms_status rms_plan_adjoint_ex(rms_plan *p, const rms_job_adj *adj, const rms_adjoint_opts *opts,
                                          rms_timing *timing)
{
    if (!p || !adj) return RMS_ERR_INVALID_ARG;
    return guard(p->ctx, [&] { return plan_adjoint_impl(p, adj, opts, timing); });
}

// ---------------------------------------------------------- allocation free
// rms_run for one job with every device buffer in the scratch array.  The
// host side uses only the stack and the context's preallocated message
// buffer: no hipMalloc / hipFree, no heap.
namespace {
struct NoAllocFail {
    rms_ctx *ctx;
    rms_status operator()(rms_status s, const char *msg) const
    {
        ctx->err.assign(msg);  // fits the capacity reserved at creation
        return s;
    }
    rms_status hip(const char *what, hipError_t e) const
    {
        ctx->err.assign(what);
        ctx->err.append(": ");
        ctx->err.append(hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? RMS_ERR_OUT_OF_MEMORY : RMS_ERR_HIP;
    }
};
}  // namespace

static rms_status run_noalloc_impl(rms_ctx *ctx, const rms_job *job, const rms_run_params *prm, rms_job_out *out,
                                   rms_timing *timing)
{
    const NoAllocFail bad{ctx};
#define NA_HIP(call)                                   \
    do {                                               \
        const hipError_t e_ = (call);                  \
        if (e_ != hipSuccess) return bad.hip(#call, e_); \
    } while (0)
    if (!job || !prm || !out) return bad(RMS_ERR_INVALID_ARG, "null argument");
    if (!ctx->scratch) return bad(RMS_ERR_STATE, "no scratch array on this device");
    const rms_run_params &P = *prm;
    if (const char *e = check_params(P)) return bad(RMS_ERR_INVALID_ARG, e);
    if (P.use_graph) return bad(RMS_ERR_INVALID_ARG, "rms_run_noalloc: use_graph must be 0");
    if (P.chunk_steps > 0 && P.chunk_steps < P.TN) return bad(RMS_ERR_INVALID_ARG, "rms_run_noalloc: one launch only (chunk_steps 0)");
    if (out->fin_g || out->fin_o || out->fin_gi || out->fin_oi || out->fin_c || out->fin_cnt)
        return bad(RMS_ERR_INVALID_ARG, "rms_run_noalloc: no final-state outputs");
    if (job->geom_id < 0 || job->geom_id >= (int32_t)ctx->geoms.size()) return bad(RMS_ERR_INVALID_ARG, "bad geom_id");
    if (job->ngrad > NOALLOC_MAX_GRAD) return bad(RMS_ERR_INVALID_ARG, "rms_run_noalloc: ngrad must be <= 2000");
    const int R = P.replicas < 1 ? 1 : P.replicas;
    const int32_t T = P.timepoints;
    const int64_t Tstep = P.TN / P.timepoints;
    const GeomDev &G = ctx->geoms[job->geom_id];
    JobDesc D;
    int sig = 0;
    bool exch = false;
    if (const char *e = fill_job(G.meta, *job, P, D, sig, exch)) return bad(RMS_ERR_INVALID_ARG, e);
    if (job->nbvec > 0) sig |= SIG_NARROW;
    if (job->ngrad > 0) sig |= SIG_PGSE;
    if (G.d_btab) sig |= SIG_BRICKS;
    if (G.wrapcells) sig |= SIG_WRAP;
    // PGSE lobe tables on the stack
    int32_t perm[NOALLOC_MAX_GRAD], trip[4 * (NOALLOC_MAX_TRIPLE + 1)];
    uint64_t keys[NOALLOC_MAX_GRAD];
    int64_t bounds[MAX_BOUNDS + 1];
    int32_t nbound = 0, ntriple = 0;
    if (job->ngrad > 0) {
        if (const char *e = pgse_tables(job->grad, job->ngrad, P.dt, P.TN, perm, trip, NOALLOC_MAX_TRIPLE, bounds,
                                        nbound, ntriple, keys))
            return bad(RMS_ERR_INVALID_ARG, e);
    }
    const int64_t nbatch = (job->nwalk + WAVE - 1) / WAVE;
    const int32_t nval = 1 + G.meta.nclass + 21 + job->nbvec;
    ScratchLayout L;
    if (const char *e = scratch_layout(nbatch, T, nval, R, job->ngrad, nbound, ntriple, job->nbvec, SCRATCH_BYTES, L))
        return bad(RMS_ERR_INVALID_ARG, e);
    unsigned char *S = ctx->scratch;
    auto dp = [&](int64_t off) { return (void *)(S + off); };
    D.cells = G.d_cells;
    D.cls = G.d_cls;
    D.btab = G.d_btab;
    D.nby = G.nby;
    D.nbz = G.nbz;
    D.wrapcells = G.wrapcells ? 1 : 0;
    D.batch0 = 0;
    D.wbase = 0;
    D.part = (double *)dp(L.part);
    D.psig = (double *)dp(L.psig);
    D.snap = (double *)dp(L.snap);
    D.counters = (unsigned long long *)dp(L.counters);
    D.grad = (const double *)dp(L.grad);
    D.gperm = (const int32_t *)dp(L.perm);
    D.triple = (const int32_t *)dp(L.triple);
    D.bounds = (const int64_t *)dp(L.bounds);
    D.qscale = (const double *)dp(L.qscale);
    D.bvec = (const double *)dp(L.bvec);
    D.nbound = nbound;
    D.ntriple = ntriple;
    if (const char *e = check_desc(D, sig)) return bad(RMS_ERR_STATE, e);
    ReduceDesc red[2];
    int nred = 1;
    red[0] = ReduceDesc{D.part, (double *)dp(L.rep), (double *)dp(L.tot), nbatch, (int64_t)T * nval};
    int64_t max_nq = (int64_t)T * nval;
    if (job->ngrad > 0) {
        red[1] = ReduceDesc{D.psig, (double *)dp(L.srep), (double *)dp(L.stot), nbatch, job->ngrad};
        nred = 2;
        max_nq = std::max<int64_t>(max_nq, job->ngrad);
    }
    // launch shape: persistent grid at the resident limit
    const int block = P.block_threads > 0 ? P.block_threads : 256;
    if (block % WAVE != 0 || block > 256) return bad(RMS_ERR_INVALID_ARG, "block_threads must be a multiple of 64 and <= 256");
    const bool f64 = P.precision == RMS_PREC_F64;
    int bpc = 0;
    if (walk_occupancy(f64, sig, block, &bpc) != hipSuccess || bpc < 1) bpc = 1;
    const int64_t need = (nbatch + block / WAVE - 1) / (block / WAVE);
    int64_t grid = P.grid_blocks > 0 ? P.grid_blocks : std::min(need, (int64_t)bpc * ctx->ncu);
    if (grid < 1) grid = 1;
    const int xcd = (P.xcd_split > 0 && grid >= 8) ? 1 : 0;
    if (xcd) grid = grid / 8 * 8;

    NA_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    // tables and descriptors in (the stack buffers stay live until the
    // synchronisation below), sentinels and counters cleared, then the walk
    NA_HIP(hipMemcpyAsync(dp(L.job), &D, sizeof(D), hipMemcpyHostToDevice, s));
    NA_HIP(hipMemcpyAsync(dp(L.red), red, sizeof(ReduceDesc) * nred, hipMemcpyHostToDevice, s));
    NA_HIP(launch_fill_zero(dp(L.batch_job), (nbatch * 4 + 7) / 8 * 8, s));  // one job: every batch is job 0
    NA_HIP(launch_fill_zero(dp(L.counters), 8 * NCOUNT, s));
    if (job->ngrad > 0) {
        NA_HIP(hipMemcpyAsync(dp(L.grad_raw), job->grad, sizeof(double) * 6 * job->ngrad, hipMemcpyHostToDevice, s));
        NA_HIP(hipMemcpyAsync(dp(L.perm), perm, sizeof(int32_t) * job->ngrad, hipMemcpyHostToDevice, s));
        NA_HIP(hipMemcpyAsync(dp(L.triple), trip, sizeof(int32_t) * 4 * (ntriple + 1), hipMemcpyHostToDevice, s));
        NA_HIP(hipMemcpyAsync(dp(L.bounds), bounds, sizeof(int64_t) * (nbound + 1), hipMemcpyHostToDevice, s));
        NA_HIP(launch_gather_rows((const double *)dp(L.grad_raw), (const int32_t *)dp(L.perm), job->ngrad, 6,
                                  (double *)dp(L.grad), s));
    } else {
        NA_HIP(launch_fill_zero(dp(L.bounds), 16, s));
        NA_HIP(launch_fill_zero(dp(L.triple), 32, s));
        NA_HIP(launch_fill_zero(dp(L.perm), 8, s));
        NA_HIP(launch_fill_zero(dp(L.grad), 48, s));
    }
    if (job->nbvec > 0) {
        NA_HIP(hipMemcpyAsync(dp(L.bval), job->bval, sizeof(double) * job->nbvec, hipMemcpyHostToDevice, s));
        NA_HIP(hipMemcpyAsync(dp(L.bvec), job->bvec, sizeof(double) * 3 * job->nbvec, hipMemcpyHostToDevice, s));
        NA_HIP(launch_qscale((const double *)dp(L.bval), job->nbvec, T, Tstep, P.dt, (double *)dp(L.qscale), s));
    } else {
        NA_HIP(launch_fill_zero(dp(L.qscale), 16, s));
        NA_HIP(launch_fill_zero(dp(L.bvec), 24, s));
    }
    LaunchArgs a{};
    a.jobs = (const JobDesc *)dp(L.job);
    a.batch_job = (const int32_t *)dp(L.batch_job);
    a.nbatch = nbatch;
    a.TN = P.TN;
    a.step0 = 0;
    a.step1 = P.TN;
    a.Tstep = Tstep;
    a.timepoints = T;
    a.dt = P.dt;
    a.st.ws = nbatch * WAVE;
    a.xcd_split = xcd;
    a.count = P.skip_counters == 0 ? 1 : 0;
    NA_HIP(hipEventRecord(ctx->ev[0], s));
    NA_HIP(launch_walk(a, f64, sig, (int)grid, block, s));
    NA_HIP(hipEventRecord(ctx->ev[1], s));
    NA_HIP(launch_reduce((const ReduceDesc *)dp(L.red), nred, max_nq, R, s));
    NA_HIP(hipEventRecord(ctx->ev[2], s));
    unsigned long long cnt[NCOUNT];
    NA_HIP(hipMemcpyAsync(cnt, dp(L.counters), sizeof(cnt), hipMemcpyDeviceToHost, s));
    if (out->rec) NA_HIP(hipMemcpyAsync(out->rec, dp(L.tot), sizeof(double) * T * nval, hipMemcpyDeviceToHost, s));
    if (out->rec_rep)
        NA_HIP(hipMemcpyAsync(out->rec_rep, dp(L.rep), sizeof(double) * R * T * nval, hipMemcpyDeviceToHost, s));
    if (job->ngrad > 0 && out->sig)
        NA_HIP(hipMemcpyAsync(out->sig, dp(L.stot), sizeof(double) * job->ngrad, hipMemcpyDeviceToHost, s));
    if (job->ngrad > 0 && out->sig_rep)
        NA_HIP(hipMemcpyAsync(out->sig_rep, dp(L.srep), sizeof(double) * R * job->ngrad, hipMemcpyDeviceToHost, s));
    NA_HIP(hipStreamSynchronize(s));
    if (cnt[5] != 0) return bad(RMS_ERR_STATE, "walkers found no non-null start voxel in 2^26 trials (walkable fraction too small)");
    for (int k = 0; k < 5; k++) out->counters[k] = cnt[k];
    if (timing) {
        std::memset(timing, 0, sizeof(*timing));
        float ms = 0.f;
        NA_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[2]));
        timing->total_ms = ms;
        NA_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
        timing->walk_ms = ms;
        NA_HIP(hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]));
        timing->reduce_ms = ms;
        timing->grid = (int32_t)grid;
        timing->block = block;
        timing->launches = 1;
        timing->walker_steps = job->nwalk * prm->TN;
    }
    return RMS_OK;
#undef NA_HIP
}

extern "C" rms_status rms_run_noalloc(rms_ctx *ctx, const rms_job *job, const rms_run_params *params,
                                      rms_job_out *out, rms_timing *timing)
{
    if (!ctx) return RMS_ERR_INVALID_ARG;
    return guard(ctx, [&] { return run_noalloc_impl(ctx, job, params, out, timing); });
}
