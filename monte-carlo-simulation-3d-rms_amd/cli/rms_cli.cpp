// rms_cli.cpp -- shared driver of the two drop-in programs.
//
// rms_main       replaces hpc_code/lib/main.cu      (narrow-pulse b-table)
// rms_main_pgse  replaces hpc_code/lib/main_PGSE.cu (PGSE gradient table)
//
// With no arguments each reads the reference's input files from the working
// directory and writes the reference's output files and stdout lines there
// (main.cu:388-621, main_PGSE.cu:414-675).  Optional flags (not in the
// reference): --seed N (default: 8 bytes of /dev/urandom, as main.cu:491-496
// reads /dev/random), --precision f64|f32 (default f64, the reference's FP64
// positions), --timepoints N (default 1000, main.cu:35), --device N,
// --dir PATH, --periodic AXES (e.g. xy: those lattice edges wrap around
// instead of being the reference's mirrors).  Inputs the reference would
// crash or hang on are rejected with a message and exit status 1.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "rms_amd.h"
#include "rms_io.h"

namespace {

struct Opts {
    bool has_seed = false;
    uint64_t seed = 0;
    int precision = RMS_PREC_F64;
    int timepoints = RMS_REF_TIMEPOINTS;
    int device = 0;
    int periodic = 0;  // bit k: axis k wraps around
    std::string dir = ".";
};

int usage(const char *prog)
{
    std::fprintf(stderr,
                 "usage: %s [--seed N] [--precision f64|f32] [--timepoints N] [--device N] [--dir PATH] "
                 "[--periodic AXES]\n",
                 prog);
    return 2;
}

int die(const std::string &msg)
{
    std::fprintf(stderr, "error: %s\n", msg.c_str());
    return 1;
}

uint64_t urandom_seed()
{
    uint64_t s = 0;
    FILE *f = std::fopen("/dev/urandom", "rb");
    if (f) {
        if (std::fread(&s, sizeof(s), 1, f) != 1) s = 0;
        std::fclose(f);
    }
    if (s == 0) s = (uint64_t)std::chrono::steady_clock::now().time_since_epoch().count();
    return s;
}

int run_body(int argc, char **argv, bool pgse);

}  // namespace

// Exceptions (a lattice too large for host memory, a failed stream) end the
// program with a message and exit status 1, like the other errors.
int run_cli(int argc, char **argv, bool pgse)
{
    try {
        return run_body(argc, argv, pgse);
    } catch (const std::bad_alloc &) {
        return die("host memory allocation failed");
    } catch (const std::exception &e) {
        return die(e.what());
    }
}

namespace {

int run_body(int argc, char **argv, bool pgse)
{
    Opts o;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&](std::string &v) {
            if (i + 1 >= argc) return false;
            v = argv[++i];
            return true;
        };
        std::string v;
        if (a == "--seed" && next(v)) {
            o.has_seed = true;
            o.seed = std::strtoull(v.c_str(), nullptr, 0);
        } else if (a == "--precision" && next(v)) {
            if (v == "f64")
                o.precision = RMS_PREC_F64;
            else if (v == "f32")
                o.precision = RMS_PREC_F32;
            else
                return usage(argv[0]);
        } else if (a == "--timepoints" && next(v)) {
            o.timepoints = std::atoi(v.c_str());
        } else if (a == "--device" && next(v)) {
            o.device = std::atoi(v.c_str());
        } else if (a == "--dir" && next(v)) {
            o.dir = v;
        } else if (a == "--periodic" && next(v)) {
            for (char ch : v) {
                const char *axes = "xyz";
                const char *q = std::strchr(axes, ch);
                if (!q || ch == 0) return usage(argv[0]);
                o.periodic |= 1 << (int)(q - axes);
            }
        } else {
            return usage(argv[0]);
        }
    }
    auto path = [&](const char *name) { return o.dir + "/" + name; };
    std::string err;
    rmsio::SimParams sp;
    if (!rmsio::read_sim_params(path("simParamInput.txt"), pgse, sp, err)) return die(err);
    if (sp.Nc < 1 || sp.Nc > RMS_NC_MAX)
        return die("number of compartments must be 1.." + std::to_string(RMS_NC_MAX) + " (reference Nc_max = 3)");
    if (sp.NPix1 < 1 || sp.NPix2 < 1 || sp.NPix3 < 1) return die("lattice dimensions must be >= 1");
    if (sp.NPar < 1) return die("particle number must be >= 1");
    if (o.timepoints < 1 || sp.TN < o.timepoints)
        return die("step number must be >= the number of time points (the reference divides by TN/1000)");

    std::vector<double> D, T2;
    if (!rmsio::read_doubles(path("diffusivity.txt"), sp.Nc, D, err)) return die(err);
    for (int i = 0; i < sp.Nc; i++) std::cout << "diffusivity=" << D[i] << std::endl;
    if (!rmsio::read_doubles(path("T2.txt"), sp.Nc, T2, err)) return die(err);

    const size_t nvox = (size_t)sp.NPix1 * sp.NPix2 * sp.NPix3;
    std::vector<uint16_t> lab16;
    uint32_t maxlab = 0;
    // fiber.bin (raw labels) takes precedence over fiber.txt when present
    const bool have_bin = std::ifstream(path("fiber.bin")).good();
    if (have_bin) {
        if (!rmsio::read_labels_bin(path("fiber.bin"), nvox, lab16, maxlab, err)) return die(err);
    } else if (!rmsio::read_labels(path("fiber.txt"), nvox, lab16, maxlab, err)) {
        return die(err);
    }
    if (maxlab >= (uint32_t)sp.Nc) return die("the lattice holds a label >= the number of compartments");
    std::vector<uint8_t> lab8;
    rms_geometry g;
    std::memset(&g, 0, sizeof(g));
    g.nx = sp.NPix1;
    g.ny = sp.NPix2;
    g.nz = sp.NPix3;
    if (maxlab < 256) {
        lab8.assign(lab16.begin(), lab16.end());
        lab16.clear();
        lab16.shrink_to_fit();
        g.label_bytes = 1;
        g.labels = lab8.data();
    } else {
        g.label_bytes = 2;
        g.labels = lab16.data();
    }
    g.nclass = sp.Nc;
    g.res = sp.res;
    for (int c = 0; c < sp.Nc; c++) {
        g.D[c] = D[c];
        g.T2[c] = T2[c];
    }
    for (int i = 0; i < RMS_NC_MAX * RMS_NC_MAX; i++) g.kappa[i] = -1.0;
    g.periodic = o.periodic;
    g.storage = RMS_STORE_AUTO;
    double step[RMS_NC_MAX], T2s[RMS_NC_MAX], P[RMS_NC_MAX * RMS_NC_MAX];
    if (rms_derive(&g, sp.dt, step, T2s, P) != RMS_OK) return die("invalid physical parameters");
    for (int i = 0; i < sp.Nc; i++)
        for (int j = 0; j < sp.Nc; j++) std::cout << "permeation probability=" << P[i * RMS_NC_MAX + j] << std::endl;

    std::vector<double> bval, bvec, grad;
    int64_t ngrad = 0;
    if (!pgse) {
        if (sp.Nbvec < 0) return die("negative number of b-vectors");
        if (sp.Nbvec > 0) {
            if (!rmsio::read_doubles(path("bval.txt"), sp.Nbvec, bval, err)) return die(err);
            if (!rmsio::read_doubles(path("bvec.txt"), 3 * (size_t)sp.Nbvec, bvec, err)) return die(err);
        }
    } else {
        if (!rmsio::read_int(path("gradient_Ngrad.txt"), ngrad, err)) return die(err);
        if (ngrad < 0 || ngrad > (1 << 20)) return die("bad number of PGSE gradients");
        if (ngrad > 0 && !rmsio::read_doubles(path("gradient_time.txt"), 6 * (size_t)ngrad, grad, err))
            return die(err);
    }

    rms_ctx *ctx = nullptr;
    rms_device_opts dopt = {o.device, 0};
    rms_status s = rms_create(&dopt, &ctx);
    if (s != RMS_OK) return die("no usable GPU (rms_create status " + std::to_string((int)s) + ")");
    int32_t gid = -1;
    if (rms_add_geometry(ctx, &g, &gid) != RMS_OK) {
        std::string m = rms_last_error(ctx);
        rms_destroy(ctx);
        return die(m);
    }
    rms_job job;
    std::memset(&job, 0, sizeof(job));
    job.geom_id = gid;
    job.nwalk = sp.NPar;
    job.seed = o.has_seed ? o.seed : urandom_seed();
    job.nbvec = pgse ? 0 : sp.Nbvec;
    job.bval = bval.empty() ? nullptr : bval.data();
    job.bvec = bvec.empty() ? nullptr : bvec.data();
    job.ngrad = (int32_t)ngrad;
    job.grad = grad.empty() ? nullptr : grad.data();
    rms_run_params rp;
    std::memset(&rp, 0, sizeof(rp));
    rp.dt = sp.dt;
    rp.TN = sp.TN;
    rp.timepoints = o.timepoints;
    rp.precision = o.precision;
    rp.replicas = 1;
    rp.skip_counters = 1;  // the reference writes no diagnostic counters
    rms_plan *plan = nullptr;
    if (rms_plan_create(ctx, &job, 1, &rp, &plan) != RMS_OK) {
        std::string m = rms_last_error(ctx);
        rms_destroy(ctx);
        return die(m);
    }
    const int T = o.timepoints;
    const int nval = rms_nval(sp.Nc, job.nbvec);
    std::vector<double> rec((size_t)T * nval), sig(ngrad > 0 ? ngrad : 1);
    rms_job_out out;
    std::memset(&out, 0, sizeof(out));
    out.rec = rec.data();
    out.sig = ngrad > 0 ? sig.data() : nullptr;
    rms_timing tm;
    auto t0 = std::chrono::steady_clock::now();
    s = rms_plan_execute(plan, &out, &tm);
    auto t1 = std::chrono::steady_clock::now();
    // main.cu:501 prints numBlocks and blockSize of its launch
    std::cout << tm.grid << std::endl << tm.block << std::endl;
    if (s != RMS_OK) {
        std::string m = rms_last_error(ctx);
        rms_plan_destroy(plan);
        rms_destroy(ctx);
        return die(m);
    }
    rms_plan_destroy(plan);
    rms_destroy(ctx);
    const double secs = std::chrono::duration<double>(t1 - t0).count();
    std::cout << "Done! Elpased time " << double((long long)secs) << " s" << std::endl;

    // split the record rows into the reference files
    std::vector<double> sig0(T), dx2((size_t)T * 6), dx4((size_t)T * 15), sigRe((size_t)T * job.nbvec),
        npc((size_t)T * sp.Nc);
    for (int r = 0; r < T; r++) {
        const double *row = rec.data() + (size_t)r * nval;
        sig0[r] = row[0];
        for (int c = 0; c < sp.Nc; c++) npc[(size_t)r * sp.Nc + c] = row[1 + c];
        for (int k = 0; k < 6; k++) dx2[(size_t)r * 6 + k] = row[1 + sp.Nc + k];
        for (int k = 0; k < 15; k++) dx4[(size_t)r * 15 + k] = row[7 + sp.Nc + k];
        for (int k = 0; k < job.nbvec; k++) sigRe[(size_t)r * job.nbvec + k] = row[22 + sp.Nc + k];
    }
    bool ok = rmsio::write_column(path("sig0.txt"), sig0, 15, err);
    ok = ok && rmsio::write_matrix(path("dx2.txt"), dx2.data(), T, 6, 15, err);
    ok = ok && rmsio::write_matrix(path("dx4.txt"), dx4.data(), T, 15, 15, err);
    if (!pgse) ok = ok && rmsio::write_matrix(path("sigRe.txt"), sigRe.data(), T, job.nbvec, 15, err);
    ok = ok && rmsio::write_matrix(path("NPar_count.txt"), npc.data(), T, sp.Nc, 0, err);
    if (pgse) {
        sig.resize(ngrad);
        ok = ok && rmsio::write_column(path("sig.txt"), sig, 15, err);
    }
    ok = ok && rmsio::write_sim_para(path("sim_para.txt"), sp.dt, sp.TN, sp.NPar, sp.res, err);
    std::vector<double> TD(T);
    const int64_t Tstep = sp.TN / T;
    for (int i = 0; i < T; i++) TD[i] = (double)(i * Tstep + 1) * sp.dt;  // main.cu:484-487
    ok = ok && rmsio::write_column(path("diff_time.txt"), TD, 0, err);
    if (!ok) return die(err);
    return 0;
}

}  // namespace
