// Host-only checks of the job-descriptor builder (csrc/rms_host.cpp), built
// with g++ and no HIP.  Each check prints "ok <name>" or "FAIL <name>: why"
// and the exit status is the number of failures.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../monte-carlo-simulation-3d-rms_amd/csrc/rms_host.h"

using namespace rms;

static int nfail = 0;

static void check(bool ok, const char *name, const char *why)
{
    if (ok) {
        std::printf("ok %s\n", name);
    } else {
        std::printf("FAIL %s: %s\n", name, why);
        nfail++;
    }
}

static rms_geometry geom(int nclass, double D1, double D2, double kap)
{
    rms_geometry g;
    std::memset(&g, 0, sizeof(g));
    g.nx = g.ny = g.nz = 8;
    g.label_bytes = 1;
    g.nclass = nclass;
    g.res = 0.1;
    g.D[1] = D1;
    g.D[2] = D2;
    for (int c = 0; c < NC; c++) g.T2[c] = INFINITY;
    for (int i = 0; i < NC * NC; i++) g.kappa[i] = -1.0;
    if (kap >= 0.0) g.kappa[1 * NC + 2] = g.kappa[2 * NC + 1] = kap;
    return g;
}

static rms_run_params params(double dt, int64_t TN, int T, int R)
{
    rms_run_params p;
    std::memset(&p, 0, sizeof(p));
    p.dt = dt;
    p.TN = TN;
    p.timepoints = T;
    p.precision = RMS_PREC_F32;
    p.replicas = R;
    return p;
}

// Lobe tables against the literal per-step comparisons of main_PGSE.cu:341-345.
static void pgse_tables_vs_masks()
{
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    bool ok = true;
    char why[256] = "";
    for (int trial = 0; trial < 200 && ok; trial++) {
        const double dt = 2e-4 * (1 + trial % 3);
        const int64_t TN = 50 + (int64_t)(U(rng) * 400);
        const int ngrad = 1 + (int)(U(rng) * 40);
        std::vector<double> grad(6 * (size_t)ngrad);
        for (int g = 0; g < ngrad; g++) {
            // lobes on the step grid, off it, and past the end of the run
            const double delta = (trial % 2 ? std::round(U(rng) * 30) : U(rng) * 30) * dt;
            const double Delta = delta + (trial % 2 ? std::round(U(rng) * 200) : U(rng) * 200) * dt;
            grad[6 * g + 0] = Delta;
            grad[6 * g + 1] = delta;
            grad[6 * g + 2] = U(rng);
            grad[6 * g + 3] = U(rng);
            grad[6 * g + 4] = U(rng);
            grad[6 * g + 5] = U(rng);
        }
        std::vector<int32_t> perm(ngrad), trip(4 * (ngrad + 1));
        std::vector<uint64_t> keys(ngrad);
        int64_t bounds[MAX_BOUNDS + 1];
        int32_t nb = 0, nt = 0;
        const char *e = pgse_tables(grad.data(), ngrad, dt, TN, perm.data(), trip.data(), ngrad, bounds, nb, nt,
                                    keys.data());
        if (e) {
            // only too many distinct boundaries may be refused
            ok = std::strstr(e, "boundaries") != nullptr && 3 * ngrad + 1 > MAX_BOUNDS;
            if (!ok) std::snprintf(why, sizeof why, "unexpected error %s", e);
            continue;
        }
        // bounds: 0, then strictly increasing, ending at TN
        if (bounds[0] != 0 || bounds[nb] != TN) {
            ok = false;
            std::snprintf(why, sizeof why, "bounds ends");
        }
        for (int k = 1; k < nb && ok; k++)
            if (!(bounds[k] < bounds[k + 1])) ok = false;
        // perm is a permutation; triples cover it in order, sentinel last
        std::vector<int> seen(ngrad, 0);
        for (int g = 0; g < ngrad && ok; g++) {
            if (perm[g] < 0 || perm[g] >= ngrad || seen[perm[g]]++) ok = false;
        }
        if (trip[4 * 0 + 3] != 0 || trip[4 * nt + 3] != ngrad) ok = false;
        for (int t = 0; t < nt && ok; t++) {
            const int ka = trip[4 * t], kb = trip[4 * t + 1], ke = trip[4 * t + 2];
            for (int s = trip[4 * t + 3]; s < trip[4 * t + 7] && ok; s++) {
                const double *gr = grad.data() + 6 * perm[s];
                for (int64_t i = 0; i < TN; i++) {
                    const double tdt = (double)(i + 1) * dt;
                    const bool l1 = tdt < gr[1];
                    const bool l2 = !l1 && (tdt >= gr[0]) && (tdt < gr[0] + gr[1]);
                    const bool w1 = i < bounds[ka];
                    const bool w2 = i >= bounds[kb] && i < bounds[ke];
                    if (l1 != w1 || l2 != w2) {
                        ok = false;
                        std::snprintf(why, sizeof why, "trial %d gradient %d step %lld: mask (%d,%d) table (%d,%d)",
                                      trial, perm[s], (long long)i, l1, l2, w1, w2);
                        break;
                    }
                }
            }
            // sorted by triple, then by original index
            if (t > 0) {
                const int *a = &trip[4 * (t - 1)], *b = &trip[4 * t];
                const bool inc = a[0] < b[0] || (a[0] == b[0] && (a[1] < b[1] || (a[1] == b[1] && a[2] < b[2])));
                if (!inc) ok = false;
            }
            for (int s = trip[4 * t + 3] + 1; s < trip[4 * t + 7]; s++)
                if (perm[s] <= perm[s - 1]) ok = false;
        }
    }
    check(ok, "pgse_tables_vs_masks", why);
}

// Descriptor checks: sentinels pass, a missing table of the chosen variant fails.
static void descriptor_tables()
{
    rms_geometry g = geom(2, 2.0, 0.0, -1.0);
    rms_job J;
    std::memset(&J, 0, sizeof(J));
    J.nwalk = 100;
    rms_run_params P = params(2e-4, 100, 10, 4);
    JobDesc D;
    int bits = 0;
    bool exch = false;
    check(fill_job(g, J, P, D, bits, exch) == nullptr && !exch && bits == 0, "fill_job_plain", "plain job refused");
    static uint64_t cells[1];
    static uint8_t cls[1];
    static double part[1], qs[2], bv[3], gr[6], ps[2], snap[3];
    static unsigned long long cnt[NCOUNT];
    static int64_t bounds[2] = {0, 0};
    static int32_t trip[8] = {0, 0, 0, 0, 0, 0, 0, 0}, gp[2] = {0, 0};
    D.cells = cells;
    D.cls = cls;
    D.part = part;
    D.counters = cnt;
    check(check_desc(D, 0) == nullptr, "desc_plain_ok", "plain descriptor refused");
    check(check_desc(D, SIG_PGSE) != nullptr, "desc_pgse_missing", "PGSE variant without lobe tables accepted");
    check(check_desc(D, SIG_NARROW) != nullptr, "desc_narrow_missing", "narrow variant without q table accepted");
    D.bounds = bounds;
    D.triple = trip;
    D.gperm = gp;
    D.grad = gr;
    D.psig = ps;
    D.snap = snap;
    D.qscale = qs;
    D.bvec = bv;
    check(check_desc(D, SIG_PGSE | SIG_NARROW) == nullptr, "desc_sentinels_ok", "sentinel tables refused");
    D.ngrad = 3;
    check(check_desc(D, SIG_NARROW) != nullptr, "desc_grad_wrong_variant", "PGSE job in a narrow plan accepted");
    D.ngrad = 0;
    D.nbound = 1;
    check(check_desc(D, SIG_PGSE) != nullptr, "desc_sentinel_bounds", "sentinel with boundaries accepted");
    D.nbound = 0;
    D.periodic = 1;
    check(check_desc(D, 0) != nullptr, "desc_periodic_variant", "periodic job in a mirror plan accepted");
}

// Steps must round below one voxel in FP32: s <= 1 - 2^-20 (as a double).
static void step_margin()
{
    const double dt = 2e-4, res = 0.1;
    rms_run_params P = params(dt, 100, 10, 1);
    rms_job J;
    std::memset(&J, 0, sizeof(J));
    J.nwalk = 1;
    auto D_for = [&](double s) { return s * s * res * res / (6.0 * dt); };
    JobDesc D;
    int bits;
    bool exch;
    rms_geometry g = geom(2, D_for(1.0 - 0x1p-22), 0.0, -1.0);
    check(fill_job(g, J, P, D, bits, exch) != nullptr, "step_margin_refused", "step 1-2^-22 accepted");
    g = geom(2, D_for(1.0 - 0x1p-18), 0.0, -1.0);
    check(fill_job(g, J, P, D, bits, exch) == nullptr, "step_margin_ok", "step 1-2^-18 refused");
    // the largest accepted step keeps fma(s, v, o) below 2 in FP32
    const float s = (float)STEP_MAX;
    const float e = std::fma(s, 1.0f, 1.0f);
    check(std::floor(e) == 1.0f, "step_margin_fp32", "fma(s, 1, 1) rounds to 2");
}

static void walker_base_rule()
{
    rms_geometry g = geom(2, 2.0, 0.0, -1.0);
    rms_run_params P = params(2e-4, 100, 10, 8);
    rms_job J;
    std::memset(&J, 0, sizeof(J));
    J.nwalk = 10;
    JobDesc D;
    int bits;
    bool exch;
    J.walker_base = 64 * 8 * 3;
    check(fill_job(g, J, P, D, bits, exch) == nullptr && D.wid0 == 64 * 8 * 3, "walker_base_ok", "aligned base");
    J.walker_base = 64 * 4;
    check(fill_job(g, J, P, D, bits, exch) != nullptr, "walker_base_unaligned", "base not a multiple of 64R");
    J.walker_base = -512;
    check(fill_job(g, J, P, D, bits, exch) != nullptr, "walker_base_negative", "negative base");
}

static void kappa_rules()
{
    Derived d;
    rms_geometry g = geom(3, 2.0, 1.5, 0.0);
    derive(g, 2e-4, d);
    check(kappa_closes_pair(g, d) && !has_exchange(g, d), "kappa_zero_closes", "kappa = 0 pair not flagged");
    g = geom(3, 2.0, 1.5, 0.01);
    derive(g, 2e-4, d);
    check(!kappa_closes_pair(g, d) && has_exchange(g, d), "kappa_open", "kappa = 0.01 pair flagged");
    g = geom(3, 2.0, 1.5, -1.0);
    derive(g, 2e-4, d);
    check(!kappa_closes_pair(g, d) && has_exchange(g, d) && d.Pcat[1 * NC + 2] == 2 && d.Pcat[2 * NC + 1] == 1,
          "kappa_reference_rule", "reference rule categories");
    g = geom(3, 2.0, 1.5, 1e-9);
    derive(g, 2e-4, d);
    check(kappa_closes_pair(g, d), "kappa_tiny_closes", "P < 1e-6 from a tiny kappa not flagged");
}

// The periodic-side rule: wrapped cell words when more than 1/256 of the
// non-null voxels lie on a periodic side.  A column of non-null voxels along
// z in a 3 x 3 x nz lattice, periodic in z only, has 2 side voxels out of nz:
// exactly 1/256 at nz = 512 (not wrapped), above it at nz = 511.
static void wrap_rule()
{
    auto column = [](int32_t nz, int32_t periodic, std::vector<uint8_t> &lab) {
        lab.assign((size_t)9 * nz, 0);
        for (int32_t z = 0; z < nz; z++) lab[(size_t)(1 * 3 + 1) * nz + z] = 1;
        rms_geometry g;
        std::memset(&g, 0, sizeof(g));
        g.nx = g.ny = 3;
        g.nz = nz;
        g.label_bytes = 1;
        g.labels = lab.data();
        g.nclass = 2;
        g.periodic = periodic;
        return g;
    };
    std::vector<uint8_t> lab;
    check(!wrap_cells_rule(column(512, 4, lab)), "wrap rule: 2 of 512 voxels on the sides is not above 1/256", "");
    check(wrap_cells_rule(column(511, 4, lab)), "wrap rule: 2 of 511 voxels on the sides is above 1/256", "");
    check(!wrap_cells_rule(column(16, 0, lab)), "wrap rule: no periodic axis", "");
    // the column touches no x or y side: periodic x, y only counts nothing
    check(!wrap_cells_rule(column(16, 3, lab)), "wrap rule: periodic axes the walkable voxels do not touch", "");
}

// Physics fields rebuilt at a perturbed D (the central differences of the
// adjoint): equal to a descriptor built from scratch with that D, and the
// variant bits they need follow the exchange categories (three non-null
// classes with equal D cross freely, P = 1, and need no draws; moving one D
// makes draws, which need the exchange variant).
static void perturbed_physics()
{
    rms_geometry g = geom(4, 1.5, 1.5, -1.0);
    g.D[3] = 1.5;
    rms_job J;
    std::memset(&J, 0, sizeof(J));
    J.nwalk = 64;
    const rms_run_params P = params(2e-4, 100, 10, 1);
    JobDesc D0, D1;
    int b0 = 0, b1 = 0;
    bool e0 = false, e1 = false;
    check(fill_job(g, J, P, D0, b0, e0) == nullptr && (b0 & SIG_EXCH) == 0 && e0,
          "equal D: walkers cross between classes without draws", "");
    rms_geometry gp = g;
    gp.D[2] = 1.5 * 1.05;
    JobDesc Dp = D0;
    check(job_physics(gp, P.dt, Dp, b1, e1) == nullptr && (b1 & SIG_EXCH) != 0, "moved D: draws need the exchange variant",
          "");
    check(fill_job(gp, J, P, D1, b0, e0) == nullptr && std::memcmp(D1.step_d, Dp.step_d, sizeof(Dp.step_d)) == 0 &&
              std::memcmp(D1.P_d, Dp.P_d, sizeof(Dp.P_d)) == 0 && D1.pcat == Dp.pcat &&
              std::memcmp(D1.dlogP, Dp.dlogP, sizeof(Dp.dlogP)) == 0 && D1.xh12 == Dp.xh12,
          "physics rebuilt at a perturbed D equal a fresh descriptor", "");
    gp.D[2] = 1e6;
    check(job_physics(gp, P.dt, Dp, b1, e1) != nullptr, "a perturbed step of a voxel or more is refused", "");
}

// Descriptors whose cell words only one kernel variant can walk.
static void storage_variants()
{
    JobDesc D;
    std::memset(&D, 0, sizeof(D));
    uint64_t cells = 0;
    uint8_t cls = 1;
    double part = 0.0;
    unsigned long long cnt = 0;
    D.cells = &cells;
    D.cls = &cls;
    D.part = &part;
    D.counters = &cnt;
    D.nclass = 2;
    D.nwalk = 1;
    check(check_desc(D, 0) == nullptr, "dense descriptor in the dense variant", "");
    uint32_t btab = 0;
    D.btab = &btab;
    D.nby = D.nbz = 1;
    check(check_desc(D, 0) != nullptr, "brick descriptor refused by the dense variant", "");
    check(check_desc(D, SIG_BRICKS) == nullptr, "brick descriptor in the brick variant", "");
    D.btab = nullptr;
    D.wrapcells = 1;
    D.periodic = 4;
    check(check_desc(D, SIG_PERIODIC) != nullptr, "wrapped cell words refused without the wrap variant", "");
    check(check_desc(D, SIG_PERIODIC | SIG_WRAP) == nullptr, "wrapped cell words in the wrap variant", "");
    D.periodic = 0;
    check(check_desc(D, SIG_PERIODIC | SIG_WRAP) != nullptr, "wrapped cell words without a periodic axis", "");
}

// Walkers of the central differences of D_c: the mask holds class c and the
// draws at faces whose P moves with D_c.  With D1 = 2 > D2 = 0.8 and a
// permeability, P_21 = min(1, sqrt(D1/D2)) pi0/(1 + pi0) with pi0 from the
// smaller step (class 2's) does not depend on D1, P_12 does (the square root),
// and both depend on D2 (pi0).  Without a permeability P_21 = 1 at every
// point.  With D1 = 1, D2 = 1.02 and no permeability, P_12 = 1 at the base
// point and below 1 at D_c (1 -+ h): no trustworthy mask.
static void fd_masks()
{
    const double dt = 2e-4, h = 0.05;
    const uint32_t b12 = 1u << (4 + 1 * NC + 2), b21 = 1u << (4 + 2 * NC + 1);
    uint32_t m = 0;
    rms_geometry g = geom(3, 2.0, 0.8, 0.2);
    check(fd_walker_mask(g, dt, 1, h, m) && m == (2u | b12), "fd mask: D1 moves P_12 only", "");
    check(fd_walker_mask(g, dt, 2, h, m) && m == (4u | b12 | b21), "fd mask: D2 moves both", "");
    g = geom(3, 2.0, 0.8, -1.0);
    check(fd_walker_mask(g, dt, 1, h, m) && m == (2u | b12), "fd mask: reference rule, D1", "");
    check(fd_walker_mask(g, dt, 2, h, m) && m == (4u | b12), "fd mask: reference rule, D2", "");
    g = geom(3, 1.0, 1.02, -1.0);
    check(!fd_walker_mask(g, dt, 1, h, m), "fd mask: P_12 leaves 1 at D1 (1 + h)", "");
    check(!fd_walker_mask(g, dt, 2, h, m), "fd mask: P_12 leaves 1 at D2 (1 - h)", "");
    g = geom(2, 2.0, 0.0, -1.0);
    check(fd_walker_mask(g, dt, 1, h, m) && m == 2u, "fd mask: one non-null class", "");
}

// The lean rule: infinite T2 in every non-null class and no drawing pair.
static void lean_rule()
{
    Derived d;
    rms_geometry g = geom(2, 2.0, 0.0, -1.0);
    derive(g, 2e-4, d);
    check(lean_walk(g, d), "lean: one class, T2 infinite", "");
    g.T2[1] = 80.0;
    derive(g, 2e-4, d);
    check(!lean_walk(g, d), "lean: a finite T2", "");
    g = geom(3, 1.5, 1.5, -1.0);  // P = 1 both ways: crossings without draws
    derive(g, 2e-4, d);
    check(lean_walk(g, d), "lean: two classes that cross without a draw", "");
    g = geom(3, 2.0, 1.5, -1.0);  // P_21 = 1, P_12 = sqrt(0.75): draws
    derive(g, 2e-4, d);
    check(!lean_walk(g, d), "lean: a drawing pair", "");
    g = geom(3, 2.0, 1.5, 0.0);  // kappa = 0: both pairs reflect
    derive(g, 2e-4, d);
    check(lean_walk(g, d), "lean: pairs closed by kappa = 0", "");
    g = geom(2, 2.0, 0.0, -1.0);
    g.T2[0] = 0.0;
    derive(g, 2e-4, d);
    check(!lean_walk(g, d), "lean: class 0 with T2 = 0 (0/0 in the weight)", "");
}

int main()
{
    pgse_tables_vs_masks();
    descriptor_tables();
    step_margin();
    walker_base_rule();
    kappa_rules();
    wrap_rule();
    perturbed_physics();
    storage_variants();
    fd_masks();
    lean_rule();
    std::printf("%d failures\n", nfail);
    return nfail;
}
