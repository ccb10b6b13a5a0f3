// rms_host.cpp -- job descriptors of the walk, built and checked on the host
// (declarations and conventions in rms_host.h).  No HIP calls, no heap.
#include "rms_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace rms {

// main.cu:429-458: T2/dt, step = sqrt(6 dt D)/res, P_ij = 0 next to class 0,
// 1 within a class, min(1, sqrt(D_j/D_i)) between classes; a permeability
// kappa_ij >= 0 scales the last by pi0/(1+pi0), pi0 = 4 kappa dt / (res
// min(s_i, s_j)).  Category 1 crosses (P > 0.999999), 0 reflects (P < 1e-6),
// 2 draws (main.cu:274-289).
const char *derive(const rms_geometry &g, double dt, Derived &d)
{
    if (g.nclass < 1 || g.nclass > NC) return "nclass must be in 1..4";
    for (int i = 0; i < NC; i++) {
        d.step[i] = d.inv_step[i] = 0.0;
        d.T2s[i] = INFINITY;
    }
    for (int i = 0; i < NC * NC; i++) {
        d.P[i] = 0.0;
        d.Pcat[i] = 0;
    }
    for (int i = 0; i < g.nclass; i++) {
        d.T2s[i] = g.T2[i] / dt;
        d.step[i] = std::sqrt(6.0 * dt * g.D[i]) / g.res;
        d.inv_step[i] = (d.step[i] > 0.0) ? 1.0 / d.step[i] : 0.0;
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
                const double kap = g.kappa[i * NC + j];
                if (kap >= 0.0) {
                    const double smin = std::fmin(d.step[i], d.step[j]);
                    const double pi0 = 4.0 * kap * dt / (g.res * smin);
                    p = p * (pi0 / (1.0 + pi0));
                }
            }
            d.P[i * NC + j] = p;
            d.Pcat[i * NC + j] = (p > 0.999999) ? 1 : ((p < 0.000001) ? 0 : 2);
        }
    }
    return nullptr;
}

// Constants of the tangent (adjoint) pass: parameter p < 3 is D of class
// p+1, p = 3 is kappa.  For the probabilistic pairs
// P = min(1, sqrt(Dj/Di)) * pi0/(1+pi0) with pi0 = 4 kappa dt / (res min(s_i, s_j));
// d log P and d log(1-P) = -P/(1-P) d log P.
void derive_tangent(const rms_geometry &g, double dt, const Derived &d, JobDesc &D)
{
    for (int c = 0; c < NC; c++) {
        D.dstep_d[c] = D.dinv_d[c] = D.rT[c] = D.invT2[c] = 0.0;
        D.dstep_f[c] = D.dinv_f[c] = 0.0f;
    }
    for (int i = 0; i < NC * NC * 4; i++) D.dlogP[i] = D.dlog1mP[i] = 0.0;
    for (int c = 0; c < g.nclass; c++) {
        const double ds = (g.D[c] > 0.0) ? d.step[c] / (2.0 * g.D[c]) : 0.0;
        const double di = (g.D[c] > 0.0) ? -d.inv_step[c] / (2.0 * g.D[c]) : 0.0;
        D.dstep_d[c] = ds;
        D.dinv_d[c] = di;
        D.dstep_f[c] = (float)ds;
        D.dinv_f[c] = (float)di;
        D.rT[c] = 1.0 / d.T2s[c];
        D.invT2[c] = 1.0 / g.T2[c];
    }
    for (int i = 1; i < g.nclass; i++) {
        for (int j = 1; j < g.nclass; j++) {
            const int ij = i * NC + j;
            if (d.Pcat[ij] != 2) continue;
            double dl[3] = {0.0, 0.0, 0.0};
            double dk = 0.0;
            const double ratio = std::sqrt(g.D[j] / g.D[i]);
            if (ratio < 1.0) {
                dl[j - 1] = dl[j - 1] + 0.5 / g.D[j];
                dl[i - 1] = dl[i - 1] - 0.5 / g.D[i];
            }
            const double kap = g.kappa[ij];
            if (kap >= 0.0) {
                const double smin = std::fmin(d.step[i], d.step[j]);
                const double pi0 = 4.0 * kap * dt / (g.res * smin);
                const int m = (d.step[i] <= d.step[j]) ? i : j;
                dl[m - 1] = dl[m - 1] - 0.5 / (g.D[m] * (1.0 + pi0));
                dk = (kap > 0.0) ? 1.0 / (kap * (1.0 + pi0)) : 0.0;
            }
            const double r = -d.P[ij] / (1.0 - d.P[ij]);
            for (int q = 0; q < 3; q++) {
                D.dlogP[ij * 4 + q] = dl[q];
                D.dlog1mP[ij * 4 + q] = r * dl[q];
            }
            D.dlogP[ij * 4 + 3] = dk;
            D.dlog1mP[ij * 4 + 3] = r * dk;
        }
    }
}

bool has_exchange(const rms_geometry &g, const Derived &d)
{
    for (int a = 1; a < g.nclass; a++)
        for (int b = 1; b < g.nclass; b++)
            if (a != b && d.Pcat[a * NC + b] != 0) return true;
    return false;
}

uint32_t geometry_flags(const rms_geometry &g)
{
    const int32_t N[3] = {g.nx, g.ny, g.nz};
    auto cls_at = [&](const int32_t q[3]) -> int32_t {
        const int64_t k = ((int64_t)q[0] * g.ny + q[1]) * g.nz + q[2];
        const uint32_t lab = g.label_bytes == 1 ? (uint32_t)((const uint8_t *)g.labels)[k]
                                                : (uint32_t)((const uint16_t *)g.labels)[k];
        return g.label_class ? (int32_t)g.label_class[lab] : (int32_t)lab;
    };
    // 1 when the voxel at p + d (wrapped on periodic axes) has class c
    auto is = [&](const int32_t p[3], const int32_t d[3], int32_t c) -> int {
        int32_t q[3];
        for (int k = 0; k < 3; k++) {
            q[k] = p[k] + d[k];
            if (q[k] >= 0 && q[k] < N[k]) continue;
            if (!((g.periodic >> k) & 1)) return 0;
            q[k] = q[k] < 0 ? q[k] + N[k] : q[k] - N[k];
        }
        return cls_at(q) == c ? 1 : 0;
    };
    static const int PA[3] = {0, 0, 1}, PB[3] = {1, 2, 2};
    uint32_t f = 0;
    int32_t p[3];
    for (p[0] = 0; p[0] < N[0]; p[0]++)
        for (p[1] = 0; p[1] < N[1]; p[1]++)
            for (p[2] = 0; p[2] < N[2]; p[2]++) {
                const int32_t c0 = cls_at(p);
                if (c0 == 0) continue;
                for (int k = 0; k < 3; k++)
                    if (!((g.periodic >> k) & 1) && (p[k] == 0 || p[k] == N[k] - 1)) f |= GEOM_MIRROR;
                if (!(f & GEOM_EDGE_ORDER))
                    for (int pr = 0; pr < 3; pr++)
                        for (int qd = 0; qd < 4; qd++) {
                            int32_t da[3] = {0, 0, 0}, db[3] = {0, 0, 0}, dab[3] = {0, 0, 0};
                            da[PA[pr]] = dab[PA[pr]] = (qd & 1) ? 1 : -1;
                            db[PB[pr]] = dab[PB[pr]] = (qd & 2) ? 1 : -1;
                            if (is(p, da, c0) + is(p, db, c0) + is(p, dab, c0) == 2) f |= GEOM_EDGE_ORDER;
                        }
                if (f == (GEOM_EDGE_ORDER | GEOM_MIRROR)) return f;
            }
    return f;
}

bool wrap_cells_rule(const rms_geometry &g)
{
    if ((g.periodic & 7) == 0) return false;
    int64_t walk = 0, side = 0;
    for (int32_t x = 0; x < g.nx; x++)
        for (int32_t y = 0; y < g.ny; y++)
            for (int32_t z = 0; z < g.nz; z++) {
                const int64_t k = ((int64_t)x * g.ny + y) * g.nz + z;
                const uint32_t lab = g.label_bytes == 1 ? (uint32_t)((const uint8_t *)g.labels)[k]
                                                        : (uint32_t)((const uint16_t *)g.labels)[k];
                const int32_t c = g.label_class ? (int32_t)g.label_class[lab] : (int32_t)lab;
                if (c == 0) continue;
                walk++;
                const bool on = (((g.periodic >> 0) & 1) && (x == 0 || x == g.nx - 1)) ||
                                (((g.periodic >> 1) & 1) && (y == 0 || y == g.ny - 1)) ||
                                (((g.periodic >> 2) & 1) && (z == 0 || z == g.nz - 1));
                side += on ? 1 : 0;
            }
    return 256 * side > walk;
}

bool lean_walk(const rms_geometry &g, const Derived &d)
{
    if (!(d.T2s[0] > 0.0)) return false;
    for (int c = 1; c < g.nclass; c++)
        if (!std::isinf(d.T2s[c])) return false;
    for (int a = 1; a < g.nclass; a++)
        for (int b = 1; b < g.nclass; b++)
            if (a != b && d.Pcat[a * NC + b] == 2) return false;
    return true;
}

bool fd_walker_mask(const rms_geometry &g, double dt, int c, double h, uint32_t &mask)
{
    mask = 1u << c;
    rms_geometry gp = g, gm = g;
    gp.D[c] = g.D[c] * (1.0 + h);  // the points of the central differences
    gm.D[c] = g.D[c] * (1.0 - h);
    Derived d0, dp, dm;
    if (derive(g, dt, d0) || derive(gp, dt, dp) || derive(gm, dt, dm)) return false;
    for (int a = 1; a < g.nclass; a++)
        for (int b = 1; b < g.nclass; b++) {
            const int ab = a * NC + b;
            if (a == b || (dp.P[ab] == d0.P[ab] && dm.P[ab] == d0.P[ab])) continue;
            if (d0.Pcat[ab] != 2 || dp.Pcat[ab] != 2 || dm.Pcat[ab] != 2) return false;
            mask |= 1u << (4 + ab);
        }
    return true;
}

bool kappa_closes_pair(const rms_geometry &g, const Derived &d)
{
    for (int a = 1; a < g.nclass; a++)
        for (int b = 1; b < g.nclass; b++) {
            if (a == b || g.kappa[a * NC + b] < 0.0 || d.Pcat[a * NC + b] != 0) continue;
            // the reference rule min(1, sqrt(Db/Da)) is positive for D > 0
            if (g.D[a] > 0.0 && g.D[b] > 0.0) return true;
        }
    return false;
}

// Threshold on the high half xhi of an FP32 step's first exchange draw: the
// draw's uniform is at least u01_f32(xhi << 16) = (256 xhi + 1) 2^-24, so
// xhi >= the smallest xhi with that bound >= P reflects whatever the low half.
// -1: the face reflects without a draw (P = 0); 65536: never decided by xhi
// alone (P = 1 crosses, and with more than two non-null classes the relation
// word does not name the other class).
int32_t exchange_threshold(int32_t nclass, int cat, float P)
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

// The PGSE sums keep sum(g) of the steps since the last flush in int32; with
// voxel indices below max(nx, ny, nz) a flush every 2^k steps, k the largest
// value <= 16 with 2^k max(nx, ny, nz) < 2^31, keeps it in range.
int32_t pgse_flush_log2(int32_t nx, int32_t ny, int32_t nz)
{
    const int64_t m = std::max<int64_t>(1, std::max<int64_t>(nx, std::max<int64_t>(ny, nz)));
    int32_t k = 16;
    while (k > 0 && (m << k) > 2147483647LL) k--;
    return k;
}

const char *check_params(const rms_run_params &P)
{
    if (!(P.dt > 0.0)) return "dt must be positive";
    if (P.timepoints < 1) return "timepoints must be >= 1";
    if (P.TN < P.timepoints) return "TN must be >= timepoints (the reference divides by TN/timepoints, main.cu:68,333)";
    if (P.TN > 2147483647LL) return "TN must be < 2^31";
    if (P.precision != RMS_PREC_F32 && P.precision != RMS_PREC_F64)
        return "precision must be RMS_PREC_F32 or RMS_PREC_F64";
    if (P.replicas > 4096) return "replicas must be <= 4096";
    return nullptr;
}

const char *fill_job(const rms_geometry &g, const rms_job &J, const rms_run_params &P, JobDesc &D, int &sig_bits,
                     bool &exchange)
{
    std::memset(&D, 0, sizeof(D));
    sig_bits = 0;
    exchange = false;
    if (J.nwalk < 1) return "nwalk must be >= 1";
    if (J.nbvec < 0 || (J.nbvec > 0 && (!J.bval || !J.bvec))) return "bad b-table";
    if (J.ngrad < 0 || (J.ngrad > 0 && !J.grad)) return "bad gradient table";
    for (int b = 0; b < J.nbvec; b++)
        if (!(J.bval[b] >= 0.0)) return "b-values must be >= 0";
    const int64_t R = P.replicas < 1 ? 1 : P.replicas;
    if (J.walker_base < 0 || J.walker_base % (WAVE * R) != 0)
        return "walker_base must be a non-negative multiple of 64 * replicas";
    if (J.walker_base > (int64_t)1 << 62 || J.nwalk > ((int64_t)1 << 62) - J.walker_base)
        return "walker ids must stay below 2^62";
    D.nx = g.nx;
    D.ny = g.ny;
    D.nz = g.nz;
    D.pg_log2 = pgse_flush_log2(g.nx, g.ny, g.nz);
    D.periodic = g.periodic;
    D.nbvec = J.nbvec;
    D.ngrad = J.ngrad;
    D.nwalk = J.nwalk;
    D.wid0 = J.walker_base;
    D.key0 = (uint32_t)J.seed;
    D.key1 = (uint32_t)(J.seed >> 32);
    D.res = g.res;
    return job_physics(g, P.dt, D, sig_bits, exchange);
}

const char *job_physics(const rms_geometry &g, double dt, JobDesc &D, int &sig_bits, bool &exchange)
{
    sig_bits = 0;
    exchange = false;
    Derived d;
    const char *e = derive(g, dt, d);
    if (e) return e;
    for (int c = 1; c < g.nclass; c++) {
        if (!(g.D[c] > 0.0)) return "non-null classes need a positive diffusivity";
        if (!(d.step[c] <= STEP_MAX))
            return "step size sqrt(6*D*dt) must be below one voxel by at least 2^-20 (main.cu:210-214 assumes "
                   "less than one)";
        if (!(g.T2[c] > 0.0)) return "T2 must be positive";
    }
    D.nclass = g.nclass;
    uint32_t pcat = 0;
    for (int i = 0; i < NC * NC; i++) pcat |= (uint32_t)(d.Pcat[i] & 3) << (2 * i);
    D.pcat = pcat;
    for (int c = 0; c < NC; c++) {
        D.step_f[c] = (float)d.step[c];
        D.inv_step_f[c] = (float)d.inv_step[c];
        D.step_d[c] = d.step[c];
        D.inv_step_d[c] = d.inv_step[c];
        D.T2s[c] = d.T2s[c];
    }
    for (int i = 0; i < NC * NC; i++) {
        D.P_f[i] = (float)d.P[i];
        D.P_d[i] = d.P[i];
    }
    D.xh12 = exchange_threshold(g.nclass, d.Pcat[1 * NC + 2], D.P_f[1 * NC + 2]);
    D.xh21 = exchange_threshold(g.nclass, d.Pcat[2 * NC + 1], D.P_f[2 * NC + 1]);
    derive_tangent(g, dt, d, D);
    exchange = has_exchange(g, d);
    // the walk kernel with faces between the two non-null classes on the fast
    // path (reflect at P = 0, draw at 0 < P < 1); it is also the only variant
    // whose march compiles exchange draws (category 2)
    bool draws = false;
    for (int i = 0; i < NC * NC; i++) draws = draws || d.Pcat[i] == 2;
    if (g.nclass == 3 || draws) sig_bits |= SIG_EXCH;
    // the walk kernel whose march wraps periodic lattice edges
    if (g.periodic != 0) sig_bits |= SIG_PERIODIC;
    return nullptr;
}

static int64_t first_false_lt(double dt, int64_t TN, double thr)
{
    int64_t lo = 0, hi = TN;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
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
        const int64_t mid = lo + (hi - lo) / 2;
        if ((double)(mid + 1) * dt >= thr)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// Insert v into the sorted distinct list b[1..n] (b[0] = 0 is kept apart);
// false when the list would exceed MAX_BOUNDS entries.
static bool insert_bound(int64_t *b, int32_t &n, int64_t v)
{
    if (v == 0) return true;
    int32_t k = 1;
    while (k <= n && b[k] < v) k++;
    if (k <= n && b[k] == v) return true;
    if (n >= MAX_BOUNDS) return false;
    for (int32_t m = n; m >= k; m--) b[m + 1] = b[m];
    b[k] = v;
    n++;
    return true;
}

static int32_t bound_index(const int64_t *b, int32_t n, int64_t v)
{
    if (v == 0) return 0;
    int32_t lo = 1, hi = n;
    while (lo < hi) {
        const int32_t mid = (lo + hi) / 2;
        if (b[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

static void lobes(const double *gr, double dt, int64_t TN, int64_t &a, int64_t &b, int64_t &e)
{
    // lobe 1: (i+1) dt < delta; lobe 2: Delta <= (i+1) dt < Delta + delta
    a = first_false_lt(dt, TN, gr[1]);
    b = first_true_ge(dt, TN, gr[0]);
    e = first_false_lt(dt, TN, gr[0] + gr[1]);
    if (b < a) b = a;
    if (e < b) e = b;
}

const char *pgse_tables(const double *grad, int32_t ngrad, double dt, int64_t TN, int32_t *perm, int32_t *triple,
                        int32_t max_triple, int64_t *bounds, int32_t &nbound, int32_t &ntriple, uint64_t *keys)
{
    nbound = 0;
    ntriple = 0;
    bounds[0] = 0;
    if (!insert_bound(bounds, nbound, TN)) return "too many distinct PGSE lobe boundaries (max 64)";
    for (int32_t g = 0; g < ngrad; g++) {
        int64_t a, b, e;
        lobes(grad + 6 * g, dt, TN, a, b, e);
        if (!insert_bound(bounds, nbound, a) || !insert_bound(bounds, nbound, b) || !insert_bound(bounds, nbound, e))
            return "too many distinct PGSE lobe boundaries (max 64)";
    }
    // sort by (ka, kb, ke, original index): indices < 2^7, gradients < 2^32
    for (int32_t g = 0; g < ngrad; g++) {
        int64_t a, b, e;
        lobes(grad + 6 * g, dt, TN, a, b, e);
        const uint64_t ka = (uint64_t)bound_index(bounds, nbound, a);
        const uint64_t kb = (uint64_t)bound_index(bounds, nbound, b);
        const uint64_t ke = (uint64_t)bound_index(bounds, nbound, e);
        keys[g] = (ka << 46) | (kb << 39) | (ke << 32) | (uint64_t)(uint32_t)g;
    }
    std::sort(keys, keys + ngrad);
    uint64_t prev = ~0ull;
    for (int32_t g = 0; g < ngrad; g++) {
        perm[g] = (int32_t)(uint32_t)keys[g];
        const uint64_t t = keys[g] >> 32;
        if (t != prev) {
            if (ntriple >= max_triple) return "too many distinct PGSE lobe triples";
            triple[4 * ntriple + 0] = (int32_t)((t >> 14) & 127u);
            triple[4 * ntriple + 1] = (int32_t)((t >> 7) & 127u);
            triple[4 * ntriple + 2] = (int32_t)(t & 127u);
            triple[4 * ntriple + 3] = g;
            ntriple++;
            prev = t;
        }
    }
    triple[4 * ntriple + 0] = 0;
    triple[4 * ntriple + 1] = 0;
    triple[4 * ntriple + 2] = 0;
    triple[4 * ntriple + 3] = ngrad;  // sentinel
    return nullptr;
}

const char *check_desc(const JobDesc &D, int sig)
{
    if (!D.cells || !D.cls) return "descriptor without a lattice";
    if (!D.part || !D.counters) return "descriptor without output buffers";
    if (D.nclass < 1 || D.nclass > NC) return "descriptor class count out of range";
    if (D.nwalk < 1) return "descriptor without walkers";
    if (sig & SIG_NARROW) {
        // qscale + r * nbvec and bvec are formed for every job of the plan
        if (!D.qscale || !D.bvec) return "narrow-pulse plan: job without a q table (or its sentinel)";
        if (D.nbvec < 0) return "negative b-table size";
    } else if (D.nbvec != 0) {
        return "job with a b-table in a plan without the narrow-pulse variant";
    }
    if (sig & SIG_PGSE) {
        // bounds[kb] for kb <= nbound and bounds[kb-1], the snapshot buffer at
        // each boundary, the triple table up to its sentinel, and the signal
        // partials are read by every job of the plan
        if (!D.bounds || !D.triple || !D.snap || !D.grad || !D.gperm || !D.psig)
            return "PGSE plan: job without lobe tables (or their sentinels)";
        if (D.nbound < 0 || D.nbound > MAX_BOUNDS) return "PGSE boundary count out of range";
        if (D.ntriple < 0 || D.ntriple > (D.ngrad > 0 ? D.ngrad : 0)) return "PGSE triple count out of range";
        if (D.ngrad == 0 && (D.nbound != 0 || D.ntriple != 0)) return "PGSE sentinel with boundaries";
    } else if (D.ngrad != 0) {
        return "job with a gradient table in a plan without the PGSE variant";
    }
    if ((sig & SIG_PERIODIC) == 0 && D.periodic != 0) return "periodic job in a plan without the periodic variant";
    if ((sig & SIG_BRICKS) == 0 && D.btab != nullptr) return "brick-stored job in a plan without the brick variant";
    // wrapped cell words let a straight move leave the lattice: only the
    // SIG_WRAP variants bring the walker back in
    if ((sig & SIG_WRAP) == 0 && D.wrapcells != 0) return "job with wrapped cell words in a plan without the wrap variant";
    if (D.wrapcells != 0 && D.periodic == 0) return "wrapped cell words without a periodic axis";
    if (D.btab != nullptr && (D.nby < 1 || D.nbz < 1)) return "brick-stored job without its brick grid";
    if ((D.periodic & ~7) != 0) return "periodic: bits 0..2 (x, y, z) only";
    return nullptr;
}

const char *scratch_layout(int64_t nbatch, int32_t timepoints, int32_t nval, int32_t replicas, int32_t ngrad,
                           int32_t nbound, int32_t ntriple, int32_t nbvec, int64_t cap, ScratchLayout &L)
{
    const char *too_big = "run needs more device memory than the RMS_SCRATCH_BYTES scratch array holds "
                          "(fewer walkers, records or replicas, or rms_run)";
    if (nbatch < 1 || timepoints < 1 || nval < 1 || replicas < 1 || ngrad < 0 || nbvec < 0) return "bad sizes";
    int64_t off = 0;
    bool over = false;
    // sizes as doubles first, so no product can overflow before the check
    auto seg = [&](double elems, double elem_bytes, double min_bytes) -> int64_t {
        double b = elems * elem_bytes;
        if (b < min_bytes) b = min_bytes;
        const int64_t o = off;
        if (over || (double)off + b + 256.0 > (double)cap) {
            over = true;
            return 0;
        }
        off += ((int64_t)b + 255) / 256 * 256;
        return o;
    };
    const double nb = (double)nbatch, T = (double)timepoints, nv = (double)nval, R = (double)replicas;
    const double ng = (double)ngrad, ws = nb * WAVE;
    L.job = seg(1, (double)sizeof(JobDesc), 0);
    L.red = seg(2, (double)sizeof(ReduceDesc), 0);
    L.batch_job = seg(nb, 4, 8);
    L.counters = seg(NCOUNT, 8, 0);
    L.part = seg(nb * T, nv * 8, 0);
    L.rep = seg(R * T, nv * 8, 0);
    L.tot = seg(T, nv * 8, 0);
    L.psig = seg(nb * ng, 8, 16);
    L.srep = seg(R * ng, 8, 16);
    L.stot = seg(ng, 8, 16);
    L.snap = seg(ngrad > 0 ? (double)(nbound + 1) * 3 * ws : 1, 8, 16);
    L.grad_raw = seg(6 * ng, 8, 48);
    L.grad = seg(6 * ng, 8, 48);
    L.perm = seg(ng, 4, 8);
    L.triple = seg(4.0 * (ntriple + 1), 4, 32);
    L.bounds = seg((double)nbound + 1, 8, 16);
    L.bval = seg((double)nbvec, 8, 16);
    L.qscale = seg(T * nbvec, 8, 16);
    L.bvec = seg(3.0 * nbvec, 8, 24);
    L.end = off;
    return over ? too_big : nullptr;
}

}  // namespace rms
