/*
 * rms_oracle.c -- CPU restatement of the RMS random-walk engine.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the CPU
 * baseline of bench.py.  Never linked into the product.
 *
 * Reference anchors (upstream repository paths):
 *   main.cu:40-56    atomAdd          -> replaced by a fixed reduction order:
 *                                         64-lane pairwise tree per batch, then
 *                                         batches summed per replica in order,
 *                                         then replicas summed in order.
 *   main.cu:58-61    setup_kernel     -> no RNG state; counter-based Philox.
 *   main.cu:63-384   propagate        -> rms_oracle_core.inc (+ records here)
 *   main.cu:429-458  derived params   -> or_derive
 *   main.cu:484-487  TD               -> or_diff_times
 *   main_PGSE.cu:336-357 PGSE phases  -> window sums (or_pgse_phase) and the
 *                                         literal loop (pgse_bruteforce=1).
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off -fopenmp).  The
 * floating-point operations are written in the exact order the HIP engine
 * uses; contraction must stay off for bit-identical results.
 */
#include "rms_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_INIT_TRIALS (1u << 26)

static __thread char g_err[512];
const char *or_last_error(void) { return g_err; }
static int fail(const char *msg)
{
    snprintf(g_err, sizeof(g_err), "%s", msg);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), as published.     */
/* ------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* Uniform in the open interval (0,1): (2k+1) * 2^-53 with a 52-bit k. */
double or_u01_f64(uint32_t a, uint32_t b)
{
    double hi = (double)(a >> 6);
    double lo = (double)(b >> 6);
    double k = hi * 67108864.0 + lo;
    return (2.0 * k + 1.0) * 1.1102230246251565e-16;
}

/* Uniform in (0,1): (2k+1) * 2^-24 with a 23-bit k. */
float or_u01_f32(uint32_t w)
{
    float k = (float)(w >> 9);
    return (2.0f * k + 1.0f) * 5.9604644775390625e-08f;
}

static inline void or_rng_block(const uint32_t key[2], uint32_t idx, uint32_t stream, uint64_t wid,
                                uint32_t out[4])
{
    uint32_t ctr[4] = {idx, stream, (uint32_t)wid, (uint32_t)(wid >> 32)};
    or_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------ */
/* Own elementary functions (Taylor polynomials, Cody-Waite reduction). */
/* The HIP engine evaluates the same sequences.                          */
/* ------------------------------------------------------------------ */
static inline void sc_poly_f64(double x, double *s, double *c)
{
    double z = x * x;
    double p = -7.647163731819816e-13;
    p = fma(p, z, 1.6059043836821613e-10);
    p = fma(p, z, -2.505210838544172e-08);
    p = fma(p, z, 2.7557319223985893e-06);
    p = fma(p, z, -0.0001984126984126984);
    p = fma(p, z, 0.008333333333333333);
    p = fma(p, z, -0.16666666666666666);
    *s = fma(x * z, p, x);
    double q = 4.779477332387385e-14;
    q = fma(q, z, -1.1470745597729725e-11);
    q = fma(q, z, 2.08767569878681e-09);
    q = fma(q, z, -2.755731922398589e-07);
    q = fma(q, z, 2.48015873015873e-05);
    q = fma(q, z, -0.001388888888888889);
    q = fma(q, z, 0.041666666666666664);
    q = fma(q, z, -0.5);
    *c = fma(z, q, 1.0);
}

static inline void sc_poly_f32(float x, float *s, float *c)
{
    float z = x * x;
    float p = 2.7557318844628753e-06f;
    p = fmaf(p, z, -0.00019841270113829523f);
    p = fmaf(p, z, 0.008333333767950535f);
    p = fmaf(p, z, -0.1666666716337204f);
    *s = fmaf(x * z, p, x);
    float q = -2.755731998149713e-07f;
    q = fmaf(q, z, 2.4801587642286904e-05f);
    q = fmaf(q, z, -0.0013888889225199819f);
    q = fmaf(q, z, 0.0416666679084301f);
    q = fmaf(q, z, -0.5f);
    *c = fmaf(z, q, 1.0f);
}

static inline void quadrant(int q, double s, double c, double *so, double *co)
{
    switch (q & 3) {
    case 0: *so = s; *co = c; break;
    case 1: *so = c; *co = -s; break;
    case 2: *so = -s; *co = -c; break;
    default: *so = -c; *co = s; break;
    }
}

/* sin(2*pi*t), cos(2*pi*t) for t in [0,1]. */
void or_sincos_turns_f64(double t, double *s, double *c)
{
    double k = floor(4.0 * t + 0.5);
    double r = t - k * 0.25;
    double x = r * 6.283185307179586;
    double sx, cx;
    sc_poly_f64(x, &sx, &cx);
    quadrant((int)k, sx, cx, s, c);
}

/* sin and cos of 2*pi*k/256, k = 0..255, as float32 roundings of the C
 * library's double sin/cos (the engine's table, rms_turn_table.h, holds the
 * same values). */
static float turn_sin[256], turn_cos[256];

__attribute__((constructor)) static void or_turn_table_init(void)
{
    for (int k = 0; k < 256; k++) {
        const double a = 2 * 3.141592653589793 * k / 256;
        turn_sin[k] = (float)sin(a);
        turn_cos[k] = (float)cos(a);
    }
}

void or_turn_table(float s[256], float c[256])
{
    for (int k = 0; k < 256; k++) {
        s[k] = turn_sin[k];
        c[k] = turn_cos[k];
    }
}

/* sin(2 pi t), cos(2 pi t) for t = m * 2^-24 (FP32 direction draw): k =
 * round(256 t), x = 2 pi (t - k/256), |x| <= pi/256, and the angle sum with
 * cos x = 1 - x^2/2, sin x = x - x^3/6. */
void or_sincos_turns_f32(float t, float *s, float *c)
{
    const uint32_t m = (uint32_t)(t * 16777216.0f);
    const uint32_t k = (m + (1u << 15)) >> 16;
    const int32_t num = (int32_t)m - (int32_t)(k << 16);
    const float x = (float)num * (6.2831854820251465f * 5.9604644775390625e-08f);
    const float as = turn_sin[k & 255u], ac = turn_cos[k & 255u];
    const float x2 = x * x;
    const float cr = fmaf(x2, -0.5f, 1.0f);
    const float sr = fmaf(x * x2, -0.16666667163372040f, x);
    *s = fmaf(as, cr, ac * sr);
    *c = fmaf(ac, cr, -(as * sr));
}

/* exp(x) for x <= 0 (T2 weights). */
double or_exp_f64(double x)
{
    if (!(x > -708.0)) return 0.0;
    double k = floor(x * 1.4426950408889634 + 0.5);
    double r = fma(-k, 0.6931471803691238, x);
    r = fma(-k, 1.9082149292705877e-10, r);
    double p = 1.6059043836821613e-10;
    p = fma(p, r, 2.08767569878681e-09);
    p = fma(p, r, 2.505210838544172e-08);
    p = fma(p, r, 2.755731922398589e-07);
    p = fma(p, r, 2.7557319223985893e-06);
    p = fma(p, r, 2.48015873015873e-05);
    p = fma(p, r, 0.0001984126984126984);
    p = fma(p, r, 0.001388888888888889);
    p = fma(p, r, 0.008333333333333333);
    p = fma(p, r, 0.041666666666666664);
    p = fma(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    int64_t e = (int64_t)k + 1023;
    uint64_t bits = (uint64_t)e << 52;
    double scale;
    memcpy(&scale, &bits, sizeof(scale));
    return p * scale;
}

/* sin and cos of an arbitrary argument (|x| up to ~1e6 at full accuracy). */
void or_sincos_f64(double x, double *s, double *c)
{
    double k = floor(x * 0.6366197723675814 + 0.5);
    double r = fma(-k, 1.5707963267341256, x);
    r = fma(-k, 6.077100506303966e-11, r);
    r = fma(-k, 2.0222662487959506e-21, r);
    double sx, cx;
    sc_poly_f64(r, &sx, &cx);
    quadrant((int)(((int64_t)k) & 3), sx, cx, s, c);
}

double or_cos_f64(double x)
{
    double s, c;
    or_sincos_f64(x, &s, &c);
    return c;
}

/* ------------------------------------------------------------------ */
/* Host derivations                                                      */
/* ------------------------------------------------------------------ */

/* main.cu:429-458.  T2 in steps, step size in voxels, permeation
 * probabilities.  kappa[i][j] < 0 selects the reference rule
 * min(1, sqrt(Dj/Di)); kappa >= 0 scales it by the symmetric elastic
 * transmission pi0/(1+pi0), pi0 = 4*kappa*dt/(res*min(s_i, s_j)), which
 * keeps s_i*P_ij = s_j*P_ji (detailed balance) and tends to the reference
 * rule as kappa -> inf.  The finite-step correction pi0/(1+pi0) is the one
 * analysis/interaction.m:438-439 applies to rejection sampling. */
int or_derive(const or_geom *g, double dt, or_derived *d)
{
    memset(d, 0, sizeof(*d));
    if (g->nclass < 1 || g->nclass > OR_NC_MAX) return fail("nclass out of range");
    d->nclass = g->nclass;
    for (int i = 0; i < g->nclass; i++) {
        d->T2s[i] = g->T2[i] / dt;
        d->step[i] = sqrt(6.0 * dt * g->D[i]) / g->res;
        d->inv_step[i] = (d->step[i] > 0.0) ? 1.0 / d->step[i] : 0.0;
    }
    for (int i = 0; i < g->nclass; i++) {
        for (int j = 0; j < g->nclass; j++) {
            double P;
            if ((i == 0) || (j == 0)) {
                P = 0.0;
            } else if (i == j) {
                P = 1.0;
            } else {
                P = fmin(1.0, sqrt(g->D[j] / g->D[i]));
                double kap = g->kappa[i * OR_NC_MAX + j];
                if (kap >= 0.0) {
                    double smin = fmin(d->step[i], d->step[j]);
                    double pi0 = 4.0 * kap * dt / (g->res * smin);
                    P = P * (pi0 / (1.0 + pi0));
                }
            }
            d->P[i * OR_NC_MAX + j] = P;
            d->Pcat[i * OR_NC_MAX + j] = (P > 0.999999) ? 1 : ((P < 0.000001) ? 0 : 2);
        }
    }
    return 1;
}

int32_t or_nval(int32_t nclass, int32_t nbvec) { return 1 + nclass + 6 + 15 + nbvec; }

/* main.cu:484-487 */
void or_diff_times(double dt, int64_t TN, int32_t timepoints, double *TD)
{
    int64_t Tstep = TN / timepoints;
    for (int32_t i = 0; i < timepoints; i++) TD[i] = (double)(i * Tstep + 1) * dt;
}

/* First step index i in [lo, TN] for which pred(i) is false, where pred is
 * monotone (true then false). */
static int64_t first_false_lt(double dt, int64_t TN, double thr)
{
    /* pred(i) = (i+1)*dt < thr */
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

/* main_PGSE.cu:341-345: lobe 1 = {i : (i+1)dt < delta}; lobe 2 = {i not in
 * lobe 1 : (i+1)dt >= Delta and (i+1)dt < Delta+delta}. */
void or_pgse_lobes(double dt, int64_t TN, int32_t ngrad, const double *grad, int64_t *abe)
{
    for (int32_t j = 0; j < ngrad; j++) {
        const double *gr = grad + 6 * j;
        int64_t a = first_false_lt(dt, TN, gr[1]);
        int64_t b = first_true_ge(dt, TN, gr[0]);
        int64_t e = first_false_lt(dt, TN, gr[0] + gr[1]);
        if (b < a) b = a;
        if (e < b) e = b;
        abe[3 * j + 0] = a;
        abe[3 * j + 1] = b;
        abe[3 * j + 2] = e;
    }
}

static inline int32_t or_class_at(const or_geom *G, int32_t x, int32_t y, int32_t z)
{
    int64_t idx = ((int64_t)x * G->ny + y) * G->nz + z;
    uint32_t lab = (G->label_bytes == 1) ? ((const uint8_t *)G->labels)[idx]
                                         : ((const uint16_t *)G->labels)[idx];
    return G->label_class ? (int32_t)G->label_class[lab] : (int32_t)lab;
}

/* Pairwise tree over 64 lanes (lane 0 of a wavefront butterfly). */
static inline double or_tree64(double *a)
{
    for (int w = 1; w < OR_BATCH; w *= 2)
        for (int l = 0; l < OR_BATCH; l += 2 * w) a[l] = a[l] + a[l + w];
    return a[0];
}

/* Record values of one walker, main.cu:341-378; column l of vals (stride 64):
 * [sig0, count[nclass], dx2[6], dx4[15], sigRe[nbvec]]. */
static inline void or_record_values(double s0, const double dx[3], int32_t c, int32_t nclass, int32_t nbvec,
                                    const double *bvec, const double *qs, double *vals, int l)
{
    double x = dx[0], y = dx[1], z = dx[2];
    int q = 0;
#define PUT(v) (vals[(q++) * OR_BATCH + l] = (v))
    PUT(s0);
    for (int cc = 0; cc < nclass; cc++) PUT((cc == c) ? 1.0 : 0.0);
    PUT(s0 * x * x);
    PUT(s0 * x * y);
    PUT(s0 * x * z);
    PUT(s0 * y * y);
    PUT(s0 * y * z);
    PUT(s0 * z * z);
    PUT(s0 * x * x * x * x);
    PUT(s0 * x * x * x * y);
    PUT(s0 * x * x * x * z);
    PUT(s0 * x * x * y * y);
    PUT(s0 * x * x * y * z);
    PUT(s0 * x * x * z * z);
    PUT(s0 * x * y * y * y);
    PUT(s0 * x * y * y * z);
    PUT(s0 * x * y * z * z);
    PUT(s0 * x * z * z * z);
    PUT(s0 * y * y * y * y);
    PUT(s0 * y * y * y * z);
    PUT(s0 * y * y * z * z);
    PUT(s0 * y * z * z * z);
    PUT(s0 * z * z * z * z);
    for (int j = 0; j < nbvec; j++) {
        double qx = qs[j] * (x * bvec[3 * j + 0] + y * bvec[3 * j + 1] + z * bvec[3 * j + 2]);
        PUT(s0 * or_cos_f64(qx));
    }
#undef PUT
}

/* Constants of the tangent (adjoint) pass, derived on the host from the
 * same quantities as or_derive.  Parameter p < 3 is D of class p+1. */
typedef struct {
    double dstep[OR_NC_MAX]; /* d step_c / d D_c = step_c / (2 D_c) */
    double dinv[OR_NC_MAX];  /* d (1/step_c) / d D_c */
    double rT[OR_NC_MAX];    /* 1 / T2s_c: rate of the T2 exponent per step in c */
    double invT2[OR_NC_MAX]; /* 1 / T2_c (ms) */
    double dlogP[OR_NC_MAX * OR_NC_MAX][3];
    double dlog1mP[OR_NC_MAX * OR_NC_MAX][3];
    double dlogPk[OR_NC_MAX * OR_NC_MAX];
    double dlog1mPk[OR_NC_MAX * OR_NC_MAX];
} or_tderiv;

/* d log P_ij and d log(1 - P_ij) with respect to D_1..D_3 and kappa for the
 * probabilistic pairs (Pcat = 2): P = min(1, sqrt(Dj/Di)) * pi0/(1+pi0),
 * pi0 = 4 kappa dt / (res min(s_i, s_j)). */
/* 1 when a pair of non-null classes is impermeable (P < 1e-6) only because
 * of its kappa: the reference rule min(1, sqrt(Dj/Di)) would let walkers
 * cross. */
static int or_kappa_closes_pair(const or_geom *g, const or_derived *d)
{
    for (int a = 1; a < d->nclass; a++)
        for (int b = 1; b < d->nclass; b++) {
            if (a == b || g->kappa[a * OR_NC_MAX + b] < 0.0 || d->Pcat[a * OR_NC_MAX + b] != 0) continue;
            if (g->D[a] > 0.0 && g->D[b] > 0.0) return 1;
        }
    return 0;
}

/* Geometry properties that add boundary terms to the D derivative: bit 0
 * when some non-null voxel has, for a pair of axes (a, b) and a quadrant
 * (sa, sb), exactly two of q + sa e_a, q + sb e_b, q + sa e_a + sb e_b in its
 * own class (a step through that voxel edge ends in a different voxel
 * depending on which face the march takes first, main.cu:199-219); bit 1
 * when some non-null voxel touches a lattice face of a non-periodic axis
 * (reaching it mirrors the origin, main.cu:302-310).  Neighbours across a
 * periodic axis wrap. */
static int or_is_class(const or_geom *G, const int32_t p[3], int da, int a, int db, int b, int32_t c)
{
    const int32_t N[3] = {G->nx, G->ny, G->nz};
    int32_t q[3] = {p[0], p[1], p[2]};
    q[a] += da;
    q[b] += db;
    for (int k = 0; k < 3; k++) {
        if (q[k] >= 0 && q[k] < N[k]) continue;
        if (!((G->periodic >> k) & 1)) return 0;
        q[k] = q[k] < 0 ? q[k] + N[k] : q[k] - N[k];
    }
    return or_class_at(G, q[0], q[1], q[2]) == c;
}

static int or_geom_flags(const or_geom *G)
{
    const int32_t N[3] = {G->nx, G->ny, G->nz};
    int f = 0;
    int32_t p[3];
    for (p[0] = 0; p[0] < N[0]; p[0]++)
        for (p[1] = 0; p[1] < N[1]; p[1]++)
            for (p[2] = 0; p[2] < N[2]; p[2]++) {
                const int32_t c0 = or_class_at(G, p[0], p[1], p[2]);
                if (c0 == 0) continue;
                for (int k = 0; k < 3; k++)
                    if (!((G->periodic >> k) & 1) && (p[k] == 0 || p[k] == N[k] - 1)) f |= 2;
                for (int a = 0; a < 3; a++)
                    for (int b = a + 1; b < 3; b++)
                        for (int sa = -1; sa <= 1; sa += 2)
                            for (int sb = -1; sb <= 1; sb += 2) {
                                const int n = or_is_class(G, p, sa, a, 0, b, c0) +
                                              or_is_class(G, p, 0, a, sb, b, c0) +
                                              or_is_class(G, p, sa, a, sb, b, c0);
                                if (n == 2) f |= 1;
                            }
            }
    return f;
}

/* 1 when walkers can move between two different non-null classes. */
static int or_has_exchange(const or_derived *d)
{
    for (int i = 1; i < d->nclass; i++)
        for (int j = 1; j < d->nclass; j++)
            if (i != j && d->Pcat[i * OR_NC_MAX + j] != 0) return 1;
    return 0;
}

static void or_derive_tangent(const or_geom *g, double dt, const or_derived *d, or_tderiv *t)
{
    memset(t, 0, sizeof(*t));
    for (int c = 0; c < g->nclass; c++) {
        t->dstep[c] = (g->D[c] > 0.0) ? d->step[c] / (2.0 * g->D[c]) : 0.0;
        t->dinv[c] = (g->D[c] > 0.0) ? -d->inv_step[c] / (2.0 * g->D[c]) : 0.0;
        t->rT[c] = 1.0 / d->T2s[c];
        t->invT2[c] = 1.0 / g->T2[c];
    }
    for (int i = 1; i < g->nclass; i++) {
        for (int j = 1; j < g->nclass; j++) {
            int ij = i * OR_NC_MAX + j;
            if (d->Pcat[ij] != 2) continue;
            double dl[3] = {0.0, 0.0, 0.0};
            double dk = 0.0;
            double ratio = sqrt(g->D[j] / g->D[i]);
            if (ratio < 1.0) {
                dl[j - 1] = dl[j - 1] + 0.5 / g->D[j];
                dl[i - 1] = dl[i - 1] - 0.5 / g->D[i];
            }
            double kap = g->kappa[ij];
            if (kap >= 0.0) {
                double smin = fmin(d->step[i], d->step[j]);
                double pi0 = 4.0 * kap * dt / (g->res * smin);
                int m = (d->step[i] <= d->step[j]) ? i : j;
                dl[m - 1] = dl[m - 1] - 0.5 / (g->D[m] * (1.0 + pi0));
                dk = (kap > 0.0) ? 1.0 / (kap * (1.0 + pi0)) : 0.0;
            }
            double P = d->P[ij];
            double r = -P / (1.0 - P);
            for (int q = 0; q < 3; q++) {
                t->dlogP[ij][q] = dl[q];
                t->dlog1mP[ij][q] = r * dl[q];
            }
            t->dlogPk[ij] = dk;
            t->dlog1mPk[ij] = r * dk;
        }
    }
}

static const int OR_DIND[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
static const int OR_WIND[15][4] = {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {0, 0, 1, 1}, {0, 0, 1, 2},
                                   {0, 0, 2, 2}, {0, 1, 1, 1}, {0, 1, 1, 2}, {0, 1, 2, 2}, {0, 2, 2, 2},
                                   {1, 1, 1, 1}, {1, 1, 1, 2}, {1, 1, 2, 2}, {1, 2, 2, 2}, {2, 2, 2, 2}};

/* F = sum_q cot_q m_q(d) over the T2-weighted columns of one record row
 * (m_q = the record value without the T2 weight: sig0, dx2, dx4, sigRe) and
 * its gradient with respect to the displacement d.  The NPar_count column is
 * not weighted; the caller adds its cotangent cot[1 + c] separately. */
static void or_adj_record(const double d[3], int c, const double *cot, int nclass, int nbvec, const double *bvec,
                          const double *qs, double *F, double g[3])
{
    double f = cot[0];
    (void)c;
    g[0] = 0.0;
    g[1] = 0.0;
    g[2] = 0.0;
    const double *c2 = cot + 1 + nclass;
    for (int q = 0; q < 6; q++) {
        int a = OR_DIND[q][0], b = OR_DIND[q][1];
        f = f + c2[q] * (d[a] * d[b]);
        g[a] = g[a] + c2[q] * d[b];
        g[b] = g[b] + c2[q] * d[a];
    }
    const double *c4 = c2 + 6;
    for (int q = 0; q < 15; q++) {
        int i0 = OR_WIND[q][0], i1 = OR_WIND[q][1], i2 = OR_WIND[q][2], i3 = OR_WIND[q][3];
        f = f + c4[q] * (((d[i0] * d[i1]) * d[i2]) * d[i3]);
        g[i0] = g[i0] + c4[q] * ((d[i1] * d[i2]) * d[i3]);
        g[i1] = g[i1] + c4[q] * ((d[i0] * d[i2]) * d[i3]);
        g[i2] = g[i2] + c4[q] * ((d[i0] * d[i1]) * d[i3]);
        g[i3] = g[i3] + c4[q] * ((d[i0] * d[i1]) * d[i2]);
    }
    const double *cs = c4 + 15;
    for (int j = 0; j < nbvec; j++) {
        const double *bv = bvec + 3 * j;
        double arg = qs[j] * (d[0] * bv[0] + d[1] * bv[1] + d[2] * bv[2]);
        double sn, cn;
        or_sincos_f64(arg, &sn, &cn);
        f = f + cs[j] * cn;
        double t = (cs[j] * sn) * qs[j];
        g[0] = g[0] - t * bv[0];
        g[1] = g[1] - t * bv[1];
        g[2] = g[2] - t * bv[2];
    }
    *F = f;
}

typedef struct {
    int nval;
    int64_t Tstep;
    double *qscale;   /* [timepoints][nbvec]: sqrt(bval/TD), main.cu:375 */
    int nbound;       /* boundaries 1..nbound (bounds[0] = 0) */
    int64_t bounds[OR_MAX_BOUNDS + 1];
    int64_t pgmask;   /* PGSE sums flush after step i when (i + 1) & pgmask == 0 */
    int32_t *gk;      /* [ngrad][3] indices (ka, kb, ke) into bounds */
    double *snapbuf;  /* per thread */
    double *phasebuf; /* per thread */
    int32_t *gorder;  /* gradients sorted by lobe triple (stable) */
    const double *cot_rec; /* adjoint: [timepoints][nval] or NULL */
    const double *cot_sig; /* adjoint: [ngrad] or NULL */
    const or_tderiv *td;   /* adjoint constants or NULL */
    int jump;              /* adjoint: D derivative across exchange faces (or_jump_step) */
    double *snapJ;    /* per thread [OR_BATCH][nb+1][3][3] tangent prefix sums */
} or_jobaux;

/* PGSE phase from window sums: phase = (|g| * (g_hat . (S1 - S2))) * dt with
 * S1 = P[ka], S2 = P[ke] - P[kb] (P = prefix sums of the displacement at the
 * lobe boundaries).  Equal to main_PGSE.cu:340-346 up to reassociation. */
static inline double or_pgse_phase(const or_jobaux *aux, const double *gr, int jg, const double *snap,
                                   double dt)
{
    const int32_t *k = aux->gk + 3 * jg;
    double df[3];
    for (int c = 0; c < 3; c++) {
        double s1 = snap[k[0] * 3 + c];
        double s2 = snap[k[2] * 3 + c] - snap[k[1] * 3 + c];
        df[c] = s1 - s2;
    }
    double dot = df[0] * gr[3] + df[1] * gr[4] + df[2] * gr[5];
    return (gr[2] * dot) * dt;
}

#define FABS fabs
#define SQRT sqrt
#define FMA fma
#define FLOOR floor
#define SINCOS_TURNS or_sincos_turns_f64
#define PIREF 0.9999999988573335
#define REAL double
#define SFX _f64
#define OR_IS_F32 0
#include "rms_oracle_core.inc"
#undef FABS
#undef SQRT
#undef FMA
#undef FLOOR
#undef SINCOS_TURNS
#undef PIREF
#undef REAL
#undef SFX
#undef OR_IS_F32

#define FABS fabsf
#define SQRT sqrtf
#define FMA fmaf
#define FLOOR floorf
#define SINCOS_TURNS or_sincos_turns_f32
#define PIREF 1.0f
#define REAL float
#define SFX _f32
#define OR_IS_F32 1
#include "rms_oracle_core.inc"
#undef FABS
#undef SQRT
#undef FMA
#undef FLOOR
#undef SINCOS_TURNS
#undef PIREF
#undef REAL
#undef SFX
#undef OR_IS_F32

/* Single step with supplied exchange draws (exact-tracer test hook). */
typedef struct {
    double *u;
} or_fixed_draws;
static double or_draw_fixed(void *ctx, int j) { return ((or_fixed_draws *)ctx)->u[j]; }

int or_march_f64(const or_geom *geom, const or_derived *d, int32_t g[3], double o[3], double v[3],
                 int32_t *cls, double u_exch[4], int32_t cnt[4], int32_t gi[3], double oi[3])
{
    or_walker_f64 w;
    memset(&w, 0, sizeof(w));
    for (int k = 0; k < 3; k++) {
        w.g[k] = g[k];
        w.o[k] = o[k];
        w.gi[k] = gi[k];
        w.oi[k] = oi[k];
    }
    w.c = *cls;
    or_dreal_f64 dr;
    or_dreal_make_f64(d, &dr);
    or_fixed_draws fd = {u_exch};
    int nd = or_marchw_f64(&w, geom, d, &dr, v, or_draw_fixed, &fd, NULL, NULL, NULL);
    (void)nd;
    for (int k = 0; k < 3; k++) {
        g[k] = w.g[k];
        o[k] = w.o[k];
        gi[k] = w.gi[k];
        oi[k] = w.oi[k];
    }
    *cls = w.c;
    for (int k = 0; k < 4; k++) cnt[k] = (int32_t)w.cnt[k];
    return nd;
}

/* One straight-move fast-path step from a given state (test entry): returns
 * 1 and updates the state (voxel, offset, origin) when the fast path
 * applies, 0 otherwise.  Exchange draws j take u_exch[j]; cnt = crossings,
 * reflections, mirrors, draws. */
int or_fast_step_f64(const or_geom *geom, const or_derived *d, int32_t g[3], double o[3], const double v[3],
                     int32_t cls, double u_exch[4], int32_t cnt[4], int32_t gi[3], double oi[3])
{
    or_walker_f64 w;
    memset(&w, 0, sizeof(w));
    for (int k = 0; k < 3; k++) {
        w.g[k] = g[k];
        w.o[k] = o[k];
        w.gi[k] = gi[k];
        w.oi[k] = oi[k];
    }
    w.c = cls;
    or_dreal_f64 dr;
    or_dreal_make_f64(d, &dr);
    or_fixed_draws fd = {u_exch};
    int r = or_fast_f64(&w, geom, d, &dr, v, or_draw_fixed, &fd, NULL, NULL, NULL);
    for (int k = 0; k < 3; k++) {
        g[k] = w.g[k];
        o[k] = w.o[k];
        gi[k] = w.gi[k];
        oi[k] = w.oi[k];
    }
    cnt[0] = (int32_t)w.cnt[0];
    cnt[1] = (int32_t)w.cnt[1];
    cnt[2] = (int32_t)w.cnt[2];
    cnt[3] = (int32_t)w.cnt[4];
    return r;
}

int32_t or_wrap_cells_rule(const or_geom *g)
{
    if ((g->periodic & 7) == 0) return 0;
    int64_t walk = 0, side = 0;
    for (int32_t x = 0; x < g->nx; x++)
        for (int32_t y = 0; y < g->ny; y++)
            for (int32_t z = 0; z < g->nz; z++) {
                if (or_class_at(g, x, y, z) == 0) continue;
                walk++;
                const int on = (((g->periodic >> 0) & 1) && (x == 0 || x == g->nx - 1)) ||
                               (((g->periodic >> 1) & 1) && (y == 0 || y == g->ny - 1)) ||
                               (((g->periodic >> 2) & 1) && (z == 0 || z == g->nz - 1));
                side += on;
            }
    return 256 * side > walk ? 1 : 0;
}

/* ------------------------------------------------------------------ */
/* Driver                                                                */
/* ------------------------------------------------------------------ */
static int validate_job(const or_job *J, const or_params *p, const or_derived *d)
{
    const or_geom *G = J->geom;
    if (G->nx < 1 || G->ny < 1 || G->nz < 1) return fail("lattice dims must be >= 1");
    if (G->label_bytes != 1 && G->label_bytes != 2) return fail("label_bytes must be 1 or 2");
    if (p->timepoints < 1 || p->TN < p->timepoints) return fail("TN must be >= timepoints");
    if (p->TN > 2147483647LL) return fail("TN too large");
    int64_t R = p->replicas < 1 ? 1 : p->replicas;
    if (J->walker_base < 0 || J->walker_base % (OR_BATCH * R) != 0)
        return fail("walker_base must be a non-negative multiple of 64 * replicas");
    for (int c = 1; c < G->nclass; c++) {
        if (!(G->D[c] > 0.0)) return fail("non-null classes need D > 0");
        if (!(d->step[c] <= 1.0 - 0x1p-20)) return fail("step size must be below one voxel by at least 2^-20");
        if (!(G->T2[c] > 0.0)) return fail("T2 must be positive");
    }
    return 1;
}

static int or_run_impl(const or_job *jobs, int32_t njobs, const or_params *p, or_out *outs, const or_adj *adj);

int or_run(const or_job *jobs, int32_t njobs, const or_params *p, or_out *outs)
{
    return or_run_impl(jobs, njobs, p, outs, NULL);
}

int or_run_adjoint(const or_job *jobs, int32_t njobs, const or_params *p, const or_adj *adj, or_out *outs)
{
    return or_run_impl(jobs, njobs, p, outs, adj);
}

static const or_jobaux *g_sort_aux;
static int cmp_triple(const void *a, const void *b)
{
    int ia = *(const int32_t *)a, ib = *(const int32_t *)b;
    const int32_t *ka = g_sort_aux->gk + 3 * ia, *kb = g_sort_aux->gk + 3 * ib;
    for (int q = 0; q < 3; q++)
        if (ka[q] != kb[q]) return ka[q] < kb[q] ? -1 : 1;
    return ia < ib ? -1 : (ia > ib);
}

/* Per-job state of a run: derived parameters, tables and partial sums. */
typedef struct {
    or_derived d;
    or_tderiv td;
    or_jobaux aux;
    double *TD;
    int64_t nbatch, b0; /* batches of the job and its first work item */
    double *part, *psig, *pgrad;
} or_jobrun;

static void or_jobrun_free(or_jobrun *jr)
{
    free(jr->aux.gorder);
    free(jr->aux.gk);
    free(jr->aux.qscale);
    free(jr->TD);
    free(jr->part);
    free(jr->psig);
    free(jr->pgrad);
}

/* Tables of one job: q scales, PGSE boundaries and lobe-triple order. */
static int or_job_setup(const or_job *J, const or_params *p, const or_adj *adj, int jb, int plan_pgse,
                        or_jobrun *jr)
{
    if (!or_derive(J->geom, p->dt, &jr->d)) return 0;
    if (!validate_job(J, p, &jr->d)) return 0;
    or_jobaux *aux = &jr->aux;
    if (adj) {
        or_derive_tangent(J->geom, p->dt, &jr->d, &jr->td);
        aux->td = &jr->td;
        aux->cot_rec = adj[jb].cot_rec;
        aux->cot_sig = adj[jb].cot_sig;
        /* boundary terms of the D derivative (exchange faces, lattice-edge
         * mirrors, voxel edges) when no job of the run has a PGSE table (the
         * continuations carry no PGSE prefix sums) */
        aux->jump = !plan_pgse && !p->no_boundary;
    }
    aux->nval = or_nval(J->geom->nclass, J->nbvec);
    aux->Tstep = p->TN / p->timepoints;
    {
        /* sum(g) between flushes stays below 2^31: the largest k <= 16 with
         * 2^k max(nx, ny, nz) < 2^31 */
        int64_t m = J->geom->nx;
        if (J->geom->ny > m) m = J->geom->ny;
        if (J->geom->nz > m) m = J->geom->nz;
        if (m < 1) m = 1;
        int k = 16;
        while (k > 0 && (m << k) > 2147483647LL) k--;
        aux->pgmask = ((int64_t)1 << k) - 1;
    }
    int T = p->timepoints;
    jr->TD = (double *)malloc(sizeof(double) * T);
    or_diff_times(p->dt, p->TN, T, jr->TD);
    int nbq = J->nbvec > 0 ? J->nbvec : 1;
    aux->qscale = (double *)calloc((size_t)T * nbq, sizeof(double));
    for (int r = 0; r < T; r++)
        for (int j = 0; j < J->nbvec; j++) aux->qscale[(size_t)r * nbq + j] = sqrt(J->bval[j] / jr->TD[r]);
    /* PGSE boundaries */
    if (J->ngrad > 0) {
        int64_t *abe = (int64_t *)malloc(sizeof(int64_t) * 3 * J->ngrad);
        or_pgse_lobes(p->dt, p->TN, J->ngrad, J->grad, abe);
        int64_t *tmp = (int64_t *)malloc(sizeof(int64_t) * (3 * (size_t)J->ngrad + 1));
        int n = 0;
        tmp[n++] = p->TN;
        for (int j = 0; j < J->ngrad; j++)
            for (int q = 0; q < 3; q++) tmp[n++] = abe[3 * j + q];
        /* sort unique, drop 0 */
        for (int a = 1; a < n; a++) {
            int64_t v = tmp[a];
            int b = a - 1;
            while (b >= 0 && tmp[b] > v) {
                tmp[b + 1] = tmp[b];
                b--;
            }
            tmp[b + 1] = v;
        }
        aux->bounds[0] = 0;
        int nb = 0;
        for (int a = 0; a < n; a++) {
            if (tmp[a] == 0) continue;
            if (nb > 0 && aux->bounds[nb] == tmp[a]) continue;
            if (nb >= OR_MAX_BOUNDS) {
                free(abe);
                free(tmp);
                return fail("too many distinct PGSE lobe boundaries");
            }
            aux->bounds[++nb] = tmp[a];
        }
        free(tmp);
        aux->nbound = nb;
        aux->gk = (int32_t *)malloc(sizeof(int32_t) * 3 * J->ngrad);
        for (int j = 0; j < J->ngrad; j++)
            for (int q = 0; q < 3; q++) {
                int64_t v = abe[3 * j + q];
                int k = 0;
                while (aux->bounds[k] != v) k++;
                aux->gk[3 * j + q] = k;
            }
        free(abe);
        aux->gorder = (int32_t *)malloc(sizeof(int32_t) * J->ngrad);
        for (int j = 0; j < J->ngrad; j++) aux->gorder[j] = j;
        g_sort_aux = aux;
        qsort(aux->gorder, J->ngrad, sizeof(int32_t), cmp_triple);
    }
    jr->nbatch = (J->nwalk + OR_BATCH - 1) / OR_BATCH;
    jr->pgrad = adj ? (double *)calloc((size_t)jr->nbatch * 16, sizeof(double)) : NULL;
    jr->part = (double *)calloc((size_t)jr->nbatch * T * aux->nval, sizeof(double));
    jr->psig = J->ngrad > 0 ? (double *)calloc((size_t)jr->nbatch * J->ngrad, sizeof(double)) : NULL;
    return 1;
}

/* Replica sums (batch order) and totals (replica order) of one job. */
static void or_job_reduce(const or_job *J, const or_params *p, const or_adj *adj, int jb, or_jobrun *jr,
                          or_out *O, int R)
{
    const int T = p->timepoints;
    const int64_t nbatch = jr->nbatch;
    size_t rv = (size_t)T * jr->aux.nval;
    double *rep = (double *)calloc((size_t)R * rv, sizeof(double));
    for (int64_t b = 0; b < nbatch; b++) {
        double *dst = rep + (size_t)(b % R) * rv;
        const double *src = jr->part + (size_t)b * rv;
        for (size_t q = 0; q < rv; q++) dst[q] = dst[q] + src[q];
    }
    for (size_t q = 0; q < rv; q++) {
        double s = 0.0;
        for (int r = 0; r < R; r++) s = s + rep[(size_t)r * rv + q];
        O->rec[q] = s;
    }
    if (O->rec_rep) memcpy(O->rec_rep, rep, sizeof(double) * R * rv);
    free(rep);
    if (J->ngrad > 0 && O->sig) {
        double *srep = (double *)calloc((size_t)R * J->ngrad, sizeof(double));
        for (int64_t b = 0; b < nbatch; b++)
            for (int j = 0; j < J->ngrad; j++)
                srep[(size_t)(b % R) * J->ngrad + j] =
                    srep[(size_t)(b % R) * J->ngrad + j] + jr->psig[(size_t)b * J->ngrad + j];
        for (int j = 0; j < J->ngrad; j++) {
            double s = 0.0;
            for (int r = 0; r < R; r++) s = s + srep[(size_t)r * J->ngrad + j];
            O->sig[j] = s;
        }
        if (O->sig_rep) memcpy(O->sig_rep, srep, sizeof(double) * R * J->ngrad);
        free(srep);
    }
    if (adj) {
        double *grep = (double *)calloc((size_t)R * 16, sizeof(double));
        for (int64_t b = 0; b < nbatch; b++)
            for (int q = 0; q < 16; q++)
                grep[(size_t)(b % R) * 16 + q] = grep[(size_t)(b % R) * 16 + q] + jr->pgrad[(size_t)b * 16 + q];
        for (int q = 0; q < 16; q++) {
            double s = 0.0;
            for (int r = 0; r < R; r++) s = s + grep[(size_t)r * 16 + q];
            adj[jb].grad[q] = s;
        }
        if (adj[jb].grad_rep) memcpy(adj[jb].grad_rep, grep, sizeof(double) * R * 16);
        /* Exchange faces, lattice-edge mirrors and voxel edges make the
         * walk's end discontinuous in D; their boundary terms (or_jump_step,
         * or_corner_step) are there for runs without PGSE tables: with one
         * the D entries are NaN when the geometry has any of them. */
        if (!jr->aux.jump && (or_has_exchange(&jr->d) || or_geom_flags(J->geom) != 0)) {
            for (int c = 1; c < OR_NC_MAX; c++) {
                adj[jb].grad[c] = NAN;
                if (adj[jb].grad_rep)
                    for (int r = 0; r < R; r++) adj[jb].grad_rep[(size_t)r * 16 + c] = NAN;
            }
        }
        /* a pair closed by its kappa draws nothing: no kappa estimate */
        if (or_kappa_closes_pair(J->geom, &jr->d)) {
            adj[jb].grad[8] = NAN;
            if (adj[jb].grad_rep)
                for (int r = 0; r < R; r++) adj[jb].grad_rep[(size_t)r * 16 + 8] = NAN;
        }
        free(grep);
    }
}

/* All jobs of a run: tables first, then every (job, batch) pair spread over
 * the OpenMP threads (each batch is computed on its own, so the thread
 * count and schedule do not change any bit), then the ordered reductions. */
static int or_run_impl(const or_job *jobs, int32_t njobs, const or_params *p, or_out *outs, const or_adj *adj)
{
    g_err[0] = 0;
    int R = p->replicas < 1 ? 1 : p->replicas;
    or_jobrun *jr = (or_jobrun *)calloc((size_t)(njobs > 0 ? njobs : 1), sizeof(or_jobrun));
    int ok = 1;
    int64_t nwork = 0;
    int maxnb = 0, maxng = 0, maxlane = 0;
    int plan_pgse = 0;
    for (int jb = 0; jb < njobs; jb++) plan_pgse = plan_pgse || jobs[jb].ngrad > 0;
    for (int jb = 0; jb < njobs && ok; jb++) {
        ok = or_job_setup(&jobs[jb], p, adj, jb, plan_pgse, &jr[jb]);
        if (!ok) break;
        jr[jb].b0 = nwork;
        nwork += jr[jb].nbatch;
        if (jr[jb].aux.nbound > maxnb) maxnb = jr[jb].aux.nbound;
        if (jobs[jb].ngrad > maxng) maxng = jobs[jb].ngrad;
        int lanew = jr[jb].aux.nval > jobs[jb].ngrad ? jr[jb].aux.nval : jobs[jb].ngrad;
        if (lanew > maxlane) maxlane = lanew;
    }
    if (!ok) {
        for (int jb = 0; jb < njobs; jb++) or_jobrun_free(&jr[jb]);
        free(jr);
        return 0;
    }
    uint64_t *cnt_tot = (uint64_t *)calloc((size_t)(njobs > 0 ? njobs : 1) * 5, sizeof(uint64_t));
    int run_ok = 1;
#ifdef _OPENMP
    int nth = p->nthreads > 0 ? p->nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nth)
#endif
    {
        uint64_t *cnt_acc = (uint64_t *)calloc((size_t)(njobs > 0 ? njobs : 1) * 5, sizeof(uint64_t));
        double *snapbuf = (double *)malloc(sizeof(double) * OR_BATCH * (maxnb + 1) * 3 + 8);
        double *phasebuf = (double *)malloc(sizeof(double) * OR_BATCH * (maxng + 1));
        const size_t nsj = (size_t)OR_BATCH * (maxnb + 1) * 9 + 8;
        double *snapJ = (double *)calloc(nsj, sizeof(double));
        double *lanebuf = (double *)malloc(sizeof(double) * OR_BATCH * (maxlane > 0 ? maxlane : 1));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int64_t w = 0; w < nwork; w++) {
            int lo = 0, hi = njobs - 1; /* the job whose batches hold work item w */
            while (lo < hi) {
                int mid = (lo + hi + 1) / 2;
                if (jr[mid].b0 <= w)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            const int jb = lo;
            const or_job *J = &jobs[jb];
            or_jobrun *r = &jr[jb];
            const int64_t b = w - r->b0;
            or_jobaux la = r->aux;
            la.snapbuf = snapbuf;
            la.phasebuf = phasebuf;
            la.snapJ = snapJ;
            memset(snapJ, 0, sizeof(double) * nsj);
            const int T = p->timepoints;
            double *pg = r->pgrad ? r->pgrad + (size_t)b * 16 : NULL;
            int rc;
            if (p->prec == OR_PREC_F32)
                rc = or_batch_f32(J, &r->d, p, &la, b, r->part + (size_t)b * T * la.nval,
                                  r->psig ? r->psig + (size_t)b * J->ngrad : NULL, lanebuf, &outs[jb],
                                  cnt_acc + 5 * jb, pg);
            else
                rc = or_batch_f64(J, &r->d, p, &la, b, r->part + (size_t)b * T * la.nval,
                                  r->psig ? r->psig + (size_t)b * J->ngrad : NULL, lanebuf, &outs[jb],
                                  cnt_acc + 5 * jb, pg);
            if (!rc) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
                run_ok = 0;
            }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        for (int q = 0; q < 5 * njobs; q++) cnt_tot[q] += cnt_acc[q];
        free(cnt_acc);
        free(snapbuf);
        free(phasebuf);
        free(snapJ);
        free(lanebuf);
    }
    if (run_ok) {
        for (int jb = 0; jb < njobs; jb++) {
            or_job_reduce(&jobs[jb], p, adj, jb, &jr[jb], &outs[jb], R);
            for (int q = 0; q < 5; q++) outs[jb].counters[q] = cnt_tot[5 * jb + q];
        }
    }
    for (int jb = 0; jb < njobs; jb++) or_jobrun_free(&jr[jb]);
    free(jr);
    free(cnt_tot);
    if (!run_ok) return fail("walker initialisation found no non-null voxel");
    return 1;
}
