/*
 * rms_oracle.h -- CPU restatement of the RMS elastic-collision random walk.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The product (librms_amd.so and its CLIs) never links it.
 *
 * What it restates (reference files, relative to the upstream repository):
 *   hpc_code/lib/main.cu:63-384       propagate(): init, step, records
 *   hpc_code/lib/main_PGSE.cu:336-357 PGSE phase accumulation and final signal
 *   hpc_code/lib/main.cu:429-458      derived parameters (T2/dt, step, Pij)
 *   hpc_code/lib/main.cu:484-487      diffusion times TD
 * plus the documented deviations that the HIP engine shares (counter-based
 * Philox stream, integer voxel + fractional offset positions, PGSE window
 * sums, deterministic batch/replica reduction order).  See rms_oracle.c.
 */
#ifndef RMS_ORACLE_H
#define RMS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_NC_MAX 4      /* classes (reference Nc_max = 3, main.cu:37) */
#define OR_NITE 4        /* face iterations per step (main.cu:36) */
#define OR_BATCH 64      /* walkers per reduction batch (one wavefront) */
#define OR_MAX_BOUNDS 64 /* distinct PGSE lobe boundaries */

enum { OR_PREC_F32 = 0, OR_PREC_F64 = 1 };

typedef struct {
    int32_t nx, ny, nz;          /* NPix1, NPix2, NPix3 */
    int32_t label_bytes;         /* 1 (uint8) or 2 (uint16) */
    const void *labels;          /* nx*ny*nz, index (x*ny + y)*nz + z */
    const uint8_t *label_class;  /* label -> class, NULL = identity */
    int32_t nlabel;              /* entries in label_class */
    int32_t nclass;              /* Nc */
    double res;                  /* voxel size, um */
    double D[OR_NC_MAX];         /* um^2/ms */
    double T2[OR_NC_MAX];        /* ms, +inf = no relaxation */
    double kappa[OR_NC_MAX * OR_NC_MAX]; /* um/ms, < 0 = diffusivity-ratio rule */
    int32_t periodic;            /* bit k: axis k wraps around instead of mirroring */
    /* 1: steps through a periodic side take the fast path (the engine's cell
     * words then name the voxel on the other side); 0: the march takes them.
     * Set by or_wrap_cells_rule. */
    int32_t wrap_cells;
} or_geom;

/* The engine's choice of wrap_cells: 1 when the lattice has a periodic axis
 * and more than 1/256 of its non-null voxels lie on a periodic side. */
int32_t or_wrap_cells_rule(const or_geom *g);

typedef struct {
    int32_t nclass;
    double step[OR_NC_MAX];      /* voxels per step */
    double inv_step[OR_NC_MAX];
    double T2s[OR_NC_MAX];       /* T2 in steps */
    double P[OR_NC_MAX * OR_NC_MAX];
    int32_t Pcat[OR_NC_MAX * OR_NC_MAX]; /* 0 reflect, 1 cross, 2 draw */
} or_derived;

typedef struct {
    const or_geom *geom;
    int64_t nwalk;
    uint64_t seed;
    int32_t nbvec;               /* narrow-pulse b-table (0 = none) */
    const double *bval;          /* ms/um^2 */
    const double *bvec;          /* 3*nbvec */
    int32_t ngrad;               /* PGSE table (0 = none) */
    const double *grad;          /* 6*ngrad rows: Delta, delta, |g|, gx, gy, gz */
    int64_t walker_base;         /* Philox id of the first walker (multiple of 64*replicas) */
    /* walk a subset: walker k of the run is walker wlist[k] of the job
     * (NULL: walker k); nwalk entries */
    const uint32_t *wlist;
} or_job;

typedef struct {
    double dt;
    int64_t TN;
    int32_t timepoints;
    int32_t prec;                /* OR_PREC_F32 / OR_PREC_F64 */
    int32_t replicas;            /* batch-means replicas (>= 1) */
    int32_t pgse_bruteforce;     /* 1: per-gradient per-step loop (slow) */
    int32_t nthreads;            /* OpenMP threads, 0 = default */
    /* adjoint without a PGSE table: 0 = pathwise plus the boundary terms of
     * D (the default), 1 = the tangent pass alone (pathwise D and the score
     * of the draws, T2, kappa) */
    int32_t no_boundary;
    int32_t reserved;
} or_params;

typedef struct {
    double *rec;          /* [timepoints * nval] totals */
    double *rec_rep;      /* [replicas * timepoints * nval] or NULL */
    double *sig;          /* [ngrad] PGSE totals or NULL */
    double *sig_rep;      /* [replicas * ngrad] or NULL */
    uint64_t counters[5]; /* crossings, reflections, edge mirrors, truncations, draws */
    /* optional final walker state, each [nwalk] (NULL = skip) */
    int32_t *fin_g;       /* [3*nwalk], axis-major */
    double *fin_o;        /* [3*nwalk] */
    int32_t *fin_gi;      /* [3*nwalk] */
    double *fin_oi;       /* [3*nwalk] */
    int32_t *fin_c;       /* [nwalk] */
    uint32_t *fin_cnt;    /* [5*nwalk] per-walker counters */
    /* adjoint pass: per walker, bit c = visited class c, bit 4 + a*4 + b =
     * drew at an exchange face from class a into class b (NULL = skip) */
    uint32_t *fin_vis;    /* [nwalk] */
} or_out;

/* ---- primitives (exported for known-answer tests) ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_u01_f64(uint32_t a, uint32_t b);
float or_u01_f32(uint32_t w);
void or_sincos_turns_f64(double t, double *s, double *c);
void or_sincos_turns_f32(float t, float *s, float *c);
void or_turn_table(float s[256], float c[256]);
double or_exp_f64(double x);
void or_sincos_f64(double x, double *s, double *c);
double or_cos_f64(double x);

/* ---- host-side derivations (main.cu:429-458, 484-487) ---- */
int or_derive(const or_geom *g, double dt, or_derived *d);
int32_t or_nval(int32_t nclass, int32_t nbvec);
void or_diff_times(double dt, int64_t TN, int32_t timepoints, double *TD);
/* PGSE lobe index intervals from main_PGSE.cu:341-345 comparisons.
 * out: per gradient (a, b, e): lobe 1 = [0,a), lobe 2 = [b,e). */
void or_pgse_lobes(double dt, int64_t TN, int32_t ngrad, const double *grad, int64_t *abe);

/* ---- single step of the face march (for the exact-tracer test) ----
 * state: g[3], o[3], v[3] in; on return g, o, v updated, returns the number
 * of face iterations used; cnt = {crossings, reflections, mirrors, truncated}.
 * Pcat/P for the class pair; draws supplied by u_exch[4]. */
int or_march_f64(const or_geom *geom, const or_derived *d, int32_t g[3], double o[3],
                 double v[3], int32_t *cls, double u_exch[4], int32_t cnt[4],
                 int32_t gi[3], double oi[3]);

/* ---- one straight-move fast-path step (for the fast-path vs march tests):
 * returns 1 (state updated, cnt = {crossings, reflections, mirrors, draws})
 * when the step is taken without the march, 0 otherwise; exchange draw j is
 * u_exch[j]. ---- */
int or_fast_step_f64(const or_geom *geom, const or_derived *d, int32_t g[3], double o[3], const double v[3],
                     int32_t cls, double u_exch[4], int32_t cnt[4], int32_t gi[3], double oi[3]);

/* ---- full simulation of a list of jobs ---- */
int or_run(const or_job *jobs, int32_t njobs, const or_params *p, or_out *outs);

/* Adjoint pass: derivatives of sum(cot_rec*rec) + sum(cot_sig*sig) with
 * respect to D[c] (grad[c], c = 1..3), T2[c] (grad[4+c]) and kappa (grad[8]).
 * D entries: pathwise, plus the score of the exchange draws and the jump of
 * the class change at exchange faces when walkers can cross between
 * non-null classes; NaN in that case when a job of the run has a PGSE table
 * (the continuations of the jump term carry no PGSE sums).  The kappa entry is
 * NaN when a pair of non-null classes is impermeable only because of its
 * kappa (no draw happens there, so the score function sees nothing). */
typedef struct {
    const double *cot_rec;  /* [timepoints*nval] or NULL */
    const double *cot_sig;  /* [ngrad] or NULL */
    double *grad;           /* [16] out */
    double *grad_rep;       /* [replicas*16] out or NULL */
} or_adj;
int or_run_adjoint(const or_job *jobs, int32_t njobs, const or_params *p, const or_adj *adj, or_out *outs);

const char *or_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
