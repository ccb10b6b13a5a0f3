/*
 * rms_amd.h -- C ABI of the MI355X RMS random-walk engine (librms_amd.so).
 *
 * The reference engine has no library interface: hpc_code/lib/main.cu and
 * main_PGSE.cu are single programs whose main() reads text files from the
 * working directory, launches propagate<<<>>> once and writes text files back
 * (main.cu:388-621, main_PGSE.cu:414-675).  This ABI is what that process
 * boundary becomes; the drop-in CLIs rms_main / rms_main_pgse are thin
 * clients of it.  Each entry point names the reference code it replaces.
 *
 * Conventions: caller-owned host buffers, device buffers owned by the
 * context, one context per host thread (contexts are not internally
 * synchronised), each context owns one HIP stream, every call returns a
 * status and rms_last_error() describes the last failure.  No C++
 * exceptions cross the ABI.
 */
#ifndef RMS_AMD_H
#define RMS_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RMS_NC_MAX 4        /* classes per geometry; reference Nc_max = 3 (main.cu:37) */
#define RMS_NITE 4          /* face iterations per step (main.cu:36) */
#define RMS_REF_TIMEPOINTS 1000 /* records per run in the reference (main.cu:35) */
#define RMS_NPARAM 16       /* adjoint parameters: dD[4], dT2[4], dkappa, pad */

typedef enum {
    RMS_OK = 0,
    RMS_ERR_INVALID_ARG = 1,
    RMS_ERR_HIP = 2,
    RMS_ERR_OUT_OF_MEMORY = 3,
    RMS_ERR_NO_DEVICE = 4,
    RMS_ERR_STATE = 5,
    RMS_ERR_UNSUPPORTED = 6
} rms_status;

enum { RMS_PREC_F32 = 0, RMS_PREC_F64 = 1 };

typedef struct rms_ctx rms_ctx;
typedef struct rms_plan rms_plan;

typedef struct {
    int32_t device;       /* HIP device ordinal */
    int32_t reserved;
} rms_device_opts;

/* A voxelised geometry with per-class physics.  Replaces the reads of
 * fiber.txt (main.cu:460-466), diffusivity.txt (:421-427), T2.txt (:429-436)
 * and the Pij construction (:444-458). */
typedef struct {
    int32_t nx, ny, nz;          /* NPix1, NPix2, NPix3 */
    int32_t label_bytes;         /* 1 = uint8, 2 = uint16 labels */
    const void *labels;          /* host, nx*ny*nz, index (x*ny + y)*nz + z */
    const uint8_t *label_class;  /* label -> class map, NULL = label is the class */
    int32_t nlabel;              /* entries of label_class */
    int32_t nclass;              /* Nc; class 0 is the null (impermeable) class */
    double res;                  /* voxel size, um */
    double D[RMS_NC_MAX];        /* um^2/ms */
    double T2[RMS_NC_MAX];       /* ms; INFINITY = no relaxation */
    /* permeability between classes (um/ms), row-major [i][j];
     * < 0 selects the reference rule P_ij = min(1, sqrt(D_j/D_i)) */
    double kappa[RMS_NC_MAX * RMS_NC_MAX];
    /* lattice edges: bit k set = axis k (x, y, z) wraps around (periodic
     * tiling); clear = the reference's mirror (main.cu:246-265, 302-310) */
    int32_t periodic;
    /* device storage of the cell words (the per-voxel neighbour relations the
     * walk reads): RMS_STORE_DENSE one 8-byte word per voxel; RMS_STORE_BRICKS
     * 2x2x4-voxel bricks, a 4-byte table entry per brick and the 16 words
     * (one 128-byte line) only of bricks where a non-null voxel has a
     * neighbour of another class or outside the lattice (about 2.6 bytes
     * per voxel in all for an axon with extra-axonal space, against 9, at
     * some 10 % of walk speed); RMS_STORE_AUTO keeps dense words while the
     * device has room and picks bricks once the dense words would leave
     * under 1/8 of its memory free.  Results are identical; the class
     * lattice (1 byte per voxel) is kept either way. */
    int32_t storage;
} rms_geometry;

enum { RMS_STORE_AUTO = 0, RMS_STORE_DENSE = 1, RMS_STORE_BRICKS = 2 };

/* One simulation on one geometry (one reference setup directory). */
typedef struct {
    int32_t geom_id;             /* from rms_add_geometry */
    int64_t nwalk;               /* NPar */
    uint64_t seed;               /* Philox key (replaces /dev/random, main.cu:491-496) */
    int32_t nbvec;               /* narrow-pulse b-table (bval.txt/bvec.txt), 0 = none */
    const double *bval;          /* [nbvec] ms/um^2 */
    const double *bvec;          /* [nbvec*3] */
    int32_t ngrad;               /* PGSE table (gradient_time.txt), 0 = none */
    const double *grad;          /* [ngrad*6] rows: Delta, delta, |g|, gx, gy, gz */
    /* id of the job's first walker in the Philox counter (reference: one
     * walker per thread index, main.cu:64-70).  0 for a whole run; a shard of
     * a larger run passes its first global walker id, a multiple of
     * 64 * replicas, so its walkers follow the same trajectories and its
     * batches fall into the same replicas as in the unsharded run. */
    int64_t walker_base;
} rms_job;

typedef struct {
    double dt;                   /* ms */
    int64_t TN;                  /* steps */
    int32_t timepoints;          /* records (reference: 1000) */
    int32_t precision;           /* RMS_PREC_F32 (FP32 offsets) or RMS_PREC_F64 */
    int32_t replicas;            /* batch-means replicas for standard errors (>= 1) */
    int64_t chunk_steps;         /* steps per launch, 0 = all in one launch */
    int32_t use_graph;           /* capture the launch sequence in a hipGraph */
    int32_t block_threads;       /* walkers per workgroup, 0 = auto (256) */
    int32_t grid_blocks;         /* workgroups, 0 = auto (resident limit) */
    int32_t xcd_split;           /* 1 = give each XCD a contiguous batch range */
    /* 1 = leave rms_job_out.counters at zero: the per-step diagnostic sums
     * (crossings, reflections, mirrors, draws; the reference has none) cost
     * about 3 % of the walk.  Final walker states always count. */
    int32_t skip_counters;
} rms_run_params;

/* Per-job outputs (caller-owned host buffers; NULL pointers are skipped).
 * Record layout per time point: [sig0, NPar_count[nclass], dx2[6], dx4[15],
 * sigRe[nbvec]] -- the sig0.txt, NPar_count.txt, dx2.txt, dx4.txt and
 * sigRe.txt columns of main.cu:549-604. */
typedef struct {
    double *rec;                 /* [timepoints * nval] */
    double *rec_rep;             /* [replicas * timepoints * nval] */
    double *sig;                 /* [ngrad] PGSE signal (sig.txt, main_PGSE.cu:648-651) */
    double *sig_rep;             /* [replicas * ngrad] */
    /* crossings, reflections, edge mirrors, truncated steps, exchange draws
     * (zero when rms_run_params.skip_counters is set) */
    uint64_t counters[5];
    /* optional final walker state, axis-major [3][nwalk] etc. */
    int32_t *fin_g;
    double *fin_o;
    int32_t *fin_gi;
    double *fin_oi;
    int32_t *fin_c;
    uint32_t *fin_cnt;           /* [5][nwalk] */
} rms_job_out;

typedef struct {
    double walk_ms;              /* hipEvent time of the walk launches */
    double reduce_ms;            /* hipEvent time of the ordered reductions */
    double total_ms;             /* hipEvent time of the whole execute */
    int32_t grid;                /* workgroups used */
    int32_t block;               /* threads per workgroup */
    int32_t launches;            /* walk launches */
    int32_t reserved;
    int64_t walker_steps;        /* walker-steps of every walk launch (walkers x steps) */
} rms_timing;

/* Adjoint outputs: derivatives of sum(cot_rec * rec) + sum(cot_sig * sig)
 * with respect to D[c] (index c), T2[c] (4 + c) and kappa (8).
 * T2: analytic (the weights only; paths do not depend on T2).
 * kappa: score function of the exchange draws against a leave-one-out
 * baseline (the mean of the other walkers of the 64-walker batch).  NaN when
 * a pair of non-null classes is impermeable (P < 1e-6) only because of its
 * kappa: no draw happens there, so the score function has nothing to sum
 * although P grows with kappa.
 * D: chosen by rms_adjoint_opts.d_estimator (below).  The pathwise
 * derivative (faces fixed, step ~ sqrt(D), reflections folded, plus the score
 * of the exchange draws) is the whole derivative only while a step's end
 * moves continuously with D.  It jumps at three places: reaching an exchange
 * face (class change), reaching a lattice-edge mirror (origin mirrored), and
 * passing a voxel edge where the face order decides the outcome (exactly two
 * of the two face neighbours and their diagonal neighbour in the voxel's
 * class, as on every staircase of a voxelised curved surface).  A job whose
 * geometry has none of these is "continuous in D". */
enum {
    /* pathwise for jobs continuous in D, central differences with common
     * random numbers for the others (the default of rms_plan_adjoint) */
    RMS_D_AUTO = 0,
    /* central differences with common random numbers for every job: two
     * forward walks per non-null class at D_c (1 +- h), same seeds, the
     * objective contracted per replica; O(h^2) bias, any table.  The walks
     * for class c take only the walkers whose walk can change with D_c: those
     * the tangent pass saw in class c or drawing at an exchange face whose P
     * depends on D_c (every other walker's walk is the same at both points,
     * so it adds nothing to the difference); fd_all_walkers walks them all. */
    RMS_D_CRN_FD = 1,
    /* pathwise plus the boundary terms of the jumps (smoothed perturbation
     * analysis): each boundary event is the derivative of its probability
     * within the step times the difference of the walks continued on either
     * side; one event per walker is kept by weighted reservoir sampling.
     * Unbiased, but each walker carries two continuations and every step
     * searches the events, and the variance is large.  The continuations
     * carry no PGSE sums: in a plan with a PGSE table the D entries of jobs
     * not continuous in D are NaN. */
    RMS_D_BOUNDARY = 2,
    /* pathwise only: NaN for jobs not continuous in D (no extra passes) */
    RMS_D_PATHWISE = 3
};

typedef struct {
    int32_t d_estimator;         /* RMS_D_* */
    int32_t fd_all_walkers;      /* 1 = central differences walk every walker */
    double fd_rel_step;          /* h of RMS_D_AUTO / RMS_D_CRN_FD, 0 = 0.05 */
} rms_adjoint_opts;

typedef struct {
    const double *cot_rec;       /* [timepoints * nval] or NULL */
    const double *cot_sig;       /* [ngrad] or NULL */
    double *grad;                /* [RMS_NPARAM] output */
    double *grad_rep;            /* [replicas * RMS_NPARAM] or NULL */
} rms_job_adj;

/* ---- context (replaces the implicit CUDA context + default stream) ---- */
rms_status rms_create(const rms_device_opts *opts, rms_ctx **out);
rms_status rms_destroy(rms_ctx *ctx);
const char *rms_last_error(const rms_ctx *ctx);
const char *rms_version(void);

/* Upload a geometry; replaces thrust::device_vector<int> d_APix (main.cu:532). */
rms_status rms_add_geometry(rms_ctx *ctx, const rms_geometry *geom, int32_t *geom_id);

/* Device bytes a geometry holds (class lattice and cell words, or brick
 * table and stored bricks), the storage chosen (RMS_STORE_DENSE or
 * RMS_STORE_BRICKS), and how steps through a periodic side are walked:
 * wrap_cells = 1 when more than 1/256 of the non-null voxels lie on a
 * periodic side (the fast path crosses into the voxel on the other side), 0
 * otherwise (the face march takes those steps).  Results are the same
 * physics either way; the two give different roundings.  No reference
 * counterpart (the reference keeps int32 labels, main.cu:532, and mirrors
 * every edge).  NULL outputs are skipped. */
rms_status rms_geometry_info(rms_ctx *ctx, int32_t geom_id, int64_t *device_bytes, int32_t *storage,
                             int32_t *wrap_cells);

/* Derived per-class quantities of a geometry at time step dt, as the
 * reference host computes them (main.cu:429-458): step[c] in voxels,
 * T2s[c] in steps, P[i*RMS_NC_MAX + j]. */
rms_status rms_derive(const rms_geometry *geom, double dt, double *step, double *T2s, double *P);

/* Record row width: 1 + nclass + 6 + 15 + nbvec. */
int32_t rms_nval(int32_t nclass, int32_t nbvec);

/* ---- plans: validate + upload tables once, execute many times ---- */
rms_status rms_plan_create(rms_ctx *ctx, const rms_job *jobs, int32_t njobs, const rms_run_params *params,
                           rms_plan **out);
/* Runs the walk (setup_kernel + propagate + copies, main.cu:498-547). */
rms_status rms_plan_execute(rms_plan *plan, rms_job_out *outs, rms_timing *timing);
rms_status rms_plan_destroy(rms_plan *plan);

/* One-shot convenience: plan_create + plan_execute + plan_destroy. */
rms_status rms_run(rms_ctx *ctx, const rms_job *jobs, int32_t njobs, const rms_run_params *params,
                   rms_job_out *outs, rms_timing *timing);

/* Derivatives of the record and signal sums (no reference counterpart: the
 * reference has no gradients).  Recomputes every trajectory from its seed;
 * D by RMS_D_AUTO. */
rms_status rms_plan_adjoint(rms_plan *plan, const rms_job_adj *adj, rms_timing *timing);
/* The same with the D estimator chosen by opts (NULL = defaults).  timing
 * covers the tangent pass; with central differences walk_ms adds the forward
 * walks at the perturbed D and launches counts them. */
rms_status rms_plan_adjoint_ex(rms_plan *plan, const rms_job_adj *adj, const rms_adjoint_opts *opts,
                               rms_timing *timing);

/* ---- allocation-free run ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same walk and reductions as rms_run for one job,
 * with every device buffer beyond the inputs (the geometry added earlier by
 * rms_add_geometry, the job's tables) and the outputs (host buffers, as for
 * rms_run) carved from one static device array of RMS_SCRATCH_BYTES per GPU,
 * which the GPU holds until the process ends.  No hipMalloc, hipFree or
 * other allocating routine is called, on the device or the host.
 * Limits: one job; one launch (chunk_steps 0 or >= TN, use_graph 0); no
 * final-state outputs; ngrad <= 2000 with at most 512 distinct lobe triples;
 * a run whose buffers exceed RMS_SCRATCH_BYTES is refused with
 * RMS_ERR_INVALID_ARG before anything is launched.
 * Every call on one GPU uses that GPU's array and leaves its contents there;
 * calls that overlap on one GPU (two threads, or two contexts on the same
 * device) race. */
#define RMS_SCRATCH_BYTES (256ull << 20)
rms_status rms_run_noalloc(rms_ctx *ctx, const rms_job *job, const rms_run_params *params, rms_job_out *out,
                           rms_timing *timing);

#ifdef __cplusplus
}
#endif
#endif
