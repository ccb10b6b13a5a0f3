// rms_types.h -- descriptors shared by the host library and the HIP kernels.
//
// One "job" is one geometry with its walkers and acquisition tables; a launch
// runs any number of jobs (the batched form).  Walkers are grouped in batches
// of 64 (one wavefront); batch b of the launch belongs to job batch_job[b].
#pragma once
#include <stdint.h>

namespace rms {

constexpr int NC = 4;          // classes per geometry (reference Nc_max = 3)
constexpr int NITE = 4;        // face iterations per step (main.cu:36)
constexpr int WAVE = 64;       // walkers per batch
constexpr int MAX_BOUNDS = 64; // distinct PGSE lobe boundaries
constexpr uint32_t INIT_TRIALS = 1u << 26;
constexpr int NCOUNT = 6;      // device counters per job (index 5: failed starts)

// Cell word of a voxel (one uint64 per voxel, index (x*ny + y)*nz + z): the
// relation of each voxel of its 3x3x3 neighbourhood to the voxel's class, two
// bits at 2*(9*(qx+1) + 3*(qy+1) + (qz+1)) for the neighbour at offset
// (qx, qy, qz) in {-1,0,1}^3 (the voxel itself is entry 13, bits 26-27, and
// holds CELL_SAME).  A straight move crosses a CELL_SAME face and reflects off
// a CELL_NULL face; a CELL_OTHER face (another non-null class) is an exchange
// face, and a CELL_EDGE face (outside the lattice) a mirror.  Bits 54-63 are
// zero.
constexpr uint64_t CELL_SAME = 0;
constexpr uint64_t CELL_NULL = 1;
constexpr uint64_t CELL_OTHER = 2;
constexpr uint64_t CELL_EDGE = 3;
constexpr uint32_t BRICK_ZERO = 0xFFFFFFFFu;

struct JobDesc {
    // geometry
    const uint64_t *cells;     // cell words, see above (dense, or the stored bricks)
    const uint8_t *cls;        // class of each voxel
    // brick storage (null: dense): entry per 2x2x4 brick, index
    // ((x>>1)*nby + (y>>1))*nbz + (z>>2); BRICK_ZERO when every non-null voxel
    // of the brick has the word 0, else the brick's slot in cells (16 words,
    // local index (x&1)*8 + (y&1)*4 + (z&3))
    const uint32_t *btab;
    int32_t nby, nbz;
    int32_t nx, ny, nz;
    int32_t pg_log2;           // PGSE sums flush at least every 2^pg_log2 steps
    int32_t nclass;
    int32_t nbvec;
    int32_t ngrad;
    int32_t ntriple;           // distinct PGSE lobe triples
    int64_t nwalk;
    int64_t batch0;            // first batch of this job in the launch
    uint32_t key0, key1;       // Philox key = job seed
    uint32_t pcat;             // 2 bits per class pair: 0 reflect, 1 cross, 2 draw
    int32_t nbound;            // PGSE boundaries 1..nbound
    // FP32 exchange faces between classes 1 and 2 (from 1, from 2): -1 reflects
    // without a draw (P = 0); otherwise the smallest high half xhi of the first
    // draw with u01_f32(xhi << 16) >= P, so xhi >= it reflects whatever the low
    // half (65536: never, P = 1 or more than two non-null classes)
    int32_t xh12, xh21;
    int32_t periodic;          // bit k: axis k wraps around instead of mirroring
    int32_t wrapcells;         // the cell words name the voxel across a periodic side (SIG bit 6 walks it)
    double res;
    float step_f[NC], inv_step_f[NC], P_f[NC * NC];
    double step_d[NC], inv_step_d[NC], P_d[NC * NC];
    double T2s[NC];            // T2 in steps (main.cu:433)
    // tangent (adjoint) data; parameter p < 3 is D of class p+1, p = 3 is kappa
    float dstep_f[NC], dinv_f[NC];    // d step_c / d D_c, d (1/step_c) / d D_c
    double dstep_d[NC], dinv_d[NC];
    double rT[NC];                    // 1 / T2s_c
    double invT2[NC];                 // 1 / T2_c (ms)
    double dlogP[NC * NC * 4];        // d log P_ij / d theta
    double dlog1mP[NC * NC * 4];      // d log (1 - P_ij) / d theta
    double *tsnap;                    // [(nbound+1)*NP*3][wstride] tangent prefix sums
    // narrow-pulse b-table (main.cu:374-378)
    const double *bvec;        // [nbvec*3]
    const double *qscale;      // [timepoints*nbvec] sqrt(bval/TD)
    // PGSE (main_PGSE.cu:340-357), window-sum form
    const double *grad;        // [ngrad*6] sorted by lobe triple
    const int32_t *gperm;      // [ngrad] original index of sorted gradient
    const int32_t *triple;     // [ntriple*4]: ka, kb, ke, first gradient; sentinel at end
    const int64_t *bounds;     // [nbound+1], bounds[0] = 0
    double *snap;              // [(nbound+1)*3][wstride] prefix sums
    // outputs
    double *part;              // [nbatch_job][timepoints][nval]
    double *psig;              // [nbatch_job][ngrad]
    unsigned long long *counters; // [NCOUNT]: the 5 of rms_job_out, then walkers whose start failed
    // tangent outputs / inputs
    const double *cot_rec;     // [timepoints][nval] cotangent of the record sums
    const double *cot_sig;     // [ngrad]
    double *pgrad;             // [nbatch_job][NPARAM]
    int64_t wbase;             // first walker slot in the state arrays
    int64_t wid0;              // Philox walker id of the job's first walker
    // walks of a subset of the job's walkers: launch walker k is walker
    // wlist[k] of the job (null: walker k), the list padded to whole batches
    const uint32_t *wlist;
    // tangent pass output per walker (null: not recorded): bit c = visited
    // class c, bit 4 + a*NC + b = drew at an exchange face from class a to b
    uint32_t *wvis;
};

// Walker state saved between chunked launches (structure of arrays).
struct StateBuf {
    int32_t *g;       // [3][ws]
    int32_t *gi;      // [3][ws]
    int32_t *c;       // [ws]
    float *of;        // [6][ws] offsets o, oi (f32 mode)
    double *od;       // [6][ws] offsets o, oi (f64 mode)
    float *tcf;       // [ws]
    double *tcd;      // [ws]
    double *tsum;     // [NC][ws]
    double *ps;       // [3][ws] PGSE running sums (FP64 part)
    int32_t *sg;      // [3][ws] PGSE sums of the voxel index since the last flush
    uint32_t *cnt;    // [5][ws]
    uint32_t *rng;    // [4][ws] cached direction block (f32 mode)
    int64_t ws;       // stride
};

struct FinalOut {
    int32_t *g;       // [3][ws] or nullptr
    double *o;
    int32_t *gi;
    double *oi;
    int32_t *c;
    uint32_t *cnt;    // [5][ws]
};

constexpr int NPARAM = 16;     // tangent parameters (D0[4], T2[4], kappa, ...)

struct LaunchArgs {
    const JobDesc *jobs;
    const int32_t *batch_job;  // [nbatch]
    int64_t nbatch;
    int64_t TN;
    int64_t step0, step1;      // this launch runs steps [step0, step1)
    int64_t Tstep;
    int32_t timepoints;
    int32_t nval;              // per job: 1 + nclass + 21 + nbvec (stride uses job value)
    double dt;
    StateBuf st;
    FinalOut fin;
    int32_t xcd_split;         // 1: partition batches per XCD
    int32_t count;             // 1: sum the per-walker diagnostic counters (crossings, reflections, ...)
};

// One ordered reduction: part[nb][nq] -> rep[R][nq] -> tot[nq].
struct ReduceDesc {
    const double *part;
    double *rep;
    double *tot;
    int64_t nb;
    int64_t nq;
};

}  // namespace rms
