// rms_host.h -- host-side construction and checking of the walk's job
// descriptors: the derived parameters of hpc_code/lib/main.cu:429-458, the
// PGSE lobe tables of main_PGSE.cu:341-345, and the checks a descriptor must
// pass before a kernel variant may read it.
//
// Plain C++ (no HIP calls) and free of heap allocation, so the same code
// builds every plan, the allocation-free entry point, and a CPU test harness.
// Functions return nullptr on success and a static message otherwise.
#pragma once
#include <stdint.h>

#include "../../include/rms_amd.h"
#include "rms_types.h"

namespace rms {

// Kernel variant bits (the SIG template argument of walk_kernel).
enum : int { SIG_NARROW = 1, SIG_PGSE = 2, SIG_EXCH = 4, SIG_PERIODIC = 8, SIG_JUMP = 16, SIG_BRICKS = 32,
             SIG_WRAP = 64, SIG_NOCOUNT = 128, SIG_LEAN = 256 };

// The adjoint variant of a plan: the boundary terms of the D derivative
// (exchange faces, lattice-edge mirrors, voxel edges) whenever no job has a
// PGSE table (the continuations carry no PGSE sums).  The jump variants
// handle exchange faces on the fast path.
inline int adjoint_sig(int sig)
{
    return (sig & SIG_PGSE) == 0 ? (sig | SIG_JUMP | SIG_EXCH) : sig;
}

// Geometry properties that add boundary terms to the D derivative (bits of
// geometry_flags):
//   GEOM_EDGE_ORDER  some non-null voxel has, for a pair of axes (a, b) and a
//                    quadrant (sa, sb), exactly two of its face neighbours
//                    q + sa e_a, q + sb e_b and its diagonal neighbour
//                    q + sa e_a + sb e_b in its own class: a step through
//                    that voxel edge then ends in a different voxel
//                    depending on which face the march takes first, as on
//                    every staircase of a voxelised curved surface;
//   GEOM_MIRROR      some non-null voxel touches a lattice face of an axis
//                    that is not periodic: reaching it mirrors the origin.
// Neighbours across a periodic axis wrap; outside the lattice otherwise is
// not the voxel's class.  Without either property (and without exchange)
// the pathwise D derivative is the whole derivative.
enum : uint32_t { GEOM_EDGE_ORDER = 1u, GEOM_MIRROR = 2u };
uint32_t geometry_flags(const rms_geometry &g);

// How steps through a periodic side are walked.  True: the cell words of
// voxels on the side name the voxels on the other side of the lattice and
// the fast path takes those steps (kernel variant SIG_WRAP: three compares
// per step); false: the words mark the side as outside the lattice and the
// march takes them (no cost while walkers stay away from the sides).  True
// when the lattice has a periodic axis and more than 1/256 of its non-null
// voxels lie on a periodic side.
bool wrap_cells_rule(const rms_geometry &g);

// Walkers whose walk can change when D_c moves to D_c (1 +- h) (the central
// differences of the adjoint's D entries): those that visit class c, and those
// that draw at an exchange face whose P changes with D_c, since the draw may
// then decide the other way.  Every other walker takes each step with the
// same length and decides each face with the same P, so its walk is the same
// at both points and adds nothing to the difference.  mask = bits of the
// tangent pass's per-walker marks (bit c, bit 4 + a*NC + b for draws from
// class a into class b) that make a walker one of them.  False when a face
// whose P changes crosses or reflects without a draw at one of the three
// points (walkers that meet it leave no mark): then every walker is walked.
bool fd_walker_mask(const rms_geometry &g, double dt, int c, double h, uint32_t &mask);


// Steps must stay below one voxel by this margin: the straight move ends at
// e = fma(s, v, o) with o in [0, 1], and e must round below 2 in FP32 and FP64.
constexpr double STEP_MAX = 1.0 - 0x1p-20;

// Allocation-free runs: capacity of the per-device scratch array and the
// table sizes whose host work space lives on the stack.
constexpr int64_t SCRATCH_BYTES = (int64_t)RMS_SCRATCH_BYTES;
constexpr int32_t NOALLOC_MAX_GRAD = 2000;   // reference Ngrad_max (main_PGSE.cu:37)
constexpr int32_t NOALLOC_MAX_TRIPLE = 512;  // distinct PGSE lobe triples

struct Derived {
    double step[NC], inv_step[NC], T2s[NC], P[NC * NC];
    int Pcat[NC * NC];
};

// main.cu:429-458 (+ the kappa extension).
const char *derive(const rms_geometry &g, double dt, Derived &d);

// Tangent constants of the adjoint pass into D.
void derive_tangent(const rms_geometry &g, double dt, const Derived &d, JobDesc &D);

// True when walkers can move between two different non-null classes.
bool has_exchange(const rms_geometry &g, const Derived &d);

// True when some pair of non-null classes is impermeable (P < 1e-6) only
// because of its kappa (the reference rule would let walkers cross).
bool kappa_closes_pair(const rms_geometry &g, const Derived &d);

// A walk that needs no T2 times and no exchange draws: every non-null
// class has an infinite T2 (class 0, never entered, a positive one), so the
// weight exp(-sum t_c / T2_c) is exp(-0) = 1 exactly, and no pair of
// non-null classes draws (P is 0 or 1 for each).  Such launches may take the
// forward variants that keep neither (SIG_LEAN).
bool lean_walk(const rms_geometry &g, const Derived &d);

int32_t exchange_threshold(int32_t nclass, int cat, float P);
int32_t pgse_flush_log2(int32_t nx, int32_t ny, int32_t nz);

// Validation of the run parameters shared by every entry point.
const char *check_params(const rms_run_params &P);

// Per-job scalars of a descriptor: lattice size, physics, seed, walker ids.
// Device table pointers are left null (the caller points them at its
// buffers or at sentinels).  sig_bits receives the variant bits the job
// needs (SIG_EXCH, SIG_PERIODIC; the table bits come from the job's tables).
const char *fill_job(const rms_geometry &g, const rms_job &J, const rms_run_params &P, JobDesc &D, int &sig_bits,
                     bool &exchange);

// The physics fields of a descriptor (nclass, steps, exchange
// probabilities and their categories and FP32 thresholds, tangent
// constants) from the geometry's per-class D, T2 and kappa at time step dt,
// with the checks of fill_job on them; sig_bits and exchange as there.  The
// central differences of the D derivative rebuild them at a perturbed D.
const char *job_physics(const rms_geometry &g, double dt, JobDesc &D, int &sig_bits, bool &exchange);

// PGSE lobe tables (window-sum form), allocation free.
//   grad [ngrad*6] rows (Delta, delta, |g|, gx, gy, gz)
//   perm [ngrad]: perm[k] = original index of the k-th sorted gradient
//   triple [4*(ntriple+1)]: (ka, kb, ke, first sorted gradient) per distinct
//     lobe triple, then a sentinel (0, 0, 0, ngrad); capacity max_triple+1
//   bounds [nbound+1]: bounds[0] = 0, then the distinct nonzero boundaries;
//     capacity MAX_BOUNDS+1
//   keys [ngrad] work space
// Lobe 1 of gradient g is steps [0, bounds[ka]), lobe 2 [bounds[kb], bounds[ke]),
// from the reference's double comparisons (i+1)*dt < delta and
// Delta <= (i+1)*dt < Delta+delta (main_PGSE.cu:341-345).
const char *pgse_tables(const double *grad, int32_t ngrad, double dt, int64_t TN, int32_t *perm, int32_t *triple,
                        int32_t max_triple, int64_t *bounds, int32_t &nbound, int32_t &ntriple, uint64_t *keys);

// Every table that walk_kernel<., sig, .> can read for this job is present
// (the job's own or a sentinel) and consistent with its sizes.
const char *check_desc(const JobDesc &D, int sig);

// Byte offsets of one allocation-free run's buffers in the scratch array,
// each 256-byte aligned.  A table the job does not have keeps a small
// segment that holds its zero sentinel.
struct ScratchLayout {
    int64_t job, red, batch_job, counters, part, rep, tot, psig, srep, stot, snap, grad_raw, grad, perm, triple,
        bounds, bval, qscale, bvec, end;
};
const char *scratch_layout(int64_t nbatch, int32_t timepoints, int32_t nval, int32_t replicas, int32_t ngrad,
                           int32_t nbound, int32_t ntriple, int32_t nbvec, int64_t cap, ScratchLayout &L);

}  // namespace rms
