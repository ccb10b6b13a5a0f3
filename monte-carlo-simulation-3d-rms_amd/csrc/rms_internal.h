// rms_internal.h -- kernel launch interface used by the host library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rms_types.h"

namespace rms {

hipError_t launch_walk(const LaunchArgs &a, int prec_f64, int sig, int grid, int block, hipStream_t s);
hipError_t walk_occupancy(int prec_f64, int sig, int block, int *blocks_per_cu);
hipError_t launch_reduce(const ReduceDesc *rd_dev, int32_t nred, int64_t max_nq, int32_t R, hipStream_t s);
hipError_t launch_fill_zero(void *p, int64_t bytes, hipStream_t s);
// Cell words and classes of a label lattice (see rms_types.h); periodic
// axes (bits 0..2) wrap the neighbourhood.
hipError_t launch_build_cells(const void *labels, int32_t label_bytes, const uint8_t *lclass, int32_t nx,
                              int32_t ny, int32_t nz, int32_t periodic, uint64_t *cells, uint8_t *cls,
                              hipStream_t s);
// Brick storage of the cell words (2x2x4 bricks, nbrick = nbx*nby*nbz in
// brick order): pass 1 counts the bricks to store per workgroup of 256
// bricks (counts[ceil(nbrick/256)]); pass 2, given the exclusive prefix sums
// of those counts, writes the brick table and the stored bricks.
hipError_t launch_brick_count(const uint64_t *cells, const uint8_t *cls, int32_t nx, int32_t ny, int32_t nz,
                              int32_t nby, int32_t nbz, int64_t nbrick, uint32_t *counts, hipStream_t s);
hipError_t launch_brick_fill(const uint64_t *cells, const uint8_t *cls, int32_t nx, int32_t ny, int32_t nz,
                             int32_t nby, int32_t nbz, int64_t nbrick, const uint32_t *first, uint32_t *btab,
                             uint64_t *store, hipStream_t s);

// Tangent (adjoint) pass: same walk with forward-mode tangents, contracted
// with the record/signal cotangents per walker.
hipError_t launch_adjoint(const LaunchArgs &a, int prec_f64, int sig, int np, int grid, int block, hipStream_t s);
hipError_t adjoint_occupancy(int prec_f64, int sig, int np, int block, int *blocks_per_cu);
// The walk kernel variant with the boundary terms (sig bit 4) for np tangent
// parameters (rms_kernels_jump.hip).
const void *jump_walk_fn(int prec_f64, int sig, int np);
// The tangent variant without a PGSE table and without the boundary terms
// (sig bits 1 and 4 clear) for np tangent parameters
// (rms_kernels_tangent.hip).
const void *tangent_walk_fn(int prec_f64, int sig, int np);
// The forward and tangent variants for plans with a brick-stored geometry
// (sig bit 5, without bit 4) (rms_kernels_bricks.hip).
const void *brick_walk_fn(int prec_f64, int sig, int np);
// The forward and tangent variants for plans with a geometry whose cell
// words are wrapped across periodic sides (sig bits 3 and 6, without bit 4)
// (rms_kernels_wrap.hip).
const void *wrap_walk_fn(int prec_f64, int sig, int np);
// The forward variant without the fast path's diagnostic counters (sig
// bit 7) for a plan variant sig (with bit 8: every T2 infinite and no
// exchange draw), or null (rms_kernels_nocount.hip).
const void *nocount_walk_fn(int prec_f64, int sig);

// Allocation-free runs: device address of this GPU's scratch array (the
// current device's copy), and the two table kernels that fill it.
hipError_t scratch_address(void **ptr);
// dst[k][c] = src[perm[k]][c] for k < nrow, c < width
hipError_t launch_gather_rows(const double *src, const int32_t *perm, int32_t nrow, int32_t width, double *dst,
                              hipStream_t s);
// out[r][j] = sqrt(bval[j] / ((r Tstep + 1) dt)), the q scale of main.cu:375
hipError_t launch_qscale(const double *bval, int32_t nbvec, int32_t T, int64_t Tstep, double dt, double *out,
                         hipStream_t s);

}  // namespace rms
