// rms_kernels.hip -- forward and tangent variants of the walk kernel
// (rms_walk.h), the ordered reductions, the lattice build, the tables of
// the allocation-free entry point, and the kernel dispatch.
#include "rms_walk.h"

namespace rms {

// Replica sums: rep[rho][q] = sum over batches b = rho, rho+R, ... of
// part[b][q], in increasing b.  One thread per (job, rho, q).
__global__ void reduce_replicas_kernel(const ReduceDesc *rd, int32_t njobs, int32_t R)
{
    const int jid = blockIdx.y;
    if (jid >= njobs) return;
    const ReduceDesc &D = rd[jid];
    const int64_t nq = D.nq;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * R) return;
    const int rho = (int)(t / nq);
    const int64_t q = t - (int64_t)rho * nq;
    double s = 0.0;
    for (int64_t b = rho; b < D.nb; b += R) s = s + D.part[b * nq + q];
    D.rep[(int64_t)rho * nq + q] = s;
}

// Totals: tot[q] = rep[0][q] + rep[1][q] + ... in order (starting from 0).
__global__ void reduce_total_kernel(const ReduceDesc *rd, int32_t njobs, int32_t R)
{
    const int jid = blockIdx.y;
    if (jid >= njobs) return;
    const ReduceDesc &D = rd[jid];
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= D.nq) return;
    double s = 0.0;
    for (int r = 0; r < R; r++) s = s + D.rep[(int64_t)r * D.nq + q];
    D.tot[q] = s;
}

// Zero fill (used instead of hipMemsetAsync so launches stay graph-safe).
__global__ void fill_zero_kernel(unsigned long long *p, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t s = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = 0ull;
}

// Cell words and classes of a label lattice: one thread per voxel reads its
// 27-voxel neighbourhood and writes each neighbour's relation to the voxel's
// class (see rms_types.h).  Along a periodic axis the neighbour beyond a
// lattice side is the voxel on the other side; along the others it is
// outside (CELL_EDGE, a mirror).
__global__ void build_cells_kernel(const void *labels, int32_t label_bytes, const uint8_t *lclass, int32_t nx,
                                   int32_t ny, int32_t nz, int32_t periodic, uint64_t *cells, uint8_t *cls_out)
{
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t lin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lin < n;
         lin += (int64_t)gridDim.x * blockDim.x) {
        const int32_t z = (int32_t)(lin % nz);
        const int64_t t = lin / nz;
        const int32_t y = (int32_t)(t % ny);
        const int32_t x = (int32_t)(t / ny);
        auto class_of = [&](int32_t xx, int32_t yy, int32_t zz) -> int32_t {
            if (periodic & 1) xx = xx < 0 ? nx - 1 : (xx >= nx ? 0 : xx);
            if (periodic & 2) yy = yy < 0 ? ny - 1 : (yy >= ny ? 0 : yy);
            if (periodic & 4) zz = zz < 0 ? nz - 1 : (zz >= nz ? 0 : zz);
            if (xx < 0 || xx >= nx || yy < 0 || yy >= ny || zz < 0 || zz >= nz) return -1;
            const int64_t k = ((int64_t)xx * ny + yy) * nz + zz;
            const uint32_t lab = label_bytes == 1 ? (uint32_t)((const uint8_t *)labels)[k]
                                                  : (uint32_t)((const uint16_t *)labels)[k];
            return lclass ? (int32_t)lclass[lab] : (int32_t)lab;
        };
        const int32_t c0 = class_of(x, y, z);
        uint64_t word = 0;
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    const int32_t cq = class_of(x + dx, y + dy, z + dz);
                    const uint64_t rel =
                        cq == c0 ? CELL_SAME : (cq == 0 ? CELL_NULL : (cq < 0 ? CELL_EDGE : CELL_OTHER));
                    word |= rel << (2 * ((dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)));
                }
        cells[lin] = word;
        cls_out[lin] = (uint8_t)c0;
    }
}

// Brick storage of the cell words.  Brick (bx, by, bz) holds the voxels
// x in {2bx, 2bx+1}, y in {2by, 2by+1}, z in 4bz..4bz+3 (those inside the
// lattice); it needs its words stored when one of its non-null voxels has a
// non-zero word (null voxels never hold a walker, so their words are never
// read).
__device__ __forceinline__ bool brick_needed(const uint64_t *cells, const uint8_t *cls, int32_t nx, int32_t ny,
                                             int32_t nz, int32_t bx, int32_t by, int32_t bz)
{
    bool need = false;
    for (int32_t q = 0; q < 16; q++) {
        const int32_t x = 2 * bx + (q >> 3), y = 2 * by + ((q >> 2) & 1), z = 4 * bz + (q & 3);
        if (x >= nx || y >= ny || z >= nz) continue;
        const int64_t lin = ((int64_t)x * ny + y) * nz + z;
        need = need || (cls[lin] != 0 && cells[lin] != 0ull);
    }
    return need;
}

// Pass 1: number of stored bricks per workgroup of 256 bricks.
__global__ void brick_count_kernel(const uint64_t *cells, const uint8_t *cls, int32_t nx, int32_t ny, int32_t nz,
                                   int32_t nby, int32_t nbz, int64_t nbrick, uint32_t *counts)
{
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool need = false;
    if (b < nbrick) {
        const int32_t bz = (int32_t)(b % nbz);
        const int64_t t = b / nbz;
        need = brick_needed(cells, cls, nx, ny, nz, (int32_t)(t / nby), (int32_t)(t % nby), bz);
    }
    const int n = __syncthreads_count(need ? 1 : 0);
    if (threadIdx.x == 0) counts[blockIdx.x] = (uint32_t)n;
}

// Pass 2: table entries and stored bricks; a workgroup's stored bricks take
// slots first[blockIdx.x], ... in brick order (the exclusive prefix sums of
// pass 1's counts).  Voxels of a stored brick outside the lattice hold 0.
__global__ void brick_fill_kernel(const uint64_t *cells, const uint8_t *cls, int32_t nx, int32_t ny, int32_t nz,
                                  int32_t nby, int32_t nbz, int64_t nbrick, const uint32_t *first, uint32_t *btab,
                                  uint64_t *store)
{
    __shared__ uint32_t wave_n[4];
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int32_t bx = 0, by = 0, bz = 0;
    bool need = false;
    if (b < nbrick) {
        bz = (int32_t)(b % nbz);
        const int64_t t = b / nbz;
        bx = (int32_t)(t / nby);
        by = (int32_t)(t % nby);
        need = brick_needed(cells, cls, nx, ny, nz, bx, by, bz);
    }
    const uint64_t bal = __ballot(need);
    const int lane = (int)(threadIdx.x & 63u), wv = (int)(threadIdx.x >> 6);
    if (lane == 0) wave_n[wv] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t before = first[blockIdx.x];
    for (int k = 0; k < wv; k++) before += wave_n[k];
    before += (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    if (b >= nbrick) return;
    btab[b] = need ? before : BRICK_ZERO;
    if (!need) return;
    for (int32_t q = 0; q < 16; q++) {
        const int32_t x = 2 * bx + (q >> 3), y = 2 * by + ((q >> 2) & 1), z = 4 * bz + (q & 3);
        const bool in = x < nx && y < ny && z < nz;
        store[(uint64_t)before * 16u + (uint32_t)q] = in ? cells[((int64_t)x * ny + y) * nz + z] : 0ull;
    }
}

hipError_t launch_brick_count(const uint64_t *cells, const uint8_t *cls, int32_t nx, int32_t ny, int32_t nz,
                              int32_t nby, int32_t nbz, int64_t nbrick, uint32_t *counts, hipStream_t s)
{
    const int64_t grid = (nbrick + 255) / 256;
    hipLaunchKernelGGL(brick_count_kernel, dim3((unsigned)grid), dim3(256), 0, s, cells, cls, nx, ny, nz, nby, nbz,
                       nbrick, counts);
    return hipGetLastError();
}

hipError_t launch_brick_fill(const uint64_t *cells, const uint8_t *cls, int32_t nx, int32_t ny, int32_t nz,
                             int32_t nby, int32_t nbz, int64_t nbrick, const uint32_t *first, uint32_t *btab,
                             uint64_t *store, hipStream_t s)
{
    const int64_t grid = (nbrick + 255) / 256;
    hipLaunchKernelGGL(brick_fill_kernel, dim3((unsigned)grid), dim3(256), 0, s, cells, cls, nx, ny, nz, nby, nbz,
                       nbrick, first, btab, store);
    return hipGetLastError();
}

// Scratch of the allocation-free entry point (rms_run_noalloc): every device
// buffer of such a run beyond its inputs and outputs.  One copy per GPU, held
// for the life of the process.  In host code `scratch` is only a handle:
// device addresses come from scratch_address(), once per device.
static __device__ __align__(256) unsigned char scratch[RMS_SCRATCH_BYTES];

hipError_t scratch_address(void **ptr) { return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)); }

// Rows of the PGSE table in lobe-triple order: dst[k] = src[perm[k]].
__global__ void gather_rows_kernel(const double *src, const int32_t *perm, int32_t nrow, int32_t width, double *dst)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nrow * width) return;
    const int32_t k = (int32_t)(t / width), c = (int32_t)(t - (int64_t)k * width);
    dst[t] = src[(int64_t)perm[k] * width + c];
}

// q scale of the narrow-pulse signal, sqrt(bval / TD) with TD = (r Tstep + 1) dt
// (main.cu:375, 486); IEEE division and square root, as on the host.
__global__ void qscale_kernel(const double *bval, int32_t nbvec, int32_t T, int64_t Tstep, double dt, double *out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)T * nbvec) return;
    const int64_t r = t / nbvec;
    const int32_t j = (int32_t)(t - r * nbvec);
    const double TD = (double)(r * Tstep + 1) * dt;
    out[t] = __dsqrt_rn(__ddiv_rn(bval[j], TD));
}

// ------------------------------------------------------------------ launch
template <typename R, int NP, int... S>
static const void *walk_fn_sig(int sig, std::integer_sequence<int, S...>)
{
    static const void *const fns[] = {(const void *)walk_kernel<R, S, NP>...};
    return fns[sig & 15];
}

// Tangent passes with a PGSE table (bit 1); those without one are in
// rms_kernels_tangent.hip, those with the boundary terms in
// rms_kernels_jump.hip.
template <typename R, int NP>
static const void *tangent_fn(int sig)
{
    static const void *const fns[] = {(const void *)walk_kernel<R, 2, NP>,  (const void *)walk_kernel<R, 3, NP>,
                                      (const void *)walk_kernel<R, 6, NP>,  (const void *)walk_kernel<R, 7, NP>,
                                      (const void *)walk_kernel<R, 10, NP>, (const void *)walk_kernel<R, 11, NP>,
                                      (const void *)walk_kernel<R, 14, NP>, (const void *)walk_kernel<R, 15, NP>};
    return (sig & 2) ? fns[(sig & 1) | ((sig >> 1) & 6)] : nullptr;
}

template <int NP>
static const void *walk_fn(int prec_f64, int sig)
{
    // sig bit 0: narrow-pulse b-table, bit 1: PGSE table, bit 2: exchange
    // faces on the fast path, bit 3: periodic lattice edges
    if constexpr (NP == 0) {
        if (prec_f64) return walk_fn_sig<double, 0>(sig, std::make_integer_sequence<int, 16>{});
        return walk_fn_sig<float, 0>(sig, std::make_integer_sequence<int, 16>{});
    } else {
        return prec_f64 ? tangent_fn<double, NP>(sig) : tangent_fn<float, NP>(sig);
    }
}

static const void *walk_fn_np(int prec_f64, int sig, int np)
{
    if ((sig & 16) && np > 0) return jump_walk_fn(prec_f64, sig, np);
    if (sig & 64) return wrap_walk_fn(prec_f64, sig, np);
    if (sig & 32) return brick_walk_fn(prec_f64, sig, np);
    if (!(sig & 2) && np > 0) return tangent_walk_fn(prec_f64, sig, np);
    switch (np) {
    case 0: return walk_fn<0>(prec_f64, sig);
    case 1: return walk_fn<1>(prec_f64, sig);
    case 2: return walk_fn<2>(prec_f64, sig);
    default: return walk_fn<3>(prec_f64, sig);
    }
}

static hipError_t launch_fn(const void *fn, const LaunchArgs &a, int grid, int block, hipStream_t s)
{
    if (!fn) return hipErrorInvalidDeviceFunction;  // a variant the host never selects
    void *args[] = {(void *)&a};
    return hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, s);
}

hipError_t launch_walk(const LaunchArgs &a, int prec_f64, int sig, int grid, int block, hipStream_t s)
{
    // a launch that reports no counters takes the variant that does not sum
    // them (and, with sig bit 8, no T2 times or draw bits); the others count
    const void *fn = !a.count ? nocount_walk_fn(prec_f64, sig) : nullptr;
    return launch_fn(fn ? fn : walk_fn_np(prec_f64, sig & 255, 0), a, grid, block, s);
}

hipError_t walk_occupancy(int prec_f64, int sig, int block, int *blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, walk_fn_np(prec_f64, sig, 0), block, 0);
}

hipError_t launch_adjoint(const LaunchArgs &a, int prec_f64, int sig, int np, int grid, int block, hipStream_t s)
{
    return launch_fn(walk_fn_np(prec_f64, sig, np < 1 ? 1 : np), a, grid, block, s);
}

hipError_t adjoint_occupancy(int prec_f64, int sig, int np, int block, int *blocks_per_cu)
{
    const void *fn = walk_fn_np(prec_f64, sig, np < 1 ? 1 : np);
    if (!fn) return hipErrorInvalidDeviceFunction;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, block, 0);
}

hipError_t launch_reduce(const ReduceDesc *rd_dev, int32_t njobs, int64_t max_nq, int32_t R, hipStream_t s)
{
    const int bs = 256;
    dim3 g1((unsigned)((max_nq * R + bs - 1) / bs), (unsigned)njobs);
    hipLaunchKernelGGL(reduce_replicas_kernel, g1, dim3(bs), 0, s, rd_dev, njobs, R);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 g2((unsigned)((max_nq + bs - 1) / bs), (unsigned)njobs);
    hipLaunchKernelGGL(reduce_total_kernel, g2, dim3(bs), 0, s, rd_dev, njobs, R);
    return hipGetLastError();
}

hipError_t launch_build_cells(const void *labels, int32_t label_bytes, const uint8_t *lclass, int32_t nx,
                              int32_t ny, int32_t nz, int32_t periodic, uint64_t *cells, uint8_t *cls, hipStream_t s)
{
    const int64_t n = (int64_t)nx * ny * nz;
    int64_t grid = (n + 255) / 256;
    if (grid > 65536) grid = 65536;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(build_cells_kernel, dim3((unsigned)grid), dim3(256), 0, s, labels, label_bytes, lclass, nx, ny,
                       nz, periodic, cells, cls);
    return hipGetLastError();
}

hipError_t launch_gather_rows(const double *src, const int32_t *perm, int32_t nrow, int32_t width, double *dst,
                              hipStream_t s)
{
    const int64_t n = (int64_t)nrow * width;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, perm, nrow, width,
                       dst);
    return hipGetLastError();
}

hipError_t launch_qscale(const double *bval, int32_t nbvec, int32_t T, int64_t Tstep, double dt, double *out,
                         hipStream_t s)
{
    const int64_t n = (int64_t)T * nbvec;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(qscale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, bval, nbvec, T, Tstep, dt,
                       out);
    return hipGetLastError();
}

hipError_t launch_fill_zero(void *p, int64_t bytes, hipStream_t s)
{
    int64_t n = bytes / 8;
    if (n <= 0) return hipSuccess;
    int grid = (int)((n + 255) / 256);
    if (grid > 65536) grid = 65536;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid), dim3(256), 0, s, (unsigned long long *)p, n);
    return hipGetLastError();
}

}  // namespace rms
