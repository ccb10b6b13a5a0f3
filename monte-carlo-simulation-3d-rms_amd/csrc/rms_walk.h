// rms_walk.h -- the RMS random-walk kernel for gfx950 (MI355X), included by
// rms_kernels.hip (forward and tangent variants) and rms_kernels_jump.hip
// (tangent variants with the boundary terms of the D derivative).
//
// One lane = one walker (hpc_code/lib/main.cu:64-70), but the grid is
// persistent: each wavefront walks batches of 64 walkers taken from a batch
// list that can span many geometries (jobs).  Walker state lives in VGPRs for
// the whole chunk of steps; the lattice is read as one 64-bit cell word per
// voxel (how each of its 26 neighbours relates to its class) through L2 /
// Infinity Cache, plus a class byte per voxel for the march and the start.
//
// A step whose straight move meets only same-class faces and faces next to
// null voxels is the straight move folded back at the null faces, decided by
// lookups in the cell word along the sequence of faces (the fast path); the
// other lanes run the face march in the same step.
//
// Per-record moments are summed across the wavefront with a fixed butterfly
// (the fixed part of a row by a reduce-scatter that adds, per value, the same
// lane pairs in the same order) and written as per-batch partials; a second
// kernel sums the partials in a
// fixed order, so results are bit-reproducible and identical to the host
// restatement.
//
// Reference anchors:
//   propagate (narrow pulse)  main.cu:63-384
//   propagate (PGSE)          main_PGSE.cu:65-410
//   atomAdd CAS loop          main.cu:40-56   -> butterfly + ordered partials
//   setup_kernel              main.cu:58-61   -> counter-based Philox, no state
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

#include "../../include/rms_amd.h"
#include "rms_device.h"
#include "rms_internal.h"
#include "rms_turn_table.h"
#include "rms_types.h"

namespace rms {

// Axis-indexed triples with named fields: code indexes them only through
// compile-time axes (for3) or value selects (pick), so walker state never
// becomes an addressable array and stays in VGPRs.
template <typename T>
struct V3 {
    T x, y, z;
};
template <int K, typename T>
__device__ __forceinline__ T &at(V3<T> &v)
{
    if constexpr (K == 0)
        return v.x;
    else if constexpr (K == 1)
        return v.y;
    else
        return v.z;
}
template <typename T>
__device__ __forceinline__ T pick(T a, T b, T c, int k)
{
    return k == 0 ? a : (k == 1 ? b : c);
}
template <int K>
struct Ax {
    static constexpr int value = K;
};
template <typename F>
__device__ __forceinline__ void for3(F &&f)
{
    f(Ax<0>{});
    f(Ax<1>{});
    f(Ax<2>{});
}

template <typename R>
struct Walker {
    V3<int32_t> g;   // voxel
    V3<R> o;         // offset inside the voxel, [0,1]
    V3<int32_t> gi;  // origin voxel (mirror-unfolded)
    V3<R> oi;        // origin offset
    // PGSE running sums of the displacement in voxels, kept split between
    // flushes: ps holds the flushed sum plus sum(o) of the later steps, sg the
    // sum(g) of those steps; a flush adds sg - n (gi + oi)
    V3<double> ps;
    V3<int32_t> sg;
    int32_t c;       // class
    int32_t lin;     // linear voxel index of g
    R tcur;          // steps spent in class c since the last flush
    R sc, isc;       // step of class c (voxels) and its reciprocal
    int32_t hc;      // reflect-sure threshold of the first exchange draw's high half from class c
    double ts0, ts1, ts2, ts3;  // steps per class
    uint32_t ncross, nrefl, nmirror, ntrunc, ndraw;
};

// Tangent state of the adjoint pass (NP pathwise parameters: D of classes
// 1..NP): position tangents, tangent of the T2 exponent E = sum t_c / T2s_c,
// score sums of the exchange draws, PGSE prefix sums of res * Jo.
template <typename R, int NP>
struct Tan {
    V3<R> Jo[NP > 0 ? NP : 1];
    double dE[NP > 0 ? NP : 1];
    double S[NP > 0 ? NP : 1];
    double Sk;
    V3<double> PJ[NP > 0 ? NP : 1];
    uint32_t vis;  // classes visited (bit c), exchange draws from class a to b (bit 4 + a*NC + b)
};

// Per-job constants held in registers for the whole batch.
template <typename R>
struct JobLocal {
    const uint64_t *cells;
    const uint8_t *cls;
    const uint32_t *btab;    // brick table (null: dense cell words)
    int32_t nby, nbz;
    int32_t nx, ny, nz, sx;  // sx = ny*nz
    int32_t nclass;
    uint32_t key0, key1, pcat;
    R st0, st1, st2, st3;    // step per class (voxels)
    R is0, is1, is2, is3;    // 1/step
    double T0, T1, T2, T3;   // T2 in steps
    double res;
    const R *P;              // permeation probabilities [NC*NC]
    R P12, P21;              // P between classes 1 and 2
    int32_t xh12, xh21;      // reflect-sure thresholds of the first draw's high half
    int32_t periodic;        // bit k: axis k wraps around (else the edge is a mirror)
    int32_t exch_fast;       // two non-null classes: a CELL_OTHER face is the other one
    // adjoint constants
    R ds1, ds2, ds3;         // d step_c / d D_c, classes 1..3
    R di1, di2, di3;         // d (1/step_c) / d D_c
    double rT0, rT1, rT2, rT3;  // 1 / T2s_c
    double sd0, sd1, sd2, sd3;  // step per class in FP64 (boundary events)
    double dsd1, dsd2, dsd3;    // d step_c / d D_c in FP64
};

template <typename R>
__device__ __forceinline__ R sel4(int c, R a0, R a1, R a2, R a3)
{
    return c == 0 ? a0 : (c == 1 ? a1 : (c == 2 ? a2 : a3));
}

template <typename R>
__device__ __forceinline__ JobLocal<R> job_local(const JobDesc &J)
{
    JobLocal<R> L;
    L.cells = J.cells;
    L.cls = J.cls;
    L.btab = J.btab;
    L.nby = J.nby;
    L.nbz = J.nbz;
    L.nx = J.nx;
    L.ny = J.ny;
    L.nz = J.nz;
    L.sx = J.ny * J.nz;
    L.nclass = J.nclass;
    L.key0 = J.key0;
    L.key1 = J.key1;
    L.pcat = J.pcat;
    if constexpr (sizeof(R) == 4) {
        L.st0 = J.step_f[0];
        L.st1 = J.step_f[1];
        L.st2 = J.step_f[2];
        L.st3 = J.step_f[3];
        L.is0 = J.inv_step_f[0];
        L.is1 = J.inv_step_f[1];
        L.is2 = J.inv_step_f[2];
        L.is3 = J.inv_step_f[3];
        L.P = J.P_f;
    } else {
        L.st0 = J.step_d[0];
        L.st1 = J.step_d[1];
        L.st2 = J.step_d[2];
        L.st3 = J.step_d[3];
        L.is0 = J.inv_step_d[0];
        L.is1 = J.inv_step_d[1];
        L.is2 = J.inv_step_d[2];
        L.is3 = J.inv_step_d[3];
        L.P = J.P_d;
    }
    L.P12 = L.P[1 * NC + 2];
    L.P21 = L.P[2 * NC + 1];
    L.xh12 = J.xh12;
    L.xh21 = J.xh21;
    L.periodic = J.periodic;
    L.exch_fast = J.nclass <= 3;
    L.T0 = J.T2s[0];
    L.T1 = J.T2s[1];
    L.T2 = J.T2s[2];
    L.T3 = J.T2s[3];
    L.res = J.res;
    if constexpr (sizeof(R) == 4) {
        L.ds1 = J.dstep_f[1];
        L.ds2 = J.dstep_f[2];
        L.ds3 = J.dstep_f[3];
        L.di1 = J.dinv_f[1];
        L.di2 = J.dinv_f[2];
        L.di3 = J.dinv_f[3];
    } else {
        L.ds1 = J.dstep_d[1];
        L.ds2 = J.dstep_d[2];
        L.ds3 = J.dstep_d[3];
        L.di1 = J.dinv_d[1];
        L.di2 = J.dinv_d[2];
        L.di3 = J.dinv_d[3];
    }
    L.rT0 = J.rT[0];
    L.rT1 = J.rT[1];
    L.rT2 = J.rT[2];
    L.rT3 = J.rT[3];
    L.sd0 = J.step_d[0];
    L.sd1 = J.step_d[1];
    L.sd2 = J.step_d[2];
    L.sd3 = J.step_d[3];
    L.dsd1 = J.dstep_d[1];
    L.dsd2 = J.dstep_d[2];
    L.dsd3 = J.dstep_d[3];
    return L;
}

template <typename R>
struct PiRef;
template <>
struct PiRef<float> {
    static constexpr float v = 1.0f;
};
template <>
struct PiRef<double> {
    static constexpr double v = 0.9999999988573335;  // Pi = 3.14159265 (main.cu:34) over pi
};

template <typename R>
__device__ __forceinline__ int32_t class_at(const JobLocal<R> &L, int32_t lin)
{
    return (int32_t)L.cls[lin];
}

// Cell word of voxel (x, y, z), linear index lin.  Dense storage reads
// cells[lin]; brick storage reads the brick's table entry and, unless the
// brick holds only zero words, the word in the brick's 128-byte slot.  The
// kernel variants with BR (plans with a brick-stored geometry) choose per
// job, with a branch uniform over the wavefront; the others read dense words
// only.
template <typename R, bool BR>
__device__ __forceinline__ uint64_t cell_at(const JobLocal<R> &L, int32_t x, int32_t y, int32_t z, int32_t lin)
{
    if constexpr (BR) {
        if (L.btab != nullptr) {
            const uint32_t e = L.btab[((x >> 1) * L.nby + (y >> 1)) * L.nbz + (z >> 2)];
            if (e == BRICK_ZERO) return 0ull;
            return L.cells[(uint64_t)e * 16u + (uint32_t)(((x & 1) << 3) | ((y & 1) << 2) | (z & 3))];
        }
    }
    return L.cells[lin];
}

template <typename R, bool BR>
__device__ __forceinline__ uint64_t cell_of(const JobLocal<R> &L, const Walker<R> &w)
{
    return cell_at<R, BR>(L, w.g.x, w.g.y, w.g.z, w.lin);
}

// Move the time spent in the current class into its FP64 total.  Every class
// total is rewritten (x + 0.0 == x for the non-negative totals) so no store
// goes through a data-dependent address.
template <typename R>
__device__ __forceinline__ void flush(Walker<R> &w)
{
    const double t = (double)w.tcur;
    w.ts0 = w.ts0 + (w.c == 0 ? t : 0.0);
    w.ts1 = w.ts1 + (w.c == 1 ? t : 0.0);
    w.ts2 = w.ts2 + (w.c == 2 ? t : 0.0);
    w.ts3 = w.ts3 + (w.c == 3 ? t : 0.0);
    w.tcur = (R)0;
}

// The origin of one axis as a double, the term the PGSE sums subtract.
template <typename R>
__device__ __forceinline__ double origin(int32_t gi, R oi)
{
    return (double)gi + (double)oi;
}

// PGSE sums after a mirror or a periodic wrap moved the origin of one axis
// from xo to xn: the npg terms since the last flush were taken with xo, the
// flush subtracts the origin at flush time, so npg (xn - xo) is added now.
__device__ __forceinline__ double mirror_ps(double ps, int32_t npg, double xo, double xn)
{
    return ps + (double)npg * (xn - xo);
}

// Rejection-sampled start voxel, main.cu:117-132 (uniforms over (0,1)).
// Returns false when INIT_TRIALS trials found no non-null voxel; the walker
// then sits in voxel (0, 0, 0) and the host reports the run as failed.
template <typename R>
__device__ __forceinline__ bool init_walker(Walker<R> &w, const JobLocal<R> &L, uint64_t wid)
{
    bool found = false;
    const double nx = (double)L.nx, ny = (double)L.ny, nz = (double)L.nz;
    w.g = {0, 0, 0};
    w.o = {(R)0.5, (R)0.5, (R)0.5};
    w.c = 0;
    w.lin = 0;
    for (uint32_t n = 0; n < INIT_TRIALS; n++) {
        U4 a = rng_block(L.key0, L.key1, n, STREAM_INIT_XY, wid);
        U4 b = rng_block(L.key0, L.key1, n, STREAM_INIT_Z, wid);
        double x0 = u01_f64(a.x, a.y) * nx;
        double x1 = u01_f64(a.z, a.w) * ny;
        double x2 = u01_f64(b.x, b.y) * nz;
        int32_t g0 = (int32_t)__builtin_floor(x0);
        int32_t g1 = (int32_t)__builtin_floor(x1);
        int32_t g2 = (int32_t)__builtin_floor(x2);
        g0 = g0 > L.nx - 1 ? L.nx - 1 : g0;
        g1 = g1 > L.ny - 1 ? L.ny - 1 : g1;
        g2 = g2 > L.nz - 1 ? L.nz - 1 : g2;
        int32_t lin = g0 * L.sx + g1 * L.nz + g2;
        int32_t c = class_at(L, lin);
        if (c != 0) {
            w.g = {g0, g1, g2};
            w.o = {(R)(x0 - (double)g0), (R)(x1 - (double)g1), (R)(x2 - (double)g2)};
            w.c = c;
            w.lin = lin;
            found = true;
            break;
        }
    }
    w.gi = w.g;
    w.oi = w.o;
    return found;
}

// Exchange draw j (0..3) of step i.  f32 layout: draw 0 is u01_f32 of the
// word (xhi << 16) | (x >> 16), where xhi holds the 16 low bits of the step's
// direction words that the direction does not use (draw_dir) and x is word x
// of block (i, 1); draws 1..3 take words y, z, w of that block.  The high half
// alone settles u < P except when P falls inside its 2^-16 interval, so the
// fast path computes the block only then or for a second draw.  f64 layout:
// half j&1 of block (i, 1 + j/2).  The block is recomputed for every draw
// (draws happen only at probabilistic membranes), which keeps no block in
// registers across the march.
__device__ __forceinline__ uint32_t exch_word0(uint32_t xhi, uint32_t x) { return (xhi << 16) | (x >> 16); }

template <typename R>
__device__ __forceinline__ R exch_uniform(const JobLocal<R> &L, uint64_t wid, int64_t i, int j, uint32_t xhi)
{
    if constexpr (sizeof(R) == 4) {
        const U4 blk = rng_block(L.key0, L.key1, (uint32_t)i, STREAM_EXCH_A, wid);
        const uint32_t word = j == 0 ? exch_word0(xhi, blk.x) : (j == 1 ? blk.y : (j == 2 ? blk.z : blk.w));
        return u01_f32(word);
    } else {
        const U4 blk = rng_block(L.key0, L.key1, (uint32_t)i, STREAM_EXCH_A + (uint32_t)(j >> 1), wid);
        return (j & 1) ? u01_f64(blk.z, blk.w) : u01_f64(blk.x, blk.y);
    }
}

// Direction of step i, main.cu:140-151.  f32 layout: one Philox block serves
// steps 2m (words x, y) and 2m+1 (words z, w); the azimuth's sin/cos come from
// the 256-angle table tab.  The polar uniform uses bits 9..31 of the first
// word and the azimuth bits 8..31 of the second; xhi returns the low bytes of
// the two words (first word's in bits 8..15), the high half of the step's
// first exchange draw.  f64 layout: xhi = 0 (unused).
template <typename R, bool XHI = true>
__device__ __forceinline__ V3<R> draw_dir(const JobLocal<R> &L, uint64_t wid, int64_t i, bool fresh_block,
                                          uint32_t &b0, uint32_t &b1, const float2 *tab, uint32_t &xhi)
{
    R u1, u2;
    if constexpr (sizeof(R) == 4) {
        // (b0, b1) carry words z, w of block i/2 from step 2m to step 2m+1
        uint32_t w0 = b0, w1 = b1;
        if (fresh_block) {
            const U4 r = rng_block(L.key0, L.key1, (uint32_t)(i >> 1), STREAM_DIR, wid);
            const bool odd = (i & 1) != 0;
            w0 = odd ? r.z : r.x;
            w1 = odd ? r.w : r.y;
            b0 = r.z;
            b1 = r.w;
        }
        // (launches whose geometries draw nothing leave the draw's bits out)
        xhi = XHI ? ((w0 & 0xFFu) << 8) | (w1 & 0xFFu) : 0u;
        u1 = u01_f32(w0);
        const R cth = (R)1 - (R)2 * u1;
        const R sth = (R)2 * sqrt_rn_mid(u1 * ((R)1 - u1));
        R sp, cp;
        sincos_turns_table(w1, tab, sp, cp);
        return V3<R>{sth * cp, sth * sp, cth};
    } else {
        xhi = 0u;
        U4 r = rng_block(L.key0, L.key1, (uint32_t)i, STREAM_DIR, wid);
        u1 = u01_f64(r.x, r.y);
        u2 = u01_f64(r.z, r.w);
        const R cth = (R)1 - (R)2 * u1;
        const R sth = (R)2 * sqrt(u1 * ((R)1 - u1));
        R sp, cp;
        sincos_turns(u2 * PiRef<R>::v, sp, cp);
        return V3<R>{sth * cp, sth * sp, cth};
    }
}

// Exchange faces met by a straight move (slots with CELL_OTHER in rels; the
// other non-null class when a geometry has two, main.cu:272-289): P = 0
// reflects without a draw, P = 1 crosses into the other class (march);
// otherwise the k-th exchange face of the step takes draw k (the march's
// numbering) and reflects when u >= P.  Returns false when the step needs the
// march (which replays the same draws); otherwise every exchange face
// reflects, ndraw counts the draws and the tangent pass adds their scores.
// With f32 draws the first draw's high half xhi bounds u to
// [u01_f32(xhi << 16), u01_f32((xhi << 16) | 0xFFFF)]; the Philox block is
// computed only when P lies inside those bounds or a second face draws.
template <typename R, int NP>
__device__ __forceinline__ bool exchange_faces(const Walker<R> &w, const JobLocal<R> &L, const JobDesc &J,
                                               uint32_t rels, uint64_t wid, int32_t i, uint32_t xhi,
                                               Tan<R, NP> &T, uint32_t &ndraw)
{
    if (!L.exch_fast) return false;
    const int32_t pidx = w.c * NC + (3 - w.c);
    const uint32_t cat = (L.pcat >> (2 * pidx)) & 3u;
    if (cat == 1u) return false;
    if (cat != 2u) return true;
    const R P = w.c == 1 ? L.P12 : L.P21;
    bool cross = false;
    if constexpr (sizeof(R) == 4) {
        ndraw = (uint32_t)__builtin_popcount((rels >> 1) & ~rels & 0x15u);  // CELL_OTHER slots
        const bool lo_cross = u01_f32((xhi << 16) | 0xFFFFu) < P;  // u < P for every low half
        const bool hi_refl = u01_f32(xhi << 16) >= P;              // u >= P for every low half
        cross = lo_cross;
        if ((!lo_cross && !hi_refl) || ndraw > 1u) {
            const U4 blk = rng_block(L.key0, L.key1, (uint32_t)i, STREAM_EXCH_A, wid);
            cross = cross || u01_f32(exch_word0(xhi, blk.x)) < P;
            cross = cross || (ndraw > 1u && u01_f32(blk.y) < P);
            cross = cross || (ndraw > 2u && u01_f32(blk.z) < P);
        }
    } else {
#pragma unroll
        for (int s = 0; s < 3; s++) {
            if (((rels >> (2 * s)) & 3u) == 2u) {
                cross = cross || exch_uniform<R>(L, wid, i, (int)ndraw, 0u) < P;
                ndraw++;
            }
        }
    }
    if (cross) return false;
    if constexpr (NP > 0) {
        if (ndraw > 0u) T.vis |= 1u << (4 + pidx);
        // scores of the reflecting Bernoulli(P) draws, one add per draw
        const double *sc4 = J.dlog1mP + 4 * pidx;
#pragma unroll
        for (uint32_t k = 0; k < 3u; k++) {
            if (k < ndraw) {
#pragma unroll
                for (int p = 0; p < NP; p++) T.S[p] = T.S[p] + sc4[p];
                T.Sk = T.Sk + sc4[3];
            }
        }
    }
    return true;
}

// Straight-move fast path.  The straight step ends at e = fma(s, v, o) and crosses
// the faces of the axes where floor(e) = c_k is +-1 (e outside [0,1)).  The crossed faces
// are met in increasing d_k / |v_k| (d_k = distance to the face), compared
// without a division: face a comes before face b (a < b) when
// d_a |v_b| <= d_b |v_a|.  Met in that order, a face whose neighbour has the
// voxel's class is crossed without a draw or a turn, one next to a null voxel
// reflects without a draw (main.cu:274-289; P = 0 next to label 0,
// :450-451), a lattice edge is a mirror, and with EXCH an exchange face is
// decided by exchange_faces.  The cell word holds every neighbour's relation
// to the voxel's class, so the step is one lookup per face along the sequence
// and the straight move folded back at the reflecting faces.  Returns false
// when the step needs the march (a crossing into another class, or three
// comparisons that are not transitive: a near tie).
template <typename R, int NP, bool EXCH, bool PER, bool WRAP, bool CNT, bool T2W>
__device__ __forceinline__ bool fast_step(Walker<R> &w, const JobLocal<R> &L, const JobDesc &J, uint64_t cell,
                                          const V3<R> &v, uint64_t wid, int32_t i, uint32_t xhi, int32_t npg,
                                          Tan<R, NP> &T)
{
    const R sc = w.sc;
    const R ex = fma(sc, v.x, w.o.x);
    const R ey = fma(sc, v.y, w.o.y);
    const R ez = fma(sc, v.z, w.o.z);
    // e is in (-1, 2): the face crossed along an axis is floor(e) in {-1, 0, 1}
    const R fx = floor(ex), fy = floor(ey), fz = floor(ez);
    const int32_t cx = (int32_t)fx, cy = (int32_t)fy, cz = (int32_t)fz;
    const R dx = cx > 0 ? (R)1 - w.o.x : w.o.x;
    const R dy = cy > 0 ? (R)1 - w.o.y : w.o.y;
    const R dz = cz > 0 ? (R)1 - w.o.z : w.o.z;
    const R avx = fabs(v.x), avy = fabs(v.y), avz = fabs(v.z);
    // face a before face b: a crossed, and b not crossed or met later
    const bool xy = cx != 0 && (cy == 0 || dx * avy <= dy * avx);
    const bool xz = cx != 0 && (cz == 0 || dx * avz <= dz * avx);
    const bool yz = cy != 0 && (cz == 0 || dy * avz <= dz * avy);
    // ranks in the sequence; a non-transitive triple gives all three rank 1
    const int32_t kx = (int32_t)!xy + (int32_t)!xz;
    const int32_t ky = (int32_t)xy + (int32_t)!yz;
    // neighbour index 9*(qx+1) + 3*(qy+1) + (qz+1), the voxel itself is 13;
    // an axis not crossed adds nothing and reads relation 0 of the current voxel
    // (bit offsets 2*index are tracked)
    const int32_t ox = 18 * cx, oy = 6 * cy, oz = 2 * cz;
    int32_t cur = 26;
    uint32_t rels = 0u;  // relation met in slot s, bits 2s..2s+1
#pragma unroll
    for (int s = 0; s < 3; s++) {
        const int32_t q = cur + (kx == s ? ox : (ky == s ? oy : oz));
        const uint32_t rel = (uint32_t)(cell >> q) & 3u;
        rels |= rel << (2 * s);
        cur = rel == 0u ? q : cur;
    }
    uint32_t ndraw = 0u, nedge = 0u, refl = rels;  // reflected slots, bits 2s
    bool rare = (rels & 0x2Au) != 0u || kx == ky;
    if constexpr (EXCH && NP == 0 && sizeof(R) == 4) {
        // one exchange face whose draw reflects whatever its low half (xhi at
        // or above the class's threshold), or that reflects without a draw, is
        // folded like a null face; other cases take the branch below
        const uint32_t om = (rels >> 1) & ~rels & 0x15u;  // CELL_OTHER slots
        const int32_t hc = w.hc;
        const bool easy = om == 0u || ((om & (om - 1u)) == 0u && (int32_t)xhi >= hc);
        rare = (rels & (rels >> 1) & 0x15u) != 0u || kx == ky || !easy;
        refl = (rels | (rels >> 1)) & 0x15u;
        ndraw = (om != 0u && hc >= 0) ? 1u : 0u;
    }
    if constexpr (WRAP) {
        // with wrapped cell words a straight move out of the lattice goes
        // through a periodic side (elsewhere the relation is CELL_EDGE,
        // already rare)
        rare = rare || (uint32_t)(w.g.x + cx) >= (uint32_t)L.nx || (uint32_t)(w.g.y + cy) >= (uint32_t)L.ny ||
               (uint32_t)(w.g.z + cz) >= (uint32_t)L.nz;
    }
    if (rare) {
        // a near tie goes to the march.  A lattice edge is a mirror: the
        // walker reflects and its origin is mirrored (main.cu:246-265,
        // 302-310).
        if (kx == ky) return false;
        const uint32_t edge = rels & (rels >> 1) & 0x15u;  // CELL_EDGE slots
        const int32_t kz = 3 - kx - ky;
        if (PER && edge != 0u) {
            // cell words that mark a periodic side as outside the lattice
            // (wrap_cells_rule false): the march crosses into the voxel on
            // the other side (decided before any state changes).  Wrapped
            // cell words name that voxel instead and never reach here.
            const uint32_t eaxes = ((edge >> (2 * kx)) & 1u) | (((edge >> (2 * ky)) & 1u) << 1) |
                                   (((edge >> (2 * kz)) & 1u) << 2);
            if ((eaxes & (uint32_t)L.periodic) != 0u) return false;
        }
        if (((rels >> 1) & ~rels & 0x15u) != 0u) {  // CELL_OTHER slots
            if constexpr (EXCH) {
                if (!exchange_faces<R, NP>(w, L, J, rels, wid, i, xhi, T, ndraw)) return false;
            } else {
                return false;
            }
        }
        if (edge != 0u) {
            if (((edge >> (2 * kx)) & 1u) != 0u) {
                const double xo = origin(w.gi.x, w.oi.x);
                w.gi.x = cx < 0 ? -w.gi.x : 2 * L.nx - w.gi.x;
                w.oi.x = -w.oi.x;
                w.ps.x = mirror_ps(w.ps.x, npg, xo, origin(w.gi.x, w.oi.x));
            }
            if (((edge >> (2 * ky)) & 1u) != 0u) {
                const double xo = origin(w.gi.y, w.oi.y);
                w.gi.y = cy < 0 ? -w.gi.y : 2 * L.ny - w.gi.y;
                w.oi.y = -w.oi.y;
                w.ps.y = mirror_ps(w.ps.y, npg, xo, origin(w.gi.y, w.oi.y));
            }
            if (((edge >> (2 * kz)) & 1u) != 0u) {
                const double xo = origin(w.gi.z, w.oi.z);
                w.gi.z = cz < 0 ? -w.gi.z : 2 * L.nz - w.gi.z;
                w.oi.z = -w.oi.z;
                w.ps.z = mirror_ps(w.ps.z, npg, xo, origin(w.gi.z, w.oi.z));
            }
            nedge = (uint32_t)__builtin_popcount(edge);
        }
        refl = (rels | (rels >> 1)) & 0x15u;
        if constexpr (WRAP) {
            // a face crossed out of the lattice is on a periodic side: the
            // cell word named the voxel on the other side, where the walker
            // goes; voxel index and origin move by one lattice length (the
            // march's rule) before the common update adds the crossing
            const int32_t ks[3] = {kx, ky, kz};
            for3([&](auto K) __attribute__((always_inline)) {
                constexpr int k = decltype(K)::value;
                const int32_t ck = k == 0 ? cx : (k == 1 ? cy : cz);
                const int32_t Nk = k == 0 ? L.nx : (k == 1 ? L.ny : L.nz);
                const bool crossed = ((refl >> (2 * ks[k])) & 1u) == 0u;
                if (crossed && (uint32_t)(at<k>(w.g) + ck) >= (uint32_t)Nk) {
                    const int32_t sh = ck < 0 ? Nk : -Nk;
                    const double xo = origin(at<k>(w.gi), at<k>(w.oi));
                    at<k>(w.g) = at<k>(w.g) + sh;
                    at<k>(w.gi) = at<k>(w.gi) + sh;
                    at<k>(w.ps) = mirror_ps(at<k>(w.ps), npg, xo, origin(at<k>(w.gi), at<k>(w.oi)));
                    w.lin += sh * (k == 0 ? L.sx : (k == 1 ? L.nz : 1));
                }
            });
        }
    }
    // every face met is now crossed (CELL_SAME, or axis not crossed) or
    // reflected (CELL_NULL, a reflecting exchange face, a mirror); the ranks
    // are a permutation of 0, 1, 2
    const int32_t kz = 3 - kx - ky;
    const bool rx = ((refl >> (2 * kx)) & 1u) != 0u;
    const bool ry = ((refl >> (2 * ky)) & 1u) != 0u;
    const bool rz = ((refl >> (2 * kz)) & 1u) != 0u;
    if constexpr (NP > 0) {
#pragma unroll
        for (int p = 0; p < NP; p++) {
            const R dsp = p == 0 ? L.ds1 : (p == 1 ? L.ds2 : L.ds3);
            const R dsc = (w.c == p + 1) ? dsp : (R)0;
            T.Jo[p].x = T.Jo[p].x + dsc * v.x;
            T.Jo[p].y = T.Jo[p].y + dsc * v.y;
            T.Jo[p].z = T.Jo[p].z + dsc * v.z;
            T.Jo[p].x = rx ? -T.Jo[p].x : T.Jo[p].x;
            T.Jo[p].y = ry ? -T.Jo[p].y : T.Jo[p].y;
            T.Jo[p].z = rz ? -T.Jo[p].z : T.Jo[p].z;
        }
    }
    const int32_t gx = rx ? 0 : cx, gy = ry ? 0 : cy, gz = rz ? 0 : cz;
    // y = e - c is the offset past the face; a reflection folds it back to 1 - y
    const R px = ex - fx, py = ey - fy, pz = ez - fz;
    w.o.x = rx ? (R)1 - px : px;
    w.o.y = ry ? (R)1 - py : py;
    w.o.z = rz ? (R)1 - pz : pz;
    w.g.x += gx;
    w.g.y += gy;
    w.g.z += gz;
    w.lin += (gx == 0 ? 0 : (gx > 0 ? L.sx : -L.sx)) + (gy == 0 ? 0 : (gy > 0 ? L.nz : -L.nz)) + gz;
    if constexpr (CNT) {
        // diagnostic sums (about 3 % of the step; variants without them skip it)
        const uint32_t nref = (uint32_t)__builtin_popcount(refl);
        w.nrefl += nref - nedge;
        w.nmirror += nedge;
        w.ndraw += ndraw;
        w.ncross += (uint32_t)(__mul24(cx, cx) + __mul24(cy, cy) + __mul24(cz, cz)) - nref;
    }
    if constexpr (T2W) w.tcur = w.tcur + (R)1;  // time in the class for its T2 weight
    return true;
}

// Face march of step i with direction v (main.cu:153-300).
//
// With a step shorter than one voxel each axis meets at most one face per
// step, so the three face distances t_k (main.cu:199-219) are computed once,
// visited in increasing (t, axis) order -- the order of the reference's strict
// `tt < tmin` scan -- and the walker advances by t_e - t_{e-1}.  A face is
// reached when f * step >= t_e - t_{e-1} (main.cu:231); the remaining
// fraction finishes the step (main.cu:237-243).  The three event slots are
// unrolled and predicated, so a wavefront never loops on its slowest lane.
// Exchange draws (0 < P < 1 between two non-null classes) exist only in
// geometries with two non-null classes, which the host runs with EXCH set;
// without it the draw is not compiled.
// Face distance in units of path (main.cu:199-219): faces farther than one
// unit of path get 2, max(0, .) clamps.
template <typename R>
__device__ __forceinline__ R face_t(R ok, R vk, R ivk)
{
    const R tmp = vk > (R)0 ? ((R)1 - ok) : (vk < (R)0 ? -ok : (R)2);
    R q = tmp * ivk;
    q = (q > (R)0) ? q : (R)0;
    return (fabs(tmp) > fabs(vk)) ? (R)2 : q;
}

template <typename R, int NP, bool EXCH, bool PER>
__device__ __forceinline__ void march_core(Walker<R> &w, const JobLocal<R> &L, const JobDesc &J, uint64_t wid,
                                           int64_t i, V3<R> &v, const V3<R> &iv, R t0, R t1, R t2, int h0, int h1,
                                           int h2, uint32_t xhi, int32_t npg, Tan<R, NP> &T);

template <typename R, int NP, bool EXCH, bool PER>
__device__ __forceinline__ void march(Walker<R> &w, const JobLocal<R> &L, const JobDesc &J, uint64_t wid,
                                      int64_t i, V3<R> v, uint32_t xhi, int32_t npg, Tan<R, NP> &T)
{
    V3<R> iv = {(R)1 / v.x, (R)1 / v.y, (R)1 / v.z};
    // face distances in units of path, main.cu:199-219
    V3<R> t;
    for3([&](auto K) __attribute__((always_inline)) {
        constexpr int k = decltype(K)::value;
        at<k>(t) = face_t(at<k>(w.o), at<k>(v), at<k>(iv));
    });
    // stable order by (t, axis)
    R t0 = t.x, t1 = t.y, t2 = t.z;
    int h0 = 0, h1 = 1, h2 = 2;
    {
        bool sw = t1 < t0;
        R tt = sw ? t1 : t0;
        t1 = sw ? t0 : t1;
        t0 = tt;
        int hh = sw ? h1 : h0;
        h1 = sw ? h0 : h1;
        h0 = hh;
        sw = t2 < t1;
        tt = sw ? t2 : t1;
        t2 = sw ? t1 : t2;
        t1 = tt;
        hh = sw ? h2 : h1;
        h2 = sw ? h1 : h2;
        h1 = hh;
        sw = t1 < t0;
        tt = sw ? t1 : t0;
        t1 = sw ? t0 : t1;
        t0 = tt;
        hh = sw ? h1 : h0;
        h1 = sw ? h0 : h1;
        h0 = hh;
    }
    march_core<R, NP, EXCH, PER>(w, L, J, wid, i, v, iv, t0, t1, t2, h0, h1, h2, xhi, npg, T);
}

// The march from face distances t0 <= t1 <= t2 (axes h0, h1, h2: the order
// the march visits them); v ends as the direction after the step's
// reflections and mirrors.
template <typename R, int NP, bool EXCH, bool PER>
__device__ __forceinline__ void march_core(Walker<R> &w, const JobLocal<R> &L, const JobDesc &J, uint64_t wid,
                                           int64_t i, V3<R> &v, const V3<R> &iv, R t0, R t1, R t2, int h0, int h1,
                                           int h2, uint32_t xhi, int32_t npg, Tan<R, NP> &T)
{
    R f = (R)1;
    R trav = (R)0;
    int nd = 0;
    bool live = true;
    uint32_t wraps = 0u;  // periodic crossings: bit 2k (down) or 2k+1 (up) for axis k
    R df[NP > 0 ? NP : 1];
#pragma unroll
    for (int p = 0; p < (NP > 0 ? NP : 1); p++) df[p] = (R)0;
#pragma unroll
    for (int e = 0; e < 3; e++) {
        const R te = e == 0 ? t0 : (e == 1 ? t1 : t2);
        const int hit = e == 0 ? h0 : (e == 1 ? h1 : h2);
        const R dte = te - trav;
        live = live && (f > (R)0) && (f * w.sc >= dte);
        if (live) {
            // reach the face, main.cu:231-236
            const R fr = dte * w.isc;
            if constexpr (NP > 0) {
                // tangents: the face is fixed, d(dte) = -d o_hit / v_hit
                const R ivh = pick(iv.x, iv.y, iv.z, hit);
                const double rTc = sel4(w.c, L.rT0, L.rT1, L.rT2, L.rT3);
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const R Jh = pick(T.Jo[p].x, T.Jo[p].y, T.Jo[p].z, hit);
                    const R dtm = -(Jh)*ivh;
                    const R dip = p == 0 ? L.di1 : (p == 1 ? L.di2 : L.di3);
                    const R dis = (w.c == p + 1) ? dip : (R)0;
                    const R dfr = dtm * w.isc + dte * dis;
                    df[p] = df[p] - dfr;
                    T.dE[p] = T.dE[p] + rTc * (double)dfr;
                    for3([&](auto K) __attribute__((always_inline)) {
                        constexpr int k = decltype(K)::value;
                        at<k>(T.Jo[p]) = (k == hit) ? (R)0 : at<k>(T.Jo[p]) + dtm * at<k>(v);
                    });
                }
            }
            f = f - fr;
            w.tcur = w.tcur + fr;
            for3([&](auto K) __attribute__((always_inline)) {
                constexpr int k = decltype(K)::value;
                const R moved = at<k>(w.o) + dte * at<k>(v);
                const R snapv = at<k>(v) > (R)0 ? (R)1 : (R)0;
                at<k>(w.o) = (k == hit) ? snapv : moved;
            });
            trav = te;
            const R vh = pick(v.x, v.y, v.z, hit);
            const int32_t gh = pick(w.g.x, w.g.y, w.g.z, hit);
            const int32_t Nh = pick(L.nx, L.ny, L.nz, hit);
            const bool up = vh > (R)0;
            const int32_t nk = gh + (up ? 1 : -1);
            bool flip;
            const bool out = nk < 0 || nk >= Nh;
            // a periodic axis continues in the voxel on the other side of the
            // lattice, with the origin moved by one lattice length
            const bool wrap = PER && out && ((L.periodic >> hit) & 1) != 0;
            if (out && !wrap) {
                // lattice-edge mirror, main.cu:249-265; origin unfold main.cu:302-310
                for3([&](auto K) __attribute__((always_inline)) {
                    constexpr int k = decltype(K)::value;
                    const bool h = (k == hit);
                    const double xo = origin(at<k>(w.gi), at<k>(w.oi));
                    const int32_t gm = (nk < 0) ? -at<k>(w.gi) : 2 * Nh - at<k>(w.gi);
                    at<k>(w.gi) = h ? gm : at<k>(w.gi);
                    at<k>(w.oi) = h ? -at<k>(w.oi) : at<k>(w.oi);
                    if (h) at<k>(w.ps) = mirror_ps(at<k>(w.ps), npg, xo, origin(at<k>(w.gi), at<k>(w.oi)));
                });
                w.nmirror++;
                flip = true;
            } else {
                const int32_t Sh = pick(L.sx, L.nz, 1, hit);
                const int32_t nkw = wrap ? (nk < 0 ? Nh - 1 : 0) : nk;
                const int32_t lin_n = w.lin + (wrap ? (nkw - gh) * Sh : (up ? Sh : -Sh));
                const int32_t cn = class_at(L, lin_n);  // label gather, main.cu:272
                const int pidx = w.c * NC + cn;
                const int cat = (int)((L.pcat >> (2 * pidx)) & 3u);
                bool cross = (cat == 1);
                if (EXCH && cat == 2) {
                    const R u = exch_uniform<R>(L, wid, i, nd, xhi);
                    nd++;
                    cross = u < L.P[pidx];
                    if constexpr (NP > 0) {
                        T.vis |= 1u << (4 + pidx);
                        // score of the Bernoulli(P) draw
                        const double *sc4 = cross ? (J.dlogP + 4 * pidx) : (J.dlog1mP + 4 * pidx);
#pragma unroll
                        for (int p = 0; p < NP; p++) T.S[p] = T.S[p] + sc4[p];
                        T.Sk = T.Sk + sc4[3];
                    }
                }
                if (cross) {
                    for3([&](auto K) __attribute__((always_inline)) {
                        constexpr int k = decltype(K)::value;
                        const bool h = (k == hit);
                        at<k>(w.g) = h ? nkw : at<k>(w.g);
                        at<k>(w.o) = h ? (up ? (R)0 : (R)1) : at<k>(w.o);
                    });
                    wraps |= wrap ? (1u << (2 * hit + (up ? 1 : 0))) : 0u;
                    w.lin = lin_n;
                    if (cn != w.c) {
                        flush(w);
                        w.c = cn;
                        if constexpr (NP > 0) T.vis |= 1u << cn;
                        w.sc = sel4(cn, L.st0, L.st1, L.st2, L.st3);
                        w.isc = sel4(cn, L.is0, L.is1, L.is2, L.is3);
                        w.hc = cn == 1 ? L.xh12 : L.xh21;
                    }
                    w.ncross++;
                } else {
                    w.nrefl++;
                }
                flip = !cross;
            }
            if (flip) {
                for3([&](auto K) __attribute__((always_inline)) {
                    constexpr int k = decltype(K)::value;
                    const bool h = (k == hit);
                    at<k>(v) = h ? -at<k>(v) : at<k>(v);
                });
            }
        }
    }
    if (PER && wraps != 0u) {
        // the origin follows each periodic crossing by one lattice length
        for3([&](auto K) __attribute__((always_inline)) {
            constexpr int k = decltype(K)::value;
            const uint32_t wk = (wraps >> (2 * k)) & 3u;
            if (wk != 0u) {
                const int32_t Nk = k == 0 ? L.nx : (k == 1 ? L.ny : L.nz);
                const double xo = origin(at<k>(w.gi), at<k>(w.oi));
                at<k>(w.gi) = at<k>(w.gi) + (wk == 2u ? -Nk : Nk);
                at<k>(w.ps) = mirror_ps(at<k>(w.ps), npg, xo, origin(at<k>(w.gi), at<k>(w.oi)));
            }
        });
    }
    if (f > (R)0) {
        // finish inside the voxel, main.cu:237-243
        const R ds = w.sc * f;
        if constexpr (NP > 0) {
            const double rTc = sel4(w.c, L.rT0, L.rT1, L.rT2, L.rT3);
#pragma unroll
            for (int p = 0; p < NP; p++) {
                const R dsp = p == 0 ? L.ds1 : (p == 1 ? L.ds2 : L.ds3);
                const R dsc = (w.c == p + 1) ? dsp : (R)0;
                const R dds = dsc * f + w.sc * df[p];
                for3([&](auto K) __attribute__((always_inline)) {
                    constexpr int k = decltype(K)::value;
                    at<k>(T.Jo[p]) = at<k>(T.Jo[p]) + dds * at<k>(v);
                });
                T.dE[p] = T.dE[p] + rTc * (double)df[p];
            }
        }
        for3([&](auto K) __attribute__((always_inline)) {
            constexpr int k = decltype(K)::value;
            at<k>(w.o) = at<k>(w.o) + ds * at<k>(v);
        });
        w.tcur = w.tcur + f;
    }
    w.ndraw += (uint32_t)nd;
}

// ---------------------------------------------------------------------------
// D derivative across exchange faces (adjoint pass).  A step reaches the
// face of axis k on side sg when sg v_k s >= dist (dist: distance to the
// face, voxels); given the azimuth of v about axis k, v_k is uniform on
// [-1, 1], so P(reach) = (1 - c)/2 with c = dist / s and its derivative is
// w = -(1/2) dc/dtheta, dc/dtheta = (d dist/dtheta - c ds/dtheta) / s.  At the
// boundary direction (v_k = sg c, the other two components rescaled) the step
// ends exactly on the face, and the two outcomes differ only when the face is
// an exchange face whose draw crosses: the walker then sits on the face in
// the other class.  Each such event adds w (F_cross - F_stay) to
// dE[F]/dtheta, F_cross / F_stay being the objective along the walks
// continued from the face in either class with the walker's own random
// stream.  One event per walker is kept by weighted reservoir sampling; at
// every record the kept pair adds W (w_p / omega) (F_cross - F_stay).  The
// host restatement (or_jump_event, or_jump_step) performs the same double
// operations in the same order.

// Reservoir uniform of candidate face m = 2k + (sg > 0) of step i.
template <typename R>
__device__ __forceinline__ double jump_uniform(const JobLocal<R> &L, uint64_t wid, int32_t i, int m)
{
    const U4 r = rng_block(L.key0, L.key1, (uint32_t)i, STREAM_JUMP + (uint32_t)(m >> 1), wid);
    return (m & 1) ? u01_f64(r.z, r.w) : u01_f64(r.x, r.y);
}

// The boundary path of face (k, sg) from the walker's state before step i,
// with step s (voxels) and the step's direction v.  True when it ends on an
// exchange face whose draw crosses; A0 then stays in the walker's class on
// the face, A1 has crossed it.  Exchange faces met before the end take their
// draws as the walk would; one that crosses drops the event (second order).
template <typename R, bool PER>
__device__ bool jump_event(const Walker<R> &w, const JobLocal<R> &L, const double v[3], int k, int sg, double s,
                           double dist, uint64_t wid, int32_t i, uint32_t xhi, Walker<R> &A0, Walker<R> &A1)
{
    const int a0 = k == 0 ? 1 : 0, a1 = k == 2 ? 1 : 2;
    const double rho2 = v[a0] * v[a0] + v[a1] * v[a1];
    if (!(rho2 > 0.0)) return false;
    const double cs = dist / s;
    const double sc = __dsqrt_rn(__ddiv_rn(1.0 - cs * cs, rho2));
    double u[3], p[3], tf[3];
    u[k] = (double)sg * cs;
    u[a0] = v[a0] * sc;
    u[a1] = v[a1] * sc;
    p[0] = (double)w.o.x;
    p[1] = (double)w.o.y;
    p[2] = (double)w.o.z;
    int32_t g[3] = {w.g.x, w.g.y, w.g.z};
    int32_t gi[3] = {w.gi.x, w.gi.y, w.gi.z};
    R oi[3] = {w.oi.x, w.oi.y, w.oi.z};
    const int32_t N[3] = {L.nx, L.ny, L.nz};
    const int32_t ci = w.c;
    int ord[2], n = 0;
    for (int e = 0; e < 2; e++) {
        const int a = e == 0 ? a0 : a1;
        tf[a] = u[a] > 0.0 ? (1.0 - p[a]) / u[a] : (u[a] < 0.0 ? p[a] / -u[a] : 2.0);
        if (tf[a] < s) ord[n++] = a;
    }
    if (n == 2 && tf[ord[1]] < tf[ord[0]]) {
        const int x = ord[0];
        ord[0] = ord[1];
        ord[1] = x;
    }
    double trav = 0.0;
    int nd = 0;
    for (int e = 0; e < n; e++) {
        const int h = ord[e];
        const double dte = tf[h] - trav;
        for (int q = 0; q < 3; q++)
            if (q != h) p[q] = p[q] + dte * u[q];
        p[h] = u[h] > 0.0 ? 1.0 : 0.0;
        trav = tf[h];
        int32_t nk = g[h] + (u[h] > 0.0 ? 1 : -1);
        int32_t wrap = 0;
        if (PER && (nk < 0 || nk >= N[h]) && ((L.periodic >> h) & 1)) {
            wrap = nk < 0 ? N[h] : -N[h];
            nk = nk < 0 ? N[h] - 1 : 0;
        }
        if (nk < 0 || nk >= N[h]) {
            // lattice-edge mirror
            u[h] = -u[h];
            gi[h] = nk < 0 ? -gi[h] : 2 * N[h] - gi[h];
            oi[h] = -oi[h];
            continue;
        }
        int32_t q3[3] = {g[0], g[1], g[2]};
        q3[h] = nk;
        const int32_t cn = class_at(L, q3[0] * L.sx + q3[1] * L.nz + q3[2]);
        bool cross;
        if (cn == ci) {
            cross = true;
        } else if (cn == 0) {
            cross = false;
        } else {
            const int pidx = ci * NC + cn;
            const uint32_t cat = (L.pcat >> (2 * pidx)) & 3u;
            if (cat == 1u) return false;
            if (cat == 2u) {
                const R ud = exch_uniform<R>(L, wid, i, nd, xhi);
                nd++;
                if (ud < L.P[pidx]) return false;
            }
            cross = false;
        }
        if (cross) {
            g[h] = nk;
            gi[h] += wrap;
            p[h] = u[h] > 0.0 ? 0.0 : 1.0;
        } else {
            u[h] = -u[h];
        }
    }
    // the end of the step lies on the face (k, sg)
    for (int q = 0; q < 3; q++)
        if (q != k) p[q] = p[q] + (s - trav) * u[q];
    p[k] = sg > 0 ? 1.0 : 0.0;
    int32_t nk = g[k] + sg;
    int32_t wrap = 0;
    if (PER && (nk < 0 || nk >= N[k]) && ((L.periodic >> k) & 1)) {
        wrap = nk < 0 ? N[k] : -N[k];
        nk = nk < 0 ? N[k] - 1 : 0;
    }
    if (nk < 0 || nk >= N[k]) {
        // a lattice-edge mirror: reaching it mirrors the origin
        // (main.cu:302-310), which flips the sign of that displacement
        // component; both continuations walk from the face, A1 with the
        // mirrored origin
        A0 = w;
        A0.g = {g[0], g[1], g[2]};
        A0.o = {(R)p[0], (R)p[1], (R)p[2]};
        A0.gi = {gi[0], gi[1], gi[2]};
        A0.oi = {oi[0], oi[1], oi[2]};
        A0.lin = g[0] * L.sx + g[1] * L.nz + g[2];
        A0.tcur = w.tcur + (R)1;
        A1 = A0;
        const int32_t gm = nk < 0 ? -gi[k] : 2 * N[k] - gi[k];
        if (k == 0) {
            A1.gi.x = gm;
            A1.oi.x = -A1.oi.x;
        } else if (k == 1) {
            A1.gi.y = gm;
            A1.oi.y = -A1.oi.y;
        } else {
            A1.gi.z = gm;
            A1.oi.z = -A1.oi.z;
        }
        return true;
    }
    int32_t q3[3] = {g[0], g[1], g[2]};
    q3[k] = nk;
    const int32_t cq = class_at(L, q3[0] * L.sx + q3[1] * L.nz + q3[2]);
    if (cq == ci || cq == 0) return false;
    const int pidx = ci * NC + cq;
    const uint32_t cat = (L.pcat >> (2 * pidx)) & 3u;
    if (cat == 0u) return false;
    if (cat == 2u) {
        const R ud = exch_uniform<R>(L, wid, i, nd, xhi);
        if (!(ud < L.P[pidx])) return false;
    }
    A0 = w;
    A0.g = {g[0], g[1], g[2]};
    A0.o = {(R)p[0], (R)p[1], (R)p[2]};
    A0.gi = {gi[0], gi[1], gi[2]};
    A0.oi = {oi[0], oi[1], oi[2]};
    A0.lin = g[0] * L.sx + g[1] * L.nz + g[2];
    A0.tcur = w.tcur + (R)1;
    A1 = A0;
    if (k == 0) {
        A1.g.x = nk;
        A1.gi.x += wrap;
        A1.o.x = sg > 0 ? (R)0 : (R)1;
    } else if (k == 1) {
        A1.g.y = nk;
        A1.gi.y += wrap;
        A1.o.y = sg > 0 ? (R)0 : (R)1;
    } else {
        A1.g.z = nk;
        A1.gi.z += wrap;
        A1.o.z = sg > 0 ? (R)0 : (R)1;
    }
    A1.lin = A1.g.x * L.sx + A1.g.y * L.nz + A1.g.z;
    flush(A1);
    A1.c = cq;
    A1.sc = sel4(cq, L.st0, L.st1, L.st2, L.st3);
    A1.isc = sel4(cq, L.is0, L.is1, L.is2, L.is3);
    A1.hc = cq == 1 ? L.xh12 : L.xh21;
    return true;
}

// ---------------------------------------------------------------------------
// The two continuations of a walker's kept event live in LDS between their
// uses (stepping, records, replacement), not in registers: the adjoint pass
// with boundary terms otherwise holds three walkers, the tangents and the
// event search at once.  One 32-bit word per field and lane, field-major
// ([field][continuation][thread], 256 threads at most): lane t touches word
// t of each row, no bank conflicts.  lin and the step of the class are
// recomputed on load.
template <typename R>
struct Cont {
    static constexpr int RW = sizeof(R) / 4;               // words per R
    static constexpr int WORDS = 3 + 3 * RW + 3 + 3 * RW + 1 + RW + 8 + 2;  // g o gi oi c tcur ts cell
    static constexpr int ROW = 2 * 256;
};

template <typename R>
__device__ __forceinline__ void put_r(uint32_t *&q, R v)
{
    if constexpr (sizeof(R) == 4) {
        q[0] = __float_as_uint(v);
    } else {
        const uint64_t b = (uint64_t)__double_as_longlong(v);
        q[0] = (uint32_t)b;
        q[Cont<R>::ROW] = (uint32_t)(b >> 32);
    }
    q += Cont<R>::RW * Cont<R>::ROW;
}

template <typename R>
__device__ __forceinline__ R get_r(const uint32_t *&q)
{
    R v;
    if constexpr (sizeof(R) == 4) {
        v = __uint_as_float(q[0]);
    } else {
        v = __longlong_as_double((long long)(((uint64_t)q[Cont<R>::ROW] << 32) | (uint64_t)q[0]));
    }
    q += Cont<R>::RW * Cont<R>::ROW;
    return v;
}

template <typename R>
__device__ __forceinline__ void store_cont(uint32_t *lds, int m, int t, const Walker<R> &a, uint64_t cell)
{
    constexpr int ROW = Cont<R>::ROW;
    uint32_t *q = lds + m * 256 + t;
    auto put_i = [&](uint32_t v) { q[0] = v; q += ROW; };
    put_i((uint32_t)a.g.x); put_i((uint32_t)a.g.y); put_i((uint32_t)a.g.z);
    put_r<R>(q, a.o.x); put_r<R>(q, a.o.y); put_r<R>(q, a.o.z);
    put_i((uint32_t)a.gi.x); put_i((uint32_t)a.gi.y); put_i((uint32_t)a.gi.z);
    put_r<R>(q, a.oi.x); put_r<R>(q, a.oi.y); put_r<R>(q, a.oi.z);
    put_i((uint32_t)a.c);
    put_r<R>(q, a.tcur);
    put_r<double>(q, a.ts0); put_r<double>(q, a.ts1); put_r<double>(q, a.ts2); put_r<double>(q, a.ts3);
    put_i((uint32_t)cell); put_i((uint32_t)(cell >> 32));
}

template <typename R>
__device__ __forceinline__ void load_cont(const uint32_t *lds, int m, int t, const JobLocal<R> &L, Walker<R> &a,
                                          uint64_t &cell)
{
    constexpr int ROW = Cont<R>::ROW;
    const uint32_t *q = lds + m * 256 + t;
    auto get_i = [&]() { const uint32_t v = q[0]; q += ROW; return v; };
    a.g.x = (int32_t)get_i(); a.g.y = (int32_t)get_i(); a.g.z = (int32_t)get_i();
    a.o.x = get_r<R>(q); a.o.y = get_r<R>(q); a.o.z = get_r<R>(q);
    a.gi.x = (int32_t)get_i(); a.gi.y = (int32_t)get_i(); a.gi.z = (int32_t)get_i();
    a.oi.x = get_r<R>(q); a.oi.y = get_r<R>(q); a.oi.z = get_r<R>(q);
    a.c = (int32_t)get_i();
    a.tcur = get_r<R>(q);
    a.ts0 = get_r<double>(q); a.ts1 = get_r<double>(q); a.ts2 = get_r<double>(q); a.ts3 = get_r<double>(q);
    const uint32_t lo = get_i(), hi = get_i();
    cell = ((uint64_t)hi << 32) | lo;
    a.lin = a.g.x * L.sx + a.g.y * L.nz + a.g.z;
    a.sc = sel4(a.c, L.st0, L.st1, L.st2, L.st3);
    a.isc = sel4(a.c, L.is0, L.is1, L.is2, L.is3);
    a.hc = a.c == 1 ? L.xh12 : L.xh21;
    a.ps = {0.0, 0.0, 0.0};
    a.sg = {0, 0, 0};
    a.ncross = a.nrefl = a.nmirror = a.ntrunc = a.ndraw = 0u;
}

// ---------------------------------------------------------------------------
// Voxel-edge events (adjoint pass).  Where a step passes a voxel edge (the
// faces of axes a and b met together) its outcome depends on which face the
// march takes first (the nearer, ties to the lower axis, main.cu:199-219)
// whenever the diagonal neighbour differs from the face neighbours, as along
// every staircase of a voxelised surface.  Given v_c (c the third axis) the
// direction's angle psi in the (a, b) plane is uniform; in the quadrant
// (sa, sb) face a comes first when psi < psi_t = atan(d_b / d_a) (d_a, d_b:
// distances to the two faces), so P(a first) moves at the rate
// (1/2 pi) dpsi_t/dtheta = (1/2 pi) (d_a d_b' - d_b d_a') / (d_a^2 + d_b^2).
// At the tie direction the path reaches the edge after t_e = r / rho
// (r = |(d_a, d_b)|, rho = sqrt(1 - v_c^2)); when t_e < s the step is marched
// with either order forced, and where the two ends differ (another class, or
// reflections off other axes) the event adds
// w (F(a first) - F(b first)) into the walker's reservoir, as an exchange-face
// event does.  Candidate m = 4 pair + quadrant draws its reservoir uniform
// from stream STREAM_CORNER + m/2.  The host restatement (or_corner_step)
// performs the same operations in the same order.

template <typename R>
__device__ __forceinline__ double corner_uniform(const JobLocal<R> &L, uint64_t wid, int32_t i, int m)
{
    const U4 r = rng_block(L.key0, L.key1, (uint32_t)i, STREAM_CORNER + (uint32_t)(m >> 1), wid);
    return (m & 1) ? u01_f64(r.z, r.w) : u01_f64(r.x, r.y);
}

// The two orders end in the same physical state when they end in the same
// class having reflected (or mirrored) off the same axes: the paths are then
// the same broken line, and the end states differ at most in how a point on
// a face is written (voxel g with o = 1 or voxel g + 1 with o = 0).
template <typename R>
__device__ __forceinline__ bool same_outcome(const Walker<R> &x, const V3<R> &vx, const Walker<R> &y,
                                             const V3<R> &vy)
{
    return x.c == y.c && vx.x == vy.x && vx.y == vy.y && vx.z == vy.z;
}

// The voxel-edge events of step i from the walker's state w before it (its
// D tangents T, direction vr, step s in voxels); updates the reservoir.
template <typename R, int NP, bool EXCH, bool PER, bool BR>
__device__ void corner_events(const Walker<R> &w, const Tan<R, NP> &T, const JobLocal<R> &L, const JobDesc &J,
                              const V3<R> &vr, double s, uint64_t wid, int32_t i, uint32_t xhi, bool &has,
                              double &Wj, double *rho, uint32_t *cont, int t)
{
    Tan<R, 0> T0;
    for (int pr = 0; pr < 3; pr++) {
        // edges of the face pairs (a, b) = (x, y), (x, z), (y, z); c the third axis
        const int a = pr == 2 ? 1 : 0, b = pr == 0 ? 1 : 2, c = 2 - pr;
        const double vc = (double)pick(vr.x, vr.y, vr.z, c);
        const double rhoc = __dsqrt_rn(1.0 - vc * vc);
        const double oa = (double)pick(w.o.x, w.o.y, w.o.z, a), ob = (double)pick(w.o.x, w.o.y, w.o.z, b);
        for (int qd = 0; qd < 4; qd++) {
            const int sa = (qd & 1) ? 1 : -1, sb = (qd & 2) ? 1 : -1;
            const double da = sa > 0 ? 1.0 - oa : oa;
            const double db = sb > 0 ? 1.0 - ob : ob;
            const double r2 = da * da + db * db;
            const double r = __dsqrt_rn(r2);
            if (!(r < s * rhoc) || !(r2 > 0.0)) continue;
            // the tie direction and the distance to the edge
            const R va = (R)((double)sa * (rhoc * (da / r)));
            const R vb = (R)((double)sb * (rhoc * (db / r)));
            const R vcr = pick(vr.x, vr.y, vr.z, c);
            V3<R> v;
            v.x = a == 0 ? va : vcr;
            v.y = a == 1 ? va : (b == 1 ? vb : vcr);
            v.z = b == 2 ? vb : vcr;
            const V3<R> iv = {(R)1 / v.x, (R)1 / v.y, (R)1 / v.z};
            const R tc = face_t(pick(w.o.x, w.o.y, w.o.z, c), pick(v.x, v.y, v.z, c), pick(iv.x, iv.y, iv.z, c));
            const R ts = (R)(r / rhoc);
            // B1: face a first, B0: face b first
            Walker<R> B0 = w, B1 = w;
            V3<R> v0 = v, v1 = v;
            if (tc < ts) {
                march_core<R, 0, EXCH, PER>(B0, L, J, wid, i, v0, iv, tc, ts, ts, c, b, a, xhi, 0, T0);
                march_core<R, 0, EXCH, PER>(B1, L, J, wid, i, v1, iv, tc, ts, ts, c, a, b, xhi, 0, T0);
            } else {
                march_core<R, 0, EXCH, PER>(B0, L, J, wid, i, v0, iv, ts, ts, tc, b, a, c, xhi, 0, T0);
                march_core<R, 0, EXCH, PER>(B1, L, J, wid, i, v1, iv, ts, ts, tc, a, b, c, xhi, 0, T0);
            }
            if (same_outcome(B0, v0, B1, v1)) continue;
            double wp[NP > 0 ? NP : 1], om = 0.0;
#pragma unroll
            for (int p = 0; p < NP; p++) {
                wp[p] = 0.0;
                if (p < L.nclass - 1) {
                    const double Ja = (double)pick(T.Jo[p].x, T.Jo[p].y, T.Jo[p].z, a);
                    const double Jb = (double)pick(T.Jo[p].x, T.Jo[p].y, T.Jo[p].z, b);
                    const double dda = sa > 0 ? -Ja : Ja, ddb = sb > 0 ? -Jb : Jb;
                    wp[p] = 0.15915494309189535 * ((da * ddb - db * dda) / r2);
                    om = om + fabs(wp[p]);
                }
            }
            if (!(om > 0.0)) continue;
            Wj = Wj + om;
            if (corner_uniform<R>(L, wid, i, 4 * pr + qd) * Wj < om) {
                has = true;
                store_cont<R>(cont, 0, t, B0, cell_of<R, BR>(L, B0));
                store_cont<R>(cont, 1, t, B1, cell_of<R, BR>(L, B1));
#pragma unroll
                for (int p = 0; p < NP; p++) rho[p] = wp[p] / om;
            }
        }
    }
}

template <typename R>
__device__ __forceinline__ V3<double> displacement(const Walker<R> &w, double res)
{
    V3<double> d;
    d.x = ((double)(w.g.x - w.gi.x) + ((double)w.o.x - (double)w.oi.x)) * res;
    d.y = ((double)(w.g.y - w.gi.y) + ((double)w.o.y - (double)w.oi.y)) * res;
    d.z = ((double)(w.g.z - w.gi.z) + ((double)w.o.z - (double)w.oi.z)) * res;
    return d;
}

// T2 weight, main.cu:334-338.
template <typename R>
__device__ __forceinline__ double t2_weight(Walker<R> &w, const JobLocal<R> &L)
{
    flush(w);
    double s = 0.0;
    s = s + (w.ts0 / L.T0);
    if (L.nclass > 1) s = s + (w.ts1 / L.T1);
    if (L.nclass > 2) s = s + (w.ts2 / L.T2);
    if (L.nclass > 3) s = s + (w.ts3 / L.T3);
    return exp_neg(-1.0 * s);
}

// Wave-sum one record value and store it from lane 0.  Lanes past the end of
// the job contribute an exact +0.0.
__device__ __forceinline__ void put(double v, bool active, double *dst, int lane)
{
    double t = wave_sum(active ? v : 0.0);
    if (lane == 0) *dst = t;
}

// The fixed part of a record row (main.cu:333-372): position 0 the weight,
// 1..4 the class indicators of classes 0..3, 5..25 the 21 weighted second and
// fourth moments in the row's order, 26..31 zero.  Products are taken left to
// right from s0.
constexpr int8_t REC_MOM[21][4] = {
    {0, 0, -1, -1}, {0, 1, -1, -1}, {0, 2, -1, -1}, {1, 1, -1, -1}, {1, 2, -1, -1}, {2, 2, -1, -1},
    {0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {0, 0, 1, 1}, {0, 0, 1, 2}, {0, 0, 2, 2},
    {0, 1, 1, 1}, {0, 1, 1, 2}, {0, 1, 2, 2}, {0, 2, 2, 2},
    {1, 1, 1, 1}, {1, 1, 1, 2}, {1, 1, 2, 2}, {1, 2, 2, 2}, {2, 2, 2, 2}};

struct RecRow {
    double s0, d[3];
    int32_t c;
    bool active;
};

template <int P>
__device__ __forceinline__ double rec_value(const RecRow &r)
{
    double v = 0.0;
    if constexpr (P == 0) {
        v = r.s0;
    } else if constexpr (P <= 4) {
        v = r.c == P - 1 ? 1.0 : 0.0;
    } else if constexpr (P <= 25) {
        v = r.s0;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (REC_MOM[P - 5][k] >= 0) v = v * r.d[REC_MOM[P - 5][k]];
    }
    return r.active ? v : 0.0;
}

// Reduce-scatter of the 32 positions over lane bits 0..LV-1: slot Q of level
// LV holds, in a lane whose bit LV-1 is b, the sum over lane pairs at distance
// 2^(LV-1) of slot Q + b*H of level LV-1 (H = 32 >> LV).  Each position meets
// exactly the pairs of wave_sum's butterfly in the same order, so its total is
// bit-identical to wave_sum of that value.
template <int LV, int Q>
__device__ __forceinline__ double rec_tree(const RecRow &r, int lane)
{
    if constexpr (LV == 0) {
        return rec_value<Q>(r);
    } else {
        constexpr int H = 32 >> LV;
        const double lo = rec_tree<LV - 1, Q>(r, lane);
        const double hi = rec_tree<LV - 1, Q + H>(r, lane);
        const bool up = ((lane >> (LV - 1)) & 1) != 0;
        return (up ? hi : lo) + xlane<1 << (LV - 1)>(up ? lo : hi);
    }
}

// Wave-sum the fixed part of a record row and store it: lane L < 32 ends with
// the total of position bitrev5(L), written to its place in the row (the class
// indicators only for the geometry's classes).
__device__ __forceinline__ void put_row(const RecRow &r, int nclass, double *dst, int lane)
{
    // an opaque copy of the lane index: the lane-bit masks of the tree are
    // formed here, not hoisted out of the step loop into scalar registers
    asm volatile("" : "+v"(lane));
    double t = rec_tree<5, 0>(r, lane);
    t = t + xlane<32>(t);
    const int p = (int)(__builtin_bitreverse32((uint32_t)lane) >> 27);
    const int off = p == 0 ? 0 : p <= 4 ? (p - 1 < nclass ? p : -1) : p <= 25 ? 1 + nclass + (p - 5) : -1;
    if (lane < 32 && off >= 0) dst[off] = t;
}

template <typename R>
__device__ __forceinline__ void load_state(Walker<R> &w, const StateBuf &st, int64_t s)
{
    const int64_t ws = st.ws;
    w.g = {st.g[s], st.g[ws + s], st.g[2 * ws + s]};
    w.gi = {st.gi[s], st.gi[ws + s], st.gi[2 * ws + s]};
    if constexpr (sizeof(R) == 4) {
        w.o = {st.of[s], st.of[ws + s], st.of[2 * ws + s]};
        w.oi = {st.of[3 * ws + s], st.of[4 * ws + s], st.of[5 * ws + s]};
        w.tcur = st.tcf[s];
    } else {
        w.o = {st.od[s], st.od[ws + s], st.od[2 * ws + s]};
        w.oi = {st.od[3 * ws + s], st.od[4 * ws + s], st.od[5 * ws + s]};
        w.tcur = st.tcd[s];
    }
    w.c = st.c[s];
    w.ts0 = st.tsum[s];
    w.ts1 = st.tsum[ws + s];
    w.ts2 = st.tsum[2 * ws + s];
    w.ts3 = st.tsum[3 * ws + s];
    w.ncross = st.cnt[s];
    w.nrefl = st.cnt[ws + s];
    w.nmirror = st.cnt[2 * ws + s];
    w.ntrunc = st.cnt[3 * ws + s];
    w.ndraw = st.cnt[4 * ws + s];
}

template <typename R>
__device__ __forceinline__ void save_state(const Walker<R> &w, const StateBuf &st, int64_t s)
{
    const int64_t ws = st.ws;
    st.g[s] = w.g.x;
    st.g[ws + s] = w.g.y;
    st.g[2 * ws + s] = w.g.z;
    st.gi[s] = w.gi.x;
    st.gi[ws + s] = w.gi.y;
    st.gi[2 * ws + s] = w.gi.z;
    if constexpr (sizeof(R) == 4) {
        st.of[s] = w.o.x;
        st.of[ws + s] = w.o.y;
        st.of[2 * ws + s] = w.o.z;
        st.of[3 * ws + s] = w.oi.x;
        st.of[4 * ws + s] = w.oi.y;
        st.of[5 * ws + s] = w.oi.z;
        st.tcf[s] = w.tcur;
    } else {
        st.od[s] = w.o.x;
        st.od[ws + s] = w.o.y;
        st.od[2 * ws + s] = w.o.z;
        st.od[3 * ws + s] = w.oi.x;
        st.od[4 * ws + s] = w.oi.y;
        st.od[5 * ws + s] = w.oi.z;
        st.tcd[s] = w.tcur;
    }
    st.c[s] = w.c;
    st.tsum[s] = w.ts0;
    st.tsum[ws + s] = w.ts1;
    st.tsum[2 * ws + s] = w.ts2;
    st.tsum[3 * ws + s] = w.ts3;
    st.cnt[s] = w.ncross;
    st.cnt[ws + s] = w.nrefl;
    st.cnt[2 * ws + s] = w.nmirror;
    st.cnt[3 * ws + s] = w.ntrunc;
    st.cnt[4 * ws + s] = w.ndraw;
}

// Contraction of one record row with its cotangent (adjoint pass):
// F = sum_q cot_q m_q(d) over the T2-weighted columns (m_q = record value
// without the weight: sig0, dx2, dx4, sigRe) and dF/dd.  The unweighted
// NPar_count column is handled by the caller.
__device__ __forceinline__ void adj_record(const double *cot, int nclass, int nbvec, const double *bvec,
                                           const double *qs, double x, double y, double z, double &F, double &g0,
                                           double &g1, double &g2)
{
    const double d[3] = {x, y, z};
    double g[3] = {0.0, 0.0, 0.0};
    double f = cot[0];
    const double *c2 = cot + 1 + nclass;
    constexpr int DI[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int ia = DI[q][0], ib = DI[q][1];
        f = f + c2[q] * (d[ia] * d[ib]);
        g[ia] = g[ia] + c2[q] * d[ib];
        g[ib] = g[ib] + c2[q] * d[ia];
    }
    const double *c4 = c2 + 6;
    constexpr int WI[15][4] = {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 2}, {0, 0, 1, 1}, {0, 0, 1, 2},
                               {0, 0, 2, 2}, {0, 1, 1, 1}, {0, 1, 1, 2}, {0, 1, 2, 2}, {0, 2, 2, 2},
                               {1, 1, 1, 1}, {1, 1, 1, 2}, {1, 1, 2, 2}, {1, 2, 2, 2}, {2, 2, 2, 2}};
#pragma unroll
    for (int q = 0; q < 15; q++) {
        const int i0 = WI[q][0], i1 = WI[q][1], i2 = WI[q][2], i3 = WI[q][3];
        f = f + c4[q] * (((d[i0] * d[i1]) * d[i2]) * d[i3]);
        g[i0] = g[i0] + c4[q] * ((d[i1] * d[i2]) * d[i3]);
        g[i1] = g[i1] + c4[q] * ((d[i0] * d[i2]) * d[i3]);
        g[i2] = g[i2] + c4[q] * ((d[i0] * d[i1]) * d[i3]);
        g[i3] = g[i3] + c4[q] * ((d[i0] * d[i1]) * d[i2]);
    }
    const double *cs = c4 + 15;
    for (int j = 0; j < nbvec; j++) {
        const double *bv = bvec + 3 * j;
        const double arg = qs[j] * (d[0] * bv[0] + d[1] * bv[1] + d[2] * bv[2]);
        double sn, cn;
        sincos_any(arg, sn, cn);
        f = f + cs[j] * cn;
        const double t = (cs[j] * sn) * qs[j];
        g[0] = g[0] - t * bv[0];
        g[1] = g[1] - t * bv[1];
        g[2] = g[2] - t * bv[2];
    }
    F = f;
    g0 = g[0];
    g1 = g[1];
    g2 = g[2];
}

// The persistent walk kernel.  SIG: bit 0 narrow-pulse b-table signal,
// bit 1 PGSE signal, bit 2 exchange faces on the fast path, bit 3 periodic
// lattice edges (the march wraps; without it every edge is a mirror), bit 4
// (adjoint only, no PGSE) the jump of the D derivative at exchange faces,
// bit 5 a geometry of the plan has its cell words in bricks, bit 6 (with
// bit 3) a geometry's cell words name the voxels across periodic sides.
// NP > 0 instantiates the adjoint pass: the same walk with tangents for D of
// classes 1..NP, contracted with the cotangents of the record rows and PGSE
// signal into per-batch parameter gradients.
//
// Occupancy targets (waves per SIMD, 512 VGPRs each): 3 for the forward walk
// and the FP32 narrow-pulse tangent pass (at most a few spills outside the
// step loop), 2 for the PGSE, FP64 and jump tangent passes.
template <typename R, int SIG, int NP>
__global__ __launch_bounds__(256, (NP == 0 || (sizeof(R) == 4 && (SIG & 18) == 0)) ? 3 : 2)
void walk_kernel(LaunchArgs a)
{
    constexpr bool JUMP = NP > 0 && (SIG & 16) != 0;
    constexpr bool EXCH = (SIG & 4) != 0;
    constexpr bool PER = (SIG & 8) != 0;
    constexpr bool BR = (SIG & 32) != 0;
    constexpr bool WRAP = PER && (SIG & 64) != 0;
    // the fast path's diagnostic counters: forward walks only (the tangent
    // passes report none), and not in the variants built without them
    // (bit 7, for launches that do not report counters)
    constexpr bool CNT = NP == 0 && (SIG & 128) == 0;
    // bit 8 (forward walks): every class's T2 is infinite and no class pair
    // draws, so the T2 weight is exp(-0) = 1 exactly and no step needs the
    // bits of an exchange draw
    constexpr bool LEAN = NP == 0 && (SIG & 256) != 0;
    // sin/cos table of the FP32 direction draw, staged in LDS
    __shared__ float2 tab[256];
    // continuations of the kept boundary events (jump variants only)
    __shared__ uint32_t cont[JUMP ? Cont<R>::WORDS * Cont<R>::ROW : 1];
    for (int k = (int)threadIdx.x; k < 256; k += (int)blockDim.x) tab[k] = turn_table[k];
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63u);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int wpb = (int)(blockDim.x >> 6);
    int64_t first, stride, hi;
    if (a.xcd_split) {
        // workgroups are dispatched round-robin over the 8 XCDs: give each
        // XCD a contiguous range of batches so a geometry stays in one L2
        const int xcd = (int)(blockIdx.x & 7u);
        const int64_t bx = (int64_t)(blockIdx.x >> 3);
        const int64_t nbx = (int64_t)(gridDim.x >> 3);
        const int64_t lo = a.nbatch * xcd / 8;
        hi = a.nbatch * (xcd + 1) / 8;
        first = lo + bx * wpb + wv;
        stride = nbx * wpb;
    } else {
        first = (int64_t)blockIdx.x * wpb + wv;
        stride = (int64_t)gridDim.x * wpb;
        hi = a.nbatch;
    }
    for (int64_t b = first; b < hi; b += stride) {
        const int jid = __builtin_amdgcn_readfirstlane(a.batch_job[b]);
        const JobDesc &J = a.jobs[jid];
        const JobLocal<R> L = job_local<R>(J);
        const int64_t lb = b - J.batch0;                   // batch within the job
        const int64_t kw = lb * WAVE + lane;               // walker of the job's launch
        const bool active = kw < J.nwalk;
        // Philox walker id (a subset run walks the listed walkers of the job)
        const uint64_t wid = (uint64_t)(J.wid0 + (J.wlist != nullptr ? (int64_t)J.wlist[kw] : kw));
        const int64_t slot = J.wbase + lb * WAVE + lane;    // state slot
        const int nbvec = (SIG & 1) ? J.nbvec : 0;
        const int nval = 1 + L.nclass + 21 + nbvec;
        const int nbound = (SIG & 2) ? J.nbound : 0;
        const int64_t ws = a.st.ws;

        Walker<R> w;
        w.ts0 = w.ts1 = w.ts2 = w.ts3 = 0.0;
        w.ncross = w.nrefl = w.nmirror = w.ntrunc = w.ndraw = 0;
        w.tcur = (R)0;
        Tan<R, NP> T;
        double GD[NP > 0 ? NP : 1], GT0 = 0.0, GT1 = 0.0, GT2 = 0.0, GT3 = 0.0, GK = 0.0;
        if constexpr (NP > 0) {
#pragma unroll
            for (int p = 0; p < NP; p++) {
                T.Jo[p] = {(R)0, (R)0, (R)0};
                T.dE[p] = 0.0;
                T.S[p] = 0.0;
                T.PJ[p] = {0.0, 0.0, 0.0};
                GD[p] = 0.0;
            }
            T.Sk = 0.0;
        }
        w.ps = {0.0, 0.0, 0.0};
        w.sg = {0, 0, 0};
        if (a.step0 == 0) {
            const bool started = init_walker<R>(w, L, wid);
            if constexpr (NP == 0) {
                const uint32_t nfail = wave_sum_u32((active && !started) ? 1u : 0u);
                if (lane == 0 && nfail != 0u) atomicAdd(J.counters + 5, (unsigned long long)nfail);
            }
            w.sc = sel4(w.c, L.st0, L.st1, L.st2, L.st3);
            w.isc = sel4(w.c, L.is0, L.is1, L.is2, L.is3);
            w.hc = w.c == 1 ? L.xh12 : L.xh21;
        } else {
            load_state<R>(w, a.st, slot);
            w.lin = w.g.x * L.sx + w.g.y * L.nz + w.g.z;
            w.sc = sel4(w.c, L.st0, L.st1, L.st2, L.st3);
            w.isc = sel4(w.c, L.is0, L.is1, L.is2, L.is3);
            w.hc = w.c == 1 ? L.xh12 : L.xh21;
            if (SIG & 2) {
                w.ps = {a.st.ps[slot], a.st.ps[ws + slot], a.st.ps[2 * ws + slot]};
                w.sg = {a.st.sg[slot], a.st.sg[ws + slot], a.st.sg[2 * ws + slot]};
            }
        }
        if constexpr (NP > 0) T.vis = 1u << w.c;
        // step counters fit int32 (TN < 2^31 is checked on the host) and are
        // wave-uniform: next record step and next PGSE boundary at or after step0
        const int32_t step0 = (int32_t)a.step0, step1 = (int32_t)a.step1, Tstep = (int32_t)a.Tstep;
        int32_t rnext = ((step0 + Tstep - 1) / Tstep) * Tstep;
        const int32_t rend = a.timepoints * Tstep;  // records only for r < timepoints
        int kb = 1;
        int32_t bnext = -1;
        // PGSE flushes after the steps i with i + 1 a lobe boundary or a
        // multiple of 2^pg_log2; i0 = the last flush point at or before step0
        const int32_t pgmask = (int32_t)((1u << J.pg_log2) - 1u);
        int32_t i0 = step0 & ~pgmask;
        if (SIG & 2) {
            while (kb <= nbound && J.bounds[kb] <= a.step0) kb++;
            bnext = __builtin_amdgcn_readfirstlane(kb <= nbound ? (int32_t)J.bounds[kb] : -1);
            // bounds[0] = 0; a job without a PGSE table (nbound = 0) in a plan
            // whose other jobs have one has no bounds array at all
            const int32_t lastb = kb > 1 ? (int32_t)J.bounds[kb - 1] : 0;
            i0 = __builtin_amdgcn_readfirstlane(lastb > i0 ? lastb : i0);
        }
        uint32_t b0 = 0, b1 = 0;
        double *part = J.part + lb * (int64_t)a.timepoints * nval;
        int32_t r = rnext / Tstep;  // record index of rnext
        uint64_t cell = cell_of<R, BR>(L, w);
        // adjoint boundary terms: the kept pair of continuations (in LDS,
        // continuation 0 without the event, 1 with it) and the reservoir
        const int tix = (int)threadIdx.x;
        bool has = false;
        double Wj = 0.0, rho[NP > 0 ? NP : 1];
#pragma unroll
        for (int p = 0; p < (NP > 0 ? NP : 1); p++) rho[p] = 0.0;

        for (int32_t i = step0; i < step1; i++) {
            const bool fresh = ((i & 1) == 0) || (i == step0);
            uint32_t xhi;
            const V3<R> v = draw_dir<R, !LEAN>(L, wid, i, fresh, b0, b1, tab, xhi);
            if constexpr (JUMP) {
                // the continuations walk with the walker's direction and draws
                if (has) {
                    Tan<R, 0> T0;
#pragma unroll 1
                    for (int m = 0; m < 2; m++) {
                        Walker<R> a;
                        uint64_t ca;
                        load_cont<R>(cont, m, tix, L, a, ca);
                        const int32_t l0 = a.lin;
                        if (!fast_step<R, 0, EXCH, PER, WRAP, false, true>(a, L, J, ca, v, wid, i, xhi, 0, T0))
                            march<R, 0, EXCH, PER>(a, L, J, wid, i, v, xhi, 0, T0);
                        if (a.lin != l0) ca = cell_of<R, BR>(L, a);
                        store_cont<R>(cont, m, tix, a, ca);
                    }
                }
                // exchange-face and lattice-edge events of this step from the
                // walker's state before it; none without another class (or a
                // lattice edge) among the 26 neighbours, which every boundary
                // path stays within
                if ((cell & 0x002AAAAAAAAAAAAAull) != 0ull) {
                    // per-lane selects of the batch's uniform values (no loads
                    // under a class branch)
                    const double sd = sel4(w.c, L.sd0, L.sd1, L.sd2, L.sd3);
                    const double dsc = sel4(w.c, 0.0, L.dsd1, L.dsd2, L.dsd3);
                    const double vd[3] = {(double)v.x, (double)v.y, (double)v.z};
                    for (int m = 0; m < 6; m++) {
                        const int k = m >> 1, sg = (m & 1) ? 1 : -1;
                        const double ok = (double)pick(w.o.x, w.o.y, w.o.z, k);
                        const double dist = sg > 0 ? 1.0 - ok : ok;
                        if (!(dist < sd)) continue;
                        Walker<R> B0, B1;
                        if (!jump_event<R, PER>(w, L, vd, k, sg, sd, dist, wid, i, xhi, B0, B1)) continue;
                        const double cs = dist / sd;
                        double wp[NP > 0 ? NP : 1], om = 0.0;
#pragma unroll
                        for (int p = 0; p < NP; p++) {
                            wp[p] = 0.0;
                            if (p < L.nclass - 1) {
                                const double Jk = (double)pick(T.Jo[p].x, T.Jo[p].y, T.Jo[p].z, k);
                                const double dd = sg > 0 ? -Jk : Jk;
                                const double ds = (w.c == p + 1) ? dsc : 0.0;
                                wp[p] = -0.5 * ((dd - cs * ds) / sd);
                                om = om + fabs(wp[p]);
                            }
                        }
                        if (!(om > 0.0)) continue;
                        Wj = Wj + om;
                        if (jump_uniform<R>(L, wid, i, m) * Wj < om) {
                            has = true;
                            store_cont<R>(cont, 0, tix, B0, cell_of<R, BR>(L, B0));
                            store_cont<R>(cont, 1, tix, B1, cell_of<R, BR>(L, B1));
#pragma unroll
                            for (int p = 0; p < NP; p++) rho[p] = wp[p] / om;
                        }
                    }
                }
                // voxel-edge events: none when all 26 neighbours are of the
                // walker's class (both orders then cross every face they
                // reach, with no reflection or class change)
                if (cell != 0ull) {
                    const double sd = sel4(w.c, L.sd0, L.sd1, L.sd2, L.sd3);
                    corner_events<R, NP, EXCH, PER, BR>(w, T, L, J, v, sd, wid, i, xhi, has, Wj, rho, cont, tix);
                }
            }
            const int32_t lin0 = w.lin;
            const int32_t npg = i - i0;  // PGSE terms since the last flush
            if (!fast_step<R, NP, EXCH, PER, WRAP, CNT, !LEAN>(w, L, J, cell, v, wid, i, xhi, npg, T))
                march<R, NP, EXCH, PER>(w, L, J, wid, i, v, xhi, npg, T);
            if (w.lin != lin0) cell = cell_of<R, BR>(L, w);  // only lanes that changed voxel read
            if (SIG & 2) {
                // running sums of the displacement in voxels, (g + o) - (gi + oi),
                // as sum(o) in FP64 and sum(g) in int32 until the next flush; the
                // snapshots at lobe boundaries are scaled to um
                w.ps.x = w.ps.x + (double)w.o.x;
                w.ps.y = w.ps.y + (double)w.o.y;
                w.ps.z = w.ps.z + (double)w.o.z;
                w.sg.x += w.g.x;
                w.sg.y += w.g.y;
                w.sg.z += w.g.z;
                if constexpr (NP > 0) {
#pragma unroll
                    for (int p = 0; p < NP; p++) {
                        T.PJ[p].x = T.PJ[p].x + (double)T.Jo[p].x;
                        T.PJ[p].y = T.PJ[p].y + (double)T.Jo[p].y;
                        T.PJ[p].z = T.PJ[p].z + (double)T.Jo[p].z;
                    }
                }
                const bool snap = i + 1 == bnext;
                if (__builtin_expect(snap || ((i + 1) & pgmask) == 0, 0)) {
                    const double n = (double)(i + 1 - i0);
                    w.ps.x = w.ps.x + ((double)w.sg.x - n * origin(w.gi.x, w.oi.x));
                    w.ps.y = w.ps.y + ((double)w.sg.y - n * origin(w.gi.y, w.oi.y));
                    w.ps.z = w.ps.z + ((double)w.sg.z - n * origin(w.gi.z, w.oi.z));
                    w.sg = {0, 0, 0};
                    i0 = i + 1;
                }
                if (snap) {
                    double *sp = J.snap + (int64_t)(kb * 3) * ws + slot;
                    sp[0] = w.ps.x * L.res;
                    sp[ws] = w.ps.y * L.res;
                    sp[2 * ws] = w.ps.z * L.res;
                    if constexpr (NP > 0) {
#pragma unroll
                        for (int p = 0; p < NP; p++) {
                            double *tp = J.tsnap + (int64_t)((kb * NP + p) * 3) * ws + slot;
                            tp[0] = T.PJ[p].x * L.res;
                            tp[ws] = T.PJ[p].y * L.res;
                            tp[2 * ws] = T.PJ[p].z * L.res;
                        }
                    }
                    kb++;
                    bnext = __builtin_amdgcn_readfirstlane(kb <= nbound ? (int32_t)J.bounds[kb] : -1);
                }
            }
            if (i == rnext && i < rend) {
                // records, main.cu:333-378
                rnext += Tstep;
                const double s0 = LEAN ? 1.0 : t2_weight<R>(w, L);
                const V3<double> d = displacement<R>(w, L.res);
                const double x = d.x, y = d.y, z = d.z;
                if constexpr (NP == 0) {
                    double *dst = part + (int64_t)r * nval;
                    put_row(RecRow{s0, {x, y, z}, w.c, active}, L.nclass, dst, lane);
                    double *m = dst + 1 + L.nclass;
                    if (SIG & 1) {
                        const double *qs = J.qscale + (int64_t)r * nbvec;
                        for (int jb = 0; jb < nbvec; jb++) {
                            const double *bv = J.bvec + 3 * jb;
                            const double qx = qs[jb] * (x * bv[0] + y * bv[1] + z * bv[2]);
                            put(s0 * cos_any(qx), active, m + 21 + jb, lane);
                        }
                    }
                } else if (J.cot_rec) {
                    // adjoint: contract this record row with its cotangent
                    double F, g0, g1, g2;
                    adj_record(J.cot_rec + (int64_t)r * nval, L.nclass, nbvec, J.bvec,
                               (SIG & 1) ? J.qscale + (int64_t)r * nbvec : nullptr, x, y, z, F, g0, g1, g2);
                    const double sF = s0 * F;
                    const double s0r = s0 * L.res;
                    // NPar_count is unweighted: only the score terms
                    const double fc = J.cot_rec[(int64_t)r * nval + 1 + w.c];
                    // the score terms against a leave-one-out baseline: the mean
                    // of the other walkers' F_r in the batch, independent of
                    // this walker's draws
                    const double Fr = sF + fc;
                    const double tot = wave_sum(active ? Fr : 0.0);
                    const int nact = __popcll(__ballot(active));
                    const double bl = nact > 1 ? (tot - Fr) / (double)(nact - 1) : 0.0;
                    const double Fc = Fr - bl;
#pragma unroll
                    for (int p = 0; p < NP; p++) {
                        if (p < L.nclass - 1) {
                            const double gJ = g0 * (double)T.Jo[p].x + g1 * (double)T.Jo[p].y + g2 * (double)T.Jo[p].z;
                            GD[p] = GD[p] + ((Fc * T.S[p] - sF * T.dE[p]) + s0r * gJ);
                        }
                    }
                    GT0 = GT0 + (sF * (w.ts0 / L.T0)) * J.invT2[0];
                    if (L.nclass > 1) GT1 = GT1 + (sF * (w.ts1 / L.T1)) * J.invT2[1];
                    if (L.nclass > 2) GT2 = GT2 + (sF * (w.ts2 / L.T2)) * J.invT2[2];
                    if (L.nclass > 3) GT3 = GT3 + (sF * (w.ts3 / L.T3)) * J.invT2[3];
                    GK = GK + Fc * T.Sk;
                    if constexpr (JUMP) {
                        if (has) {
                            // the kept event: W (w_p / omega) (F_cross - F_stay)
                            const double *qs = (SIG & 1) ? J.qscale + (int64_t)r * nbvec : nullptr;
                            const double *cr = J.cot_rec + (int64_t)r * nval;
                            double Fm[2];
#pragma unroll 1
                            for (int m = 0; m < 2; m++) {
                                Walker<R> A;
                                uint64_t ca;
                                load_cont<R>(cont, m, tix, L, A, ca);
                                const double s0m = t2_weight<R>(A, L);  // flushes A
                                store_cont<R>(cont, m, tix, A, ca);
                                const V3<double> dm = displacement<R>(A, L.res);
                                double F0, h0, h1, h2;
                                adj_record(cr, L.nclass, nbvec, J.bvec, qs, dm.x, dm.y, dm.z, F0, h0, h1, h2);
                                Fm[m] = s0m * F0 + cr[1 + A.c];
                            }
                            const double dF = Fm[1] - Fm[0];
#pragma unroll
                            for (int p = 0; p < NP; p++)
                                if (p < L.nclass - 1) GD[p] = GD[p] + (Wj * rho[p]) * dF;
                        }
                    }
                }
                r++;
            }
        }
        if ((SIG & 2) && a.step1 == a.TN) {
            // final PGSE signal, main_PGSE.cu:348-357
            const double s0 = LEAN ? 1.0 : t2_weight<R>(w, L);
            double *psig = J.psig + lb * (int64_t)J.ngrad;
            double Fs = 0.0, H[NP > 0 ? NP : 1];
#pragma unroll
            for (int p = 0; p < (NP > 0 ? NP : 1); p++) H[p] = 0.0;
            for (int t = 0; t < J.ntriple; t++) {
                const int32_t *tr = J.triple + 4 * t;
                const int ka = tr[0], kk = tr[1], ke = tr[2], g0 = tr[3], g1 = tr[7];
                double df[3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const double s1 = ka == 0 ? 0.0 : J.snap[(int64_t)(ka * 3 + k) * ws + slot];
                    const double se = ke == 0 ? 0.0 : J.snap[(int64_t)(ke * 3 + k) * ws + slot];
                    const double sb = kk == 0 ? 0.0 : J.snap[(int64_t)(kk * 3 + k) * ws + slot];
                    const double s2 = se - sb;
                    df[k] = s1 - s2;
                }
                V3<double> dJ[NP > 0 ? NP : 1];
                if constexpr (NP > 0) {
#pragma unroll
                    for (int p = 0; p < NP; p++) {
                        double q3[3];
#pragma unroll
                        for (int k = 0; k < 3; k++) {
                            const double s1 = ka == 0 ? 0.0 : J.tsnap[(int64_t)((ka * NP + p) * 3 + k) * ws + slot];
                            const double se = ke == 0 ? 0.0 : J.tsnap[(int64_t)((ke * NP + p) * 3 + k) * ws + slot];
                            const double sb = kk == 0 ? 0.0 : J.tsnap[(int64_t)((kk * NP + p) * 3 + k) * ws + slot];
                            const double s2 = se - sb;
                            q3[k] = s1 - s2;
                        }
                        dJ[p] = {q3[0], q3[1], q3[2]};
                    }
                }
                for (int jg = g0; jg < g1; jg++) {
                    const double *gr = J.grad + 6 * jg;
                    const double dot = df[0] * gr[3] + df[1] * gr[4] + df[2] * gr[5];
                    const double ph = (gr[2] * dot) * a.dt;
                    if constexpr (NP == 0) {
                        const double tsum = wave_sum(active ? s0 * cos_any(ph) : 0.0);
                        if (lane == 0) psig[J.gperm[jg]] = tsum;
                    } else if (J.cot_sig) {
                        double sn, cn;
                        sincos_any(ph, sn, cn);
                        const double cs = J.cot_sig[J.gperm[jg]];
                        Fs = Fs + cs * cn;
                        const double wgt = cs * sn;
#pragma unroll
                        for (int p = 0; p < NP; p++) {
                            const double dotJ = dJ[p].x * gr[3] + dJ[p].y * gr[4] + dJ[p].z * gr[5];
                            const double dph = (gr[2] * dotJ) * a.dt;
                            H[p] = H[p] + wgt * dph;
                        }
                    }
                }
            }
            if constexpr (NP > 0) {
                if (J.cot_sig) {
                    const double sFs = s0 * Fs;
                    // leave-one-out baseline of the score terms, as for records
                    const double tot = wave_sum(active ? sFs : 0.0);
                    const int nact = __popcll(__ballot(active));
                    const double bl = nact > 1 ? (tot - sFs) / (double)(nact - 1) : 0.0;
                    const double Fc = sFs - bl;
#pragma unroll
                    for (int p = 0; p < NP; p++)
                        if (p < L.nclass - 1) GD[p] = GD[p] + ((Fc * T.S[p] - sFs * T.dE[p]) - s0 * H[p]);
                    GT0 = GT0 + (sFs * (w.ts0 / L.T0)) * J.invT2[0];
                    if (L.nclass > 1) GT1 = GT1 + (sFs * (w.ts1 / L.T1)) * J.invT2[1];
                    if (L.nclass > 2) GT2 = GT2 + (sFs * (w.ts2 / L.T2)) * J.invT2[2];
                    if (L.nclass > 3) GT3 = GT3 + (sFs * (w.ts3 / L.T3)) * J.invT2[3];
                    GK = GK + Fc * T.Sk;
                }
            }
        }
        if constexpr (NP > 0) {
            // per-batch gradient partials [16]: D[1..3], T2[0..3] at 4.., kappa at 8
            double *pg = J.pgrad + lb * NPARAM;
#pragma unroll
            for (int p = 0; p < NP; p++) put(GD[p], active, pg + 1 + p, lane);
            if (J.wvis != nullptr && active) J.wvis[kw] = T.vis;
            put(GT0, active, pg + 4, lane);
            put(GT1, active, pg + 5, lane);
            put(GT2, active, pg + 6, lane);
            put(GT3, active, pg + 7, lane);
            put(GK, active, pg + 8, lane);
        } else if (a.step1 == a.TN) {
            // diagnostics: exact integer sums (only counted when the launch
            // asks; the march's own increments are not reported otherwise)
            if (a.count) {
                const uint32_t cs[5] = {w.ncross, w.nrefl, w.nmirror, w.ntrunc, w.ndraw};
#pragma unroll
                for (int k = 0; k < 5; k++) {
                    const uint32_t t = wave_sum_u32(active ? cs[k] : 0u);
                    if (lane == 0) atomicAdd(J.counters + k, (unsigned long long)t);
                }
            }
            if (a.fin.g && active) {
                a.fin.g[slot] = w.g.x;
                a.fin.g[ws + slot] = w.g.y;
                a.fin.g[2 * ws + slot] = w.g.z;
                a.fin.o[slot] = (double)w.o.x;
                a.fin.o[ws + slot] = (double)w.o.y;
                a.fin.o[2 * ws + slot] = (double)w.o.z;
                a.fin.gi[slot] = w.gi.x;
                a.fin.gi[ws + slot] = w.gi.y;
                a.fin.gi[2 * ws + slot] = w.gi.z;
                a.fin.oi[slot] = (double)w.oi.x;
                a.fin.oi[ws + slot] = (double)w.oi.y;
                a.fin.oi[2 * ws + slot] = (double)w.oi.z;
                a.fin.c[slot] = w.c;
                a.fin.cnt[slot] = w.ncross;
                a.fin.cnt[ws + slot] = w.nrefl;
                a.fin.cnt[2 * ws + slot] = w.nmirror;
                a.fin.cnt[3 * ws + slot] = w.ntrunc;
                a.fin.cnt[4 * ws + slot] = w.ndraw;
            }
        } else {
            save_state<R>(w, a.st, slot);
            if (SIG & 2) {
                a.st.ps[slot] = w.ps.x;
                a.st.ps[ws + slot] = w.ps.y;
                a.st.ps[2 * ws + slot] = w.ps.z;
                a.st.sg[slot] = w.sg.x;
                a.st.sg[ws + slot] = w.sg.y;
                a.st.sg[2 * ws + slot] = w.sg.z;
            }
        }
    }
}

}  // namespace rms
