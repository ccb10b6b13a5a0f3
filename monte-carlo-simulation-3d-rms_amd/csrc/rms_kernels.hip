// rms_kernels.hip -- the RMS random-walk engine for gfx950 (MI355X).
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
// and written as per-batch partials; a second kernel sums the partials in a
// fixed order, so results are bit-reproducible and identical to the host
// restatement.
//
// Reference anchors:
//   propagate (narrow pulse)  main.cu:63-384
//   propagate (PGSE)          main_PGSE.cu:65-410
//   atomAdd CAS loop          main.cu:40-56   -> butterfly + ordered partials
//   setup_kernel              main.cu:58-61   -> counter-based Philox, no state
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
};

// Per-job constants held in registers for the whole batch.
template <typename R>
struct JobLocal {
    const uint64_t *cells;
    const uint8_t *cls;
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
template <typename R>
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
        xhi = ((w0 & 0xFFu) << 8) | (w1 & 0xFFu);
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
template <typename R, int NP, bool EXCH, bool PER>
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
        const int32_t hc = w.c == 1 ? L.xh12 : L.xh21;
        const bool easy = om == 0u || ((om & (om - 1u)) == 0u && (int32_t)xhi >= hc);
        rare = (rels & (rels >> 1) & 0x15u) != 0u || kx == ky || !easy;
        refl = (rels | (rels >> 1)) & 0x15u;
        ndraw = (om != 0u && hc >= 0) ? 1u : 0u;
    }
    if (rare) {
        // a near tie goes to the march.  A lattice edge is a mirror: the
        // walker reflects and its origin is mirrored (main.cu:246-265,
        // 302-310).
        if (kx == ky) return false;
        const uint32_t edge = rels & (rels >> 1) & 0x15u;  // CELL_EDGE slots
        const int32_t kz = 3 - kx - ky;
        if (PER && edge != 0u) {
            // a periodic edge wraps around: the march crosses into the voxel
            // on the other side (decided before any state changes)
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
    const uint32_t nref = (uint32_t)__builtin_popcount(refl);
    w.nrefl += nref - nedge;
    w.nmirror += nedge;
    w.ndraw += ndraw;
    w.ncross += (uint32_t)(__mul24(cx, cx) + __mul24(cy, cy) + __mul24(cz, cz)) - nref;
    w.tcur = w.tcur + (R)1;
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
template <typename R, int NP, bool EXCH, bool PER>
__device__ __forceinline__ void march(Walker<R> &w, const JobLocal<R> &L, const JobDesc &J, uint64_t wid,
                                      int64_t i, V3<R> v, uint32_t xhi, int32_t npg, Tan<R, NP> &T)
{
    V3<R> iv = {(R)1 / v.x, (R)1 / v.y, (R)1 / v.z};
    // face distances in units of path, main.cu:199-219
    V3<R> t;
    for3([&](auto K) __attribute__((always_inline)) {
        constexpr int k = decltype(K)::value;
        const R vk = at<k>(v);
        const R ok = at<k>(w.o);
        const R tmp = vk > (R)0 ? ((R)1 - ok) : (vk < (R)0 ? -ok : (R)2);
        R q = tmp * at<k>(iv);
        q = (q > (R)0) ? q : (R)0;
        at<k>(t) = (fabs(tmp) > fabs(vk)) ? (R)2 : q;
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
                        w.sc = sel4(cn, L.st0, L.st1, L.st2, L.st3);
                        w.isc = sel4(cn, L.is0, L.is1, L.is2, L.is3);
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
// lattice edges (the march wraps; without it every edge is a mirror).  NP > 0 instantiates the adjoint pass: the same walk
// with tangents for D of classes 1..NP, contracted with the cotangents of the
// record rows and PGSE signal into per-batch parameter gradients.
//
// Occupancy targets (waves per SIMD, 512 VGPRs each): 3 for the forward walk
// and the FP32 narrow-pulse tangent pass (at most a few spills outside the
// step loop), 2 for the PGSE and FP64 tangent passes.
template <typename R, int SIG, int NP>
__global__ __launch_bounds__(256, (NP == 0 || (sizeof(R) == 4 && (SIG & 2) == 0)) ? 3 : 2)
void walk_kernel(LaunchArgs a)
{
    // sin/cos table of the FP32 direction draw, staged in LDS
    __shared__ float2 tab[256];
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
        const bool active = lb * WAVE + lane < J.nwalk;
        const uint64_t wid = (uint64_t)(J.wid0 + lb * WAVE + lane);  // Philox walker id
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
        } else {
            load_state<R>(w, a.st, slot);
            w.lin = w.g.x * L.sx + w.g.y * L.nz + w.g.z;
            w.sc = sel4(w.c, L.st0, L.st1, L.st2, L.st3);
            w.isc = sel4(w.c, L.is0, L.is1, L.is2, L.is3);
            if (SIG & 2) {
                w.ps = {a.st.ps[slot], a.st.ps[ws + slot], a.st.ps[2 * ws + slot]};
                w.sg = {a.st.sg[slot], a.st.sg[ws + slot], a.st.sg[2 * ws + slot]};
            }
        }
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
        uint64_t cell = L.cells[w.lin];

        for (int32_t i = step0; i < step1; i++) {
            const bool fresh = ((i & 1) == 0) || (i == step0);
            uint32_t xhi;
            const V3<R> v = draw_dir<R>(L, wid, i, fresh, b0, b1, tab, xhi);
            const int32_t lin0 = w.lin;
            const int32_t npg = i - i0;  // PGSE terms since the last flush
            if (!fast_step<R, NP, (SIG & 4) != 0, (SIG & 8) != 0>(w, L, J, cell, v, wid, i, xhi, npg, T))
                march<R, NP, (SIG & 4) != 0, (SIG & 8) != 0>(w, L, J, wid, i, v, xhi, npg, T);
            if (w.lin != lin0) cell = L.cells[w.lin];  // only lanes that changed voxel read
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
                const double s0 = t2_weight<R>(w, L);
                const V3<double> d = displacement<R>(w, L.res);
                const double x = d.x, y = d.y, z = d.z;
                if constexpr (NP == 0) {
                    double *dst = part + (int64_t)r * nval;
                    put(s0, active, dst + 0, lane);
                    for (int cc = 0; cc < L.nclass; cc++) put(cc == w.c ? 1.0 : 0.0, active, dst + 1 + cc, lane);
                    double *m = dst + 1 + L.nclass;
                    put(s0 * x * x, active, m + 0, lane);
                    put(s0 * x * y, active, m + 1, lane);
                    put(s0 * x * z, active, m + 2, lane);
                    put(s0 * y * y, active, m + 3, lane);
                    put(s0 * y * z, active, m + 4, lane);
                    put(s0 * z * z, active, m + 5, lane);
                    put(s0 * x * x * x * x, active, m + 6, lane);
                    put(s0 * x * x * x * y, active, m + 7, lane);
                    put(s0 * x * x * x * z, active, m + 8, lane);
                    put(s0 * x * x * y * y, active, m + 9, lane);
                    put(s0 * x * x * y * z, active, m + 10, lane);
                    put(s0 * x * x * z * z, active, m + 11, lane);
                    put(s0 * x * y * y * y, active, m + 12, lane);
                    put(s0 * x * y * y * z, active, m + 13, lane);
                    put(s0 * x * y * z * z, active, m + 14, lane);
                    put(s0 * x * z * z * z, active, m + 15, lane);
                    put(s0 * y * y * y * y, active, m + 16, lane);
                    put(s0 * y * y * y * z, active, m + 17, lane);
                    put(s0 * y * y * z * z, active, m + 18, lane);
                    put(s0 * y * z * z * z, active, m + 19, lane);
                    put(s0 * z * z * z * z, active, m + 20, lane);
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
#pragma unroll
                    for (int p = 0; p < NP; p++) {
                        if (p < L.nclass - 1) {
                            const double gJ = g0 * (double)T.Jo[p].x + g1 * (double)T.Jo[p].y + g2 * (double)T.Jo[p].z;
                            GD[p] = GD[p] + ((sF * (T.S[p] - T.dE[p]) + s0r * gJ) + fc * T.S[p]);
                        }
                    }
                    GT0 = GT0 + (sF * (w.ts0 / L.T0)) * J.invT2[0];
                    if (L.nclass > 1) GT1 = GT1 + (sF * (w.ts1 / L.T1)) * J.invT2[1];
                    if (L.nclass > 2) GT2 = GT2 + (sF * (w.ts2 / L.T2)) * J.invT2[2];
                    if (L.nclass > 3) GT3 = GT3 + (sF * (w.ts3 / L.T3)) * J.invT2[3];
                    GK = GK + (sF + fc) * T.Sk;
                }
                r++;
            }
        }
        if ((SIG & 2) && a.step1 == a.TN) {
            // final PGSE signal, main_PGSE.cu:348-357
            const double s0 = t2_weight<R>(w, L);
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
#pragma unroll
                    for (int p = 0; p < NP; p++)
                        if (p < L.nclass - 1) GD[p] = GD[p] + (sFs * (T.S[p] - T.dE[p]) - s0 * H[p]);
                    GT0 = GT0 + (sFs * (w.ts0 / L.T0)) * J.invT2[0];
                    if (L.nclass > 1) GT1 = GT1 + (sFs * (w.ts1 / L.T1)) * J.invT2[1];
                    if (L.nclass > 2) GT2 = GT2 + (sFs * (w.ts2 / L.T2)) * J.invT2[2];
                    if (L.nclass > 3) GT3 = GT3 + (sFs * (w.ts3 / L.T3)) * J.invT2[3];
                    GK = GK + sFs * T.Sk;
                }
            }
        }
        if constexpr (NP > 0) {
            // per-batch gradient partials [16]: D[1..3], T2[0..3] at 4.., kappa at 8
            double *pg = J.pgrad + lb * NPARAM;
#pragma unroll
            for (int p = 0; p < NP; p++) put(GD[p], active, pg + 1 + p, lane);
            put(GT0, active, pg + 4, lane);
            put(GT1, active, pg + 5, lane);
            put(GT2, active, pg + 6, lane);
            put(GT3, active, pg + 7, lane);
            put(GK, active, pg + 8, lane);
        } else if (a.step1 == a.TN) {
            // diagnostics: exact integer sums
            const uint32_t cs[5] = {w.ncross, w.nrefl, w.nmirror, w.ntrunc, w.ndraw};
#pragma unroll
            for (int k = 0; k < 5; k++) {
                const uint32_t t = wave_sum_u32(active ? cs[k] : 0u);
                if (lane == 0) atomicAdd(J.counters + k, (unsigned long long)t);
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
// class (see rms_types.h).
__global__ void build_cells_kernel(const void *labels, int32_t label_bytes, const uint8_t *lclass, int32_t nx,
                                   int32_t ny, int32_t nz, uint64_t *cells, uint8_t *cls_out)
{
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t lin = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lin < n;
         lin += (int64_t)gridDim.x * blockDim.x) {
        const int32_t z = (int32_t)(lin % nz);
        const int64_t t = lin / nz;
        const int32_t y = (int32_t)(t % ny);
        const int32_t x = (int32_t)(t / ny);
        auto class_of = [&](int32_t xx, int32_t yy, int32_t zz) -> int32_t {
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

template <int NP>
static const void *walk_fn(int prec_f64, int sig)
{
    // sig bit 0: narrow-pulse b-table, bit 1: PGSE table, bit 2: exchange
    // faces on the fast path, bit 3: periodic lattice edges
    if (prec_f64) return walk_fn_sig<double, NP>(sig, std::make_integer_sequence<int, 16>{});
    return walk_fn_sig<float, NP>(sig, std::make_integer_sequence<int, 16>{});
}

static const void *walk_fn_np(int prec_f64, int sig, int np)
{
    switch (np) {
    case 0: return walk_fn<0>(prec_f64, sig);
    case 1: return walk_fn<1>(prec_f64, sig);
    case 2: return walk_fn<2>(prec_f64, sig);
    default: return walk_fn<3>(prec_f64, sig);
    }
}

static hipError_t launch_fn(const void *fn, const LaunchArgs &a, int grid, int block, hipStream_t s)
{
    void *args[] = {(void *)&a};
    return hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, s);
}

hipError_t launch_walk(const LaunchArgs &a, int prec_f64, int sig, int grid, int block, hipStream_t s)
{
    return launch_fn(walk_fn_np(prec_f64, sig, 0), a, grid, block, s);
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
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, walk_fn_np(prec_f64, sig, np < 1 ? 1 : np),
                                                        block, 0);
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
                              int32_t ny, int32_t nz, uint64_t *cells, uint8_t *cls, hipStream_t s)
{
    const int64_t n = (int64_t)nx * ny * nz;
    int64_t grid = (n + 255) / 256;
    if (grid > 65536) grid = 65536;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(build_cells_kernel, dim3((unsigned)grid), dim3(256), 0, s, labels, label_bytes, lclass, nx, ny,
                       nz, cells, cls);
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
