"""The record rows' wavefront reduce-scatter (rec_tree / put_row in
csrc/rms_walk.h), modelled lane by lane in numpy: every position's total
equals the butterfly (wave_sum: pairs at lane distance 1, 2, 4, ..., 32) bit
for bit, and lane L < 32 ends with position bitrev5(L).  The inputs span many
binades, so that another pairing order gives other totals (checked too)."""
import numpy as np


def _butterfly(x, masks=(1, 2, 4, 8, 16, 32)):
    """x: [64 lanes, n values]; every lane ends with the tree total."""
    lanes = np.arange(64)
    for m in masks:
        x = x + x[lanes ^ m]
    return x


def _reduce_scatter(v):
    """v: [64 lanes, 32 positions] -> [64] the lane's kept total."""
    lanes = np.arange(64)
    slots = v.copy()
    for lv in range(1, 6):
        h = 32 >> lv
        lo, hi = slots[:, :h], slots[:, h:2 * h]
        up = ((lanes >> (lv - 1)) & 1).astype(bool)[:, None]
        send = np.where(up, lo, hi)
        keep = np.where(up, hi, lo)
        slots = keep + send[lanes ^ (1 << (lv - 1))]
    t = slots[:, 0]
    return t + t[lanes ^ 32]


def _bitrev5(x):
    return int(f"{x & 31:05b}"[::-1], 2)


def test_reduce_scatter_equals_the_butterfly():
    rng = np.random.default_rng(5)
    for trial in range(20):
        v = rng.standard_normal((64, 32)) * np.exp2(rng.integers(-30, 30, (64, 32)))
        v[rng.random((64, 32)) < 0.2] = 0.0      # inactive lanes add +0.0
        v[:, 26:] = 0.0                           # padding positions
        ref = _butterfly(v)[0]
        got = _reduce_scatter(v)
        for lane in range(64):
            p = _bitrev5(lane)
            assert got[lane].tobytes() == ref[p].tobytes(), (trial, lane, p)
        # the inputs tell pairing orders apart: the reversed order differs
        other = _butterfly(v, masks=(32, 16, 8, 4, 2, 1))[0]
        assert np.any(other != ref)


def test_store_offsets_cover_the_row_once():
    """Lanes 0..31 write position bitrev5(L) to its place: weight at 0, the
    class indicators of the geometry's classes at 1..nclass, the 21 moments
    after them; padding and absent classes are not written."""
    for nclass in range(1, 5):
        offs = []
        for lane in range(32):
            p = _bitrev5(lane)
            off = 0 if p == 0 else (p if p - 1 < nclass else -1) if p <= 4 else 1 + nclass + (p - 5) if p <= 25 else -1
            if off >= 0:
                offs.append(off)
        assert sorted(offs) == list(range(1 + nclass + 21))
