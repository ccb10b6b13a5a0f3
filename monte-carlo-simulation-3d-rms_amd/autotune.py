"""Launch-shape and packing search for the persistent walk kernel.

The walk is VALU-bound with a latency-bound label gather, so the knobs that
matter are occupancy (workgroup size x resident workgroups), the
batch -> XCD assignment (a geometry whose batches stay on one XCD keeps its
lattice in that XCD's L2) and, for many geometries, the order of jobs in the
batch list.  `tune` times short probe runs (the same jobs with few steps)
for each candidate and returns the fastest; results are cached per
workload signature in a JSON file.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

CACHE = os.path.join(os.path.expanduser("~"), ".cache", "rms_amd_autotune.json")


def signature(jobs, precision, sig_kind):
    h = hashlib.sha1()
    for jb in jobs:
        g = jb["geom"]
        h.update(str((g.labels.shape, g.labels.dtype.str, int(jb["nwalk"]) // 4096, g.nclass)).encode())
    h.update(f"{precision}:{sig_kind}".encode())
    return h.hexdigest()[:16]


def pack_order(jobs, mode="size_desc"):
    """Order of jobs in the batch list: 'input', 'size_desc' (largest lattice
    first, so each XCD's contiguous range holds few large lattices) or
    'walkers_desc'."""
    idx = list(range(len(jobs)))
    if mode == "size_desc":
        idx.sort(key=lambda k: -jobs[k]["geom"].labels.size)
    elif mode == "walkers_desc":
        idx.sort(key=lambda k: -int(jobs[k]["nwalk"]))
    return idx


def candidates(njobs):
    out = []
    for block in (64, 128, 256):
        for occ_mult in (1.0, 0.5):
            for xs in ((1, -1) if njobs > 1 else (-1,)):
                for pack in (("size_desc", "input") if njobs > 1 else ("input",)):
                    out.append(dict(block_threads=block, occ_mult=occ_mult, xcd_split=xs, pack=pack))
    return out


def tune(engine, jobs, dt, timepoints=10, precision="f32", probe_steps=400, reps=2, use_cache=True, verbose=False):
    """Return the best {block_threads, grid_blocks, xcd_split, pack} for the
    jobs.  Probe runs use `probe_steps` steps and the full walker counts."""
    sig_kind = ("n" if any(j.get("bval") is not None for j in jobs) else "") + \
               ("p" if any(j.get("grad") is not None for j in jobs) else "")
    key = signature(jobs, precision, sig_kind)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            cache = json.load(open(CACHE))
        except Exception:
            cache = {}
    if use_cache and key in cache:
        return cache[key]
    best = None
    results = []
    for cand in candidates(len(jobs)):
        order = pack_order(jobs, cand["pack"])
        ordered = [jobs[k] for k in order]
        # resident grid from a probe plan at full occupancy
        p = engine.plan(ordered, dt, probe_steps, timepoints=min(timepoints, probe_steps), precision=precision,
                        replicas=1, block_threads=cand["block_threads"], xcd_split=cand["xcd_split"], counters=False)
        p.execute(want_replicas=False)
        grid_full = p.timing["grid"]
        p.close()
        grid = max(8, int(grid_full * cand["occ_mult"]) // 8 * 8)
        p = engine.plan(ordered, dt, probe_steps, timepoints=min(timepoints, probe_steps), precision=precision,
                        replicas=1, block_threads=cand["block_threads"], grid_blocks=grid,
                        xcd_split=cand["xcd_split"], counters=False)
        ts = []
        for _ in range(reps):
            p.execute(want_replicas=False)
            ts.append(p.timing["walk_ms"])
        p.close()
        t = float(np.min(ts))
        rec = dict(cand, grid_blocks=grid, walk_ms=t)
        results.append(rec)
        if verbose:
            print("autotune", rec, flush=True)
        if best is None or t < best["walk_ms"]:
            best = rec
    best = dict(best)
    best["order"] = pack_order(jobs, best["pack"])
    if use_cache:
        cache[key] = best
        try:
            os.makedirs(os.path.dirname(CACHE), exist_ok=True)
            json.dump(cache, open(CACHE, "w"))
        except OSError:
            pass
    best["_all"] = results
    return best


def timed(fn, *a, **kw):
    t0 = time.perf_counter()
    r = fn(*a, **kw)
    return r, time.perf_counter() - t0
