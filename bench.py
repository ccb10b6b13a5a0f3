"""Throughput benchmark: walker-steps/s of the RMS random walk on one MI355X.

Workload of the headline number (one "step" of this benchmark = one full
simulation run): a single synthetic beaded, undulating axon (r0 = 0.8 um,
caliber CV 0.2, undulation 0.5 um) voxelised at 0.1 um in a 64x64x1000 uint8
periodic lattice; 1e6 intra-axonal walkers, D0 = 2 um^2/ms, dt = 2e-4 ms, 1e5
steps; PGSE table of 30 directions x b in {0.5,1,2,3} ms/um^2 x Delta in
{10,20,40} ms at delta = 5 ms (360 rows); FP32 in-voxel offsets with FP64
moment accumulation; moments recorded every 100 steps (1000 records, as the
reference's main.cu).

Prints one JSON line with the throughput, the kernel's position against the
VALU roofline and the CPU baseline (the oracle, FP64 reference semantics,
built for this host and run on all its cores, on bounded samples of every
workload).

Beside the headline the line carries ("forms") every workload of the metric,
each with the standard error of D(t_end) (and K, S) from 32 batch-means
replicas and the cost per geometry of reaching a 1 % standard error of
D(t_end) on this GPU (walkers needed x steps / rate), the figure that puts the
forms side by side at matched Monte Carlo error:
  plain         the headline run (periodic lattice edges)
  plain_mirror  the same with the reference's mirror edges (main.cu:246-265)
  planes        impermeable planes 5 um apart, 32x32x52 uint8, x/y periodic,
                1e4 walkers x 1e4 steps, records every 100 steps
  batched       256 axons with extra-axonal shells (gamma radii mean 0.6 um,
                CV 0.4, lengths 20-200 um, uneven uint16 lattices), kappa in
                [0, 0.02] um/ms, T2 80/60 ms, D 2.0/1.5 um^2/ms, 1e7 walkers,
                5e4 steps, the PGSE table, one persistent launch with the
                autotuned shape and order, replayed as a captured hipGraph
  gradient      the axon with extra-axonal space, kappa 0.01 um/ms, 1e6
                walkers, 5e4 steps, 20 records, narrow-pulse 64 directions x b
                in {0.5,1,2,3}: forward pass plus the adjoint pass of
                sum_t [D_t(z) + K_t(z) + sum_j S_j(t)]
  sweep         the parameter sweep (caliber CV, undulation and kappa drawn per
                geometry, 70x70x1000 uint16 lattices, 2e4 steps, 20 records,
                batched forward plus adjoint) on its first --sweep-geoms of the
                4096 geometries (default 512) with the full run's 9766 walkers
                per geometry; --sweep-geoms 4096 runs all of it (~40 GB of
                labels, a few minutes of setup)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# VALU peak of MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz (non-packed lane ops/s)
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the walk kernel on the headline workload, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE passes scaled by a 1 GiB device copy (x2.0 reads, x1.0 writes); the summaries with the
# commands are in profiles/.  Keyed by (precision, walkers, steps, periodic axes).
MEASURED_TRAFFIC = {("f32", 1_000_000, 100_000, ""): 2.007e10, ("f32", 1_000_000, 100_000, "xyz"): 2.120e10}


def ops_per_step(s, pgse=True):
    """Algorithmic lane-ops per walker-step: Philox block (98) + u01 (6) +
    direction (17) + reciprocals (12) + 28 per face iteration (1+1.5s) +
    6 per crossing (1.5s) + bookkeeping (3); PGSE window sums add 7."""
    return 136 + 28 * (1 + 1.5 * s) + 9 * s + (7 if pgse else 0)


def _jackknife(fn, rec_rep, sig_rep=None):
    """Delete-one jackknife over the batch-means replicas of fn(rec, sig)."""
    R = rec_rep.shape[0]
    rt = rec_rep.sum(0)
    st = None if sig_rep is None else sig_rep.sum(0)
    est = np.asarray(fn(rt, st))
    loo = np.stack([np.asarray(fn(rt - rec_rep[r], None if st is None else st - sig_rep[r])) for r in range(R)])
    se = np.sqrt((R - 1) / R * np.sum((loo - loo.mean(0)) ** 2, axis=0))
    return est, se


def _stat_summary(A, r, dt, TN, T, nclass, ngrad, axis=(0.0, 0.0, 1.0)):
    """D and K along `axis` at the last record and the first PGSE signals
    S = sig / sig0(t=0), with jackknife standard errors."""
    TD = A.diff_times(dt, TN, T)

    def fn(rec, sig):
        s = A.split(rec, nclass)
        K, D = A.akc_mom(s["sig0"], s["dx2"], s["dx4"], TD, list(axis))
        out = [D[-1, 0], K[-1, 0]]
        if sig is not None:
            out += list(sig[:3] / rec[0, 0])
        return np.array(out)

    est, se = _jackknife(fn, r["rec_rep"], r.get("sig_rep") if ngrad else None)
    out = {"D_end": float(est[0]), "D_end_se": float(se[0]), "K_end": float(est[1]), "K_end_se": float(se[1])}
    if ngrad:
        out["S_first"] = [float(x) for x in est[2:]]
        out["S_first_se"] = [float(x) for x in se[2:]]
    return out


def _se_cost(rate, walkers_per_geom, TN, rel_se, target=0.01):
    """Seconds of this GPU per geometry to reach a relative standard error
    `target` of D(t_end): walkers needed N (rel_se / target)^2, times TN
    steps, at the form's walker-steps/s."""
    return float(walkers_per_geom * (rel_se / target) ** 2 * TN / rate)


def _geom(rms, args):
    """Geometry factory with the --storage choice of the cell words."""
    def make(*a, **k):
        return rms.Geometry(*a, storage=args.storage, **k)
    return make


def _hbm_bytes(eng, jobs):
    """Device bytes of the jobs' geometries and the storages chosen."""
    info = [eng.geometry_info(j["geom"]) for j in jobs]
    return int(sum(b for b, _ in info)), sorted({s for _, s in info})


def _timed_runs(plan, n, want_last=True, replicas_untimed=False):
    """n timed runs; the replica sums (for standard errors) come with the
    last one, or, with replicas_untimed, from one more run outside the timed
    region (a many-job plan's replica block is large: 256 jobs x 32 replicas
    x 1000 records x 25 values is 1.6 GB)."""
    import torch
    ks = []
    t0 = time.perf_counter()
    r = None
    for k in range(n):
        r = plan.execute(want_replicas=(want_last and not replicas_untimed and k == n - 1))
        ks.append(plan.timing["walk_ms"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if want_last and replicas_untimed:
        r = plan.execute(want_replicas=True)
    return r, el, float(np.mean(ks))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--walkers", type=int, default=1_000_000)
    ap.add_argument("--tn", type=int, default=100_000)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--storage", default="auto", choices=["auto", "dense", "bricks"],
                    help="device layout of the cell words (auto: dense while HBM has room, then bricks)")
    ap.add_argument("--cpu-configs", default="c1,c2,c3,c4,c5")
    ap.add_argument("--graph", action="store_true", help="replay the headline run as a captured hipGraph")
    ap.add_argument("--periodic", default="xyz", help="axes of the single-axon lattice whose edges wrap around "
                    "(the workload's periodic lattice); '' gives the reference's mirrors")
    ap.add_argument("--forms", default="plain_mirror,planes,batched,gradient,sweep",
                    help="comma list of forms to time beside the headline")
    ap.add_argument("--batched-walkers", type=int, default=10_000_000)
    ap.add_argument("--batched-tn", type=int, default=50_000)
    ap.add_argument("--grad-walkers", type=int, default=1_000_000)
    ap.add_argument("--grad-tn", type=int, default=50_000)
    ap.add_argument("--sweep-geoms", type=int, default=512)
    ap.add_argument("--sweep-walkers-per-geom", type=int, default=40_000_000 // 4096)
    ap.add_argument("--sweep-tn", type=int, default=20_000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "read_bytes": n * 4, "write_bytes": n * 4}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the walk shards by walker "
              "ranges with one FP64 all-reduce at the end; that multi-GPU driver is not part of this package.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import __graft_entry__ as ge

    rms = ge.load_package()
    from rms_amd import analysis as A

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    W = rms.workloads
    dt, D0, res = W.DT, W.C2["D0"], W.RES
    lab = W.c2_labels()
    grad = W.pgse_table_c2()
    TN = args.tn
    timepoints = min(1000, TN)
    s_vox = math.sqrt(6 * dt * D0) / res
    ops = ops_per_step(s_vox, pgse=True)
    eng = rms.Engine(0)
    jobs = W.c2_jobs(_geom(rms, args), args.walkers, args.periodic, lab)
    # the diagnostic counters (crossings, reflections, ...) are not part of
    # the reference's outputs: the timed runs skip them
    plan = eng.plan(jobs, dt, TN, timepoints=timepoints, precision=args.precision, replicas=32, use_graph=args.graph,
                    counters=False)
    for _ in range(args.warmup):
        plan.execute(want_replicas=False)
    torch.cuda.synchronize()
    walk_ms = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = plan.execute(want_replicas=(k == args.steps - 1))
        walk_ms.append(plan.timing["walk_ms"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    steps_total = float(args.walkers) * TN
    value = steps_total * args.steps / el
    kern_ms = float(np.mean(walk_ms))
    kern_rate = steps_total / (kern_ms / 1e3)
    achieved = kern_rate * ops / 1e12
    traffic = MEASURED_TRAFFIC.get((args.precision, args.walkers, TN, "".join(sorted(args.periodic))))
    res0 = r[0]
    TD = A.diff_times(dt, TN, timepoints)
    K, D = A.akc_mom(res0["sig0"], res0["dx2"], res0["dx4"], TD, [0, 0, 1])
    S = res0["sig"] / res0["rec"][-1, 0]
    edges = f"periodic {args.periodic}" if args.periodic else "mirror (reference)"
    plain_stats = _stat_summary(A, res0, dt, TN, timepoints, 2, grad.shape[0])
    line = {
        "metric": "walker-steps/s on 1 MI355X (plain vs batched vs +gradient) at fixed MC std-error",
        "value": value,
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {
            "workload": "single beaded undulating axon, 64x64x1000 uint8 lattice at 0.1 um, "
                        f"{args.walkers} walkers x {TN} steps, PGSE 360 rows, plain form",
            "walkers": args.walkers, "TN": TN, "dt_ms": dt, "D0": D0, "records": timepoints,
            "lattice_edges": edges,
            "positions": args.precision + " in-voxel offsets + int32 voxel", "accumulation": "fp64",
            "graph": bool(args.graph), "grid": plan.timing["grid"], "block": plan.timing["block"],
        },
        "roofline": {
            "bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tlane-op/s",
            "frac": achieved / VALU_PEAK_TOPS,
            "traffic": traffic,
            "ops_per_walker_step": ops, "kernel_ms_avg": kern_ms, "kernel_walker_steps_per_s": kern_rate,
            # the same launch against the HBM roofline: measured bytes / kernel time
            "hbm": None if traffic is None else {
                "achieved": traffic / (kern_ms / 1e3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": traffic / (kern_ms / 1e3) / 1e9 / HBM_PEAK_GBS},
        },
        "physics": {"D_axial_end": float(D[-1, 0]), "K_axial_end": float(K[-1, 0]),
                    "S_first_rows": [float(x) for x in S[:3]]},
    }
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    rel = plain_stats["D_end_se"] / plain_stats["D_end"]
    line["forms"] = {"plain": dict(walker_steps_per_s=value, kernel_ms=kern_ms, geometries=1,
                                   walkers_per_geometry=args.walkers, TN=TN, lattice_edges=edges,
                                   roofline_frac=achieved / VALU_PEAK_TOPS, ops_per_walker_step=ops,
                                   D_end_rel_se=rel,
                                   seconds_per_geometry_at_1pct_se=_se_cost(value, args.walkers, TN, rel),
                                   **plain_stats)}
    plan.close()
    print(f"bench: plain {value:.4g} walker-steps/s", file=sys.stderr, flush=True)
    form_fns = {"plain_mirror": form_plain_mirror, "planes": form_planes, "batched": form_batched,
                "gradient": form_gradient, "sweep": form_sweep}
    for f in forms:
        if f in form_fns:
            line["forms"][f] = form_fns[f](rms, eng, args, A)
            print(f"bench: {f} done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args, line)
    print(json.dumps(line), flush=True)
    eng.close()


def form_plain_mirror(rms, eng, args, A):
    """The headline workload with the reference's mirror edges."""
    W = rms.workloads
    dt, TN, T = W.DT, args.tn, min(1000, args.tn)
    plan = eng.plan(W.c2_jobs(_geom(rms, args), args.walkers, ""), dt, TN, timepoints=T, precision=args.precision,
                    replicas=32, counters=False)
    plan.execute(want_replicas=False)
    n = max(2, min(args.steps, 3))
    r, el, kms = _timed_runs(plan, n)
    plan.close()
    rate = float(args.walkers) * TN * n / el
    st = _stat_summary(A, r[0], dt, TN, T, 2, 360)
    ops = ops_per_step(math.sqrt(6 * dt * W.C2["D0"]) / W.RES)
    krate = float(args.walkers) * TN / (kms / 1e3)
    rel = st["D_end_se"] / st["D_end"]
    return dict(walker_steps_per_s=rate, kernel_ms=kms, kernel_walker_steps_per_s=krate, geometries=1,
                walkers_per_geometry=args.walkers, TN=TN, lattice_edges="mirror (reference)",
                roofline_frac=krate * ops / 1e12 / VALU_PEAK_TOPS, D_end_rel_se=rel,
                seconds_per_geometry_at_1pct_se=_se_cost(rate, args.walkers, TN, rel), **st)


def form_planes(rms, eng, args, A):
    """Impermeable planes at full size (1e4 walkers x 1e4 steps, FP32 offsets):
    a latency-bound launch of 157 wavefronts; D_perp, D_par and K_perp at the
    last record."""
    W = rms.workloads
    c = W.C1
    plan = eng.plan(W.c1_jobs(_geom(rms, args)), W.DT, c["TN"], timepoints=c["timepoints"], precision=args.precision,
                    replicas=32, counters=False)
    plan.execute(want_replicas=False)
    n = max(3, args.steps)
    r, el, kms = _timed_runs(plan, n)
    plan.close()
    ws = float(c["nwalk"]) * c["TN"]
    st_perp = _stat_summary(A, r[0], W.DT, c["TN"], c["timepoints"], 2, 0, axis=(0.0, 0.0, 1.0))
    st_par = _stat_summary(A, r[0], W.DT, c["TN"], c["timepoints"], 2, 0, axis=(1.0, 0.0, 0.0))
    rate = ws * n / el
    rel = st_perp["D_end_se"] / st_perp["D_end"]
    return dict(walker_steps_per_s=rate, kernel_ms=kms, kernel_walker_steps_per_s=ws / (kms / 1e3), geometries=1,
                walkers_per_geometry=c["nwalk"], TN=c["TN"], records=c["timepoints"],
                lattice="32x32x52 uint8, label-0 planes at z = 0, 51 (gap 5.0 um), x/y periodic",
                D_perp_end=st_perp["D_end"], D_perp_end_se=st_perp["D_end_se"], K_perp_end=st_perp["K_end"],
                K_perp_end_se=st_perp["K_end_se"], D_par_end=st_par["D_end"], D_par_end_se=st_par["D_end_se"],
                D_end_rel_se=rel, seconds_per_geometry_at_1pct_se=_se_cost(rate, c["nwalk"], c["TN"], rel))


def form_batched(rms, eng, args, A):
    """Bundle of 256 axons in one persistent launch (autotuned shape and
    order), replayed as a captured hipGraph."""
    from rms_amd import autotune
    W = rms.workloads
    dt, TN = W.DT, args.batched_tn
    T = min(1000, TN)
    jobs = W.c3_jobs(_geom(rms, args), args.batched_walkers)
    nw = np.array([j["nwalk"] for j in jobs])
    t0 = time.perf_counter()
    best = autotune.tune(eng, jobs, dt, timepoints=10, precision=args.precision, probe_steps=200, reps=1,
                         use_cache=False)
    tune_s = time.perf_counter() - t0
    ordered = [jobs[k] for k in best["order"]]
    plan = eng.plan(ordered, dt, TN, timepoints=T, precision=args.precision, replicas=32,
                    block_threads=best["block_threads"], grid_blocks=best["grid_blocks"],
                    xcd_split=best["xcd_split"], use_graph=True, counters=False)
    plan.execute(want_replicas=False)
    n = max(1, min(args.steps, 2))
    r, el, kms = _timed_runs(plan, n, replicas_untimed=True)
    plan.close()
    total = float(nw.sum()) * TN
    hbm, stores = _hbm_bytes(eng, jobs)
    stats = [_stat_summary(A, rr, dt, TN, T, 3, 360) for rr in r]
    rel = np.array([s["D_end_se"] / s["D_end"] for s in stats])
    s = np.array([math.sqrt(6 * dt * 2.0) / W.RES, math.sqrt(6 * dt * 1.5) / W.RES])
    ops = ops_per_step(float(s.mean()))  # the two compartments' steps, equal weight
    rate = total * n / el
    return {
        "walker_steps_per_s": rate, "kernel_ms": kms,
        "kernel_walker_steps_per_s": total / (kms / 1e3), "graph": True,
        "roofline_frac_estimate": total / (kms / 1e3) * ops / 1e12 / VALU_PEAK_TOPS,
        "geometries": len(jobs), "walkers": int(nw.sum()), "walkers_per_geometry_median": float(np.median(nw)),
        "TN": TN, "records": T, "voxels": int(sum(j["geom"].labels.size for j in jobs)), "label_bytes": 2,
        "lattice_bytes_in_hbm": hbm, "storage": stores,
        "autotune": {k: best[k] for k in ("block_threads", "grid_blocks", "xcd_split", "pack")},
        "autotune_s": tune_s,
        "geometry0": stats[best["order"].index(0)],
        "D_end_rel_se_median": float(np.median(rel)),
        "seconds_per_geometry_at_1pct_se": float(np.median([_se_cost(rate, w, TN, e) for w, e in zip(nw, rel)])),
    }


def form_gradient(rms, eng, args, A):
    """Forward plus adjoint pass on the axon with extra-axonal space."""
    import torch
    W = rms.workloads
    c = W.C4
    dt, TN, T, Wn = W.DT, args.grad_tn, c["timepoints"], args.grad_walkers
    jobs = W.c4_jobs(_geom(rms, args), Wn)
    nb = int(jobs[0]["bval"].size)
    TD = A.diff_times(dt, TN, T)
    plan = eng.plan(jobs, dt, TN, timepoints=T, precision=args.precision, replicas=32, counters=False)

    def once(want):
        out = plan.execute(want_replicas=want)[0]
        cot = W.c4_cotangent(A, out["rec"], 3, TD, nb)
        adj = plan.adjoint([cot], None, want_replicas=want)[0]
        return out, adj

    once(False)
    torch.cuda.synchronize()
    nrun = max(1, min(args.steps, 2))
    fwd_ms, adj_ms = [], []
    t0 = time.perf_counter()
    for k in range(nrun):
        out, adj = once(k == nrun - 1)
        fwd_ms.append(plan.timing["walk_ms"])
        adj_ms.append(plan.adj_timing["walk_ms"])
    # walker-steps of the central differences' walks (the adjoint call's
    # total less its tangent pass)
    fd_ws = plan.adj_timing["walker_steps"] - float(Wn) * TN
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # the same cotangent through the boundary-term estimator of D (pathwise
    # plus the jumps, reservoir-sampled), once, for its values and cost
    cot = W.c4_cotangent(A, out["rec"], 3, TD, nb)
    t1 = time.perf_counter()
    bnd = plan.adjoint([cot], None, want_replicas=True, estimator="boundary")[0]
    bnd_s = time.perf_counter() - t1
    bnd_ms = plan.adj_timing["walk_ms"]
    plan.close()
    gr = adj["grad"]
    se = np.sqrt(32) * np.std(adj["grad_rep"], axis=0, ddof=1)
    se_b = np.sqrt(32) * np.std(bnd["grad_rep"], axis=0, ddof=1)
    names = {"D_intra": 1, "D_extra": 2, "T2_intra": 5, "T2_extra": 6, "kappa": 8}
    fin = lambda x: None if not np.isfinite(x) else float(x)  # noqa: E731

    def fS(rec, sig):
        s = A.split(rec, 3, nb)
        K, D = A.akc_mom(s["sig0"], s["dx2"], s["dx4"], TD, [0.0, 0.0, 1.0])
        S = A.narrow_pulse_signal(s["sig0"], s["sigRe"])
        return np.array([D[-1, 0], K[-1, 0], S[-1, c["ndir"]], S[-1, 2 * c["ndir"]]])

    est, ses = _jackknife(fS, out["rec_rep"])
    rate = float(Wn) * TN * nrun / el
    s = np.array([math.sqrt(6 * dt * 2.0) / W.RES, math.sqrt(6 * dt * 1.5) / W.RES])
    # per walker-step of the problem: the forward walk, the tangent pass (the
    # walk plus ~40 lane-ops of tangents) and the forward walks of the central
    # differences at D (1 +- h), on the walkers whose walks can change with D_c
    ops1 = ops_per_step(float(s.mean()), pgse=False)
    ops_all = 2 * ops1 + 40 + fd_ws / (float(Wn) * TN) * ops1
    kern = float(Wn) * TN / ((np.mean(fwd_ms) + np.mean(adj_ms)) / 1e3)
    rel = float(ses[0] / est[0])
    return {
        "walker_steps_per_s": rate, "kernel_walker_steps_per_s": kern,
        "forward_kernel_ms": float(np.mean(fwd_ms)), "adjoint_kernel_ms": float(np.mean(adj_ms)),
        "adjoint": "tangent pass (T2 analytic, kappa score, D pathwise) + central differences of D with common "
                   "random numbers at D (1 +- 0.05): 4 forward walks of the walkers whose walks can change with D_c",
        "central_difference_walker_steps": fd_ws,
        "central_difference_fraction_of_4_walks": fd_ws / (4.0 * Wn * TN),
        "roofline_frac_estimate": kern * ops_all / 1e12 / VALU_PEAK_TOPS,
        "geometries": 1, "walkers_per_geometry": Wn, "TN": TN, "records": T, "bvecs": nb, "kappa": c["kappa"],
        "objective": "sum_t [D_t(z) + K_t(z) + sum_j S_j(t)], narrow-pulse S = sigRe/sig0",
        "gradient": {k: fin(gr[v]) for k, v in names.items()},
        "gradient_se": {k: fin(se[v]) for k, v in names.items()},
        "boundary_estimator": {"D_intra": fin(bnd["grad"][1]), "D_intra_se": fin(se_b[1]),
                               "D_extra": fin(bnd["grad"][2]), "D_extra_se": fin(se_b[2]),
                               "adjoint_kernel_ms": float(bnd_ms), "seconds": bnd_s},
        "D_end": float(est[0]), "D_end_se": float(ses[0]), "K_end": float(est[1]), "K_end_se": float(ses[1]),
        "S_end_b1_first": float(est[2]), "S_end_b1_first_se": float(ses[2]),
        "S_end_b2_first": float(est[3]), "S_end_b2_first_se": float(ses[3]),
        "D_end_rel_se": rel, "seconds_per_geometry_at_1pct_se": _se_cost(rate, Wn, TN, rel),
    }


def form_sweep(rms, eng, args, A):
    """Parameter sweep: the first --sweep-geoms of the 4096 geometries, each
    with the full run's walkers per geometry, in one batched forward launch
    plus its adjoint."""
    import torch
    W = rms.workloads
    c = W.C5
    ng, nwg, TN, T = args.sweep_geoms, args.sweep_walkers_per_geom, args.sweep_tn, c["timepoints"]
    t0 = time.perf_counter()
    jobs = []
    for k0 in range(0, ng, 64):
        for jb in W.c5_jobs(_geom(rms, args), range(k0, min(ng, k0 + 64)), nwg):
            eng.add_geometry(jb["geom"])
            jb["geom"].labels = jb["geom"].labels[:0, :0, :0]  # the device holds the lattice now
            jobs.append(jb)
        print(f"sweep: {len(jobs)}/{ng} geometries uploaded ({time.perf_counter() - t0:.0f} s)", file=sys.stderr,
              flush=True)
    setup_s = time.perf_counter() - t0
    hbm, stores = _hbm_bytes(eng, jobs)
    TD = A.diff_times(W.DT, TN, T)
    nb = int(jobs[0]["bval"].size)
    plan = eng.plan(jobs, W.DT, TN, timepoints=T, precision=args.precision, replicas=32, counters=False)
    t0 = time.perf_counter()
    outs = plan.execute(want_replicas=True)
    cots = [W.c4_cotangent(A, o["rec"], 3, TD, nb) for o in outs]
    adj = plan.adjoint(cots, None)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    fwd_ms, adj_ms = plan.timing["walk_ms"], plan.adj_timing["walk_ms"]
    fd_ws = plan.adj_timing["walker_steps"] - float(nwg) * ng * TN
    plan.close()
    dk = np.array([a["grad"][8] for a in adj])
    dT2 = np.array([a["grad"][5] for a in adj])
    dD = np.array([a["grad"][1] for a in adj])
    stats = [_stat_summary(A, o, W.DT, TN, T, 3, 0) for o in outs]
    Dend = np.array([s["D_end"] for s in stats])
    rel = np.array([s["D_end_se"] / s["D_end"] for s in stats])
    total = float(nwg) * ng * TN
    rate = total / el
    return {
        "walker_steps_per_s": rate, "forward_walker_steps_per_s": total / (fwd_ms / 1e3),
        "forward_kernel_ms": float(fwd_ms), "adjoint_kernel_ms": float(adj_ms),
        "geometries": ng, "of_geometries": c["ngeom"], "walkers_per_geometry": nwg, "walkers": int(nwg * ng),
        "TN": TN, "records": T, "bvecs": nb, "voxels": ng * 70 * 70 * 1000, "label_bytes": 2,
        "lattice_bytes_in_hbm": hbm, "storage": stores,
        "lattice_bytes_per_voxel": hbm / (ng * 70 * 70 * 1000), "setup_s": setup_s,
        "D_end_median": float(np.median(Dend)), "D_end_range": [float(np.min(Dend)), float(np.max(Dend))],
        "dobj_dkappa_median": float(np.nanmedian(dk)), "dobj_dT2_intra_median": float(np.median(dT2)),
        "dobj_dD_intra_median": float(np.nanmedian(dD)),
        "adjoint": "tangent pass + central differences of D with common random numbers (4 forward walks of the "
                   "walkers whose walks can change with D_c)",
        "central_difference_fraction_of_4_walks": fd_ws / (4.0 * nwg * ng * TN),
        "D_end_rel_se_median": float(np.median(rel)),
        "seconds_per_geometry_at_1pct_se": _se_cost(rate, nwg, TN, float(np.median(rel))),
    }


def _cpu_threads():
    """CPUs this process may use: its affinity mask, capped by a cgroup quota
    and by OMP_NUM_THREADS when those are set."""
    n = len(os.sched_getaffinity(0))
    caps = {"affinity": n}
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            caps["cgroup_quota"] = float(q) / float(per)
            n = min(n, max(1, int(float(q) / float(per))))
    except (OSError, ValueError):
        pass
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        caps["OMP_NUM_THREADS"] = int(env)
        n = min(n, int(env))
    return n, caps


def cpu_baseline(args, line):
    """The oracle (FP64 reference semantics, OpenMP over every batch of every
    geometry), built with -march=native for this host and run in a child
    process with OMP_PROC_BIND=close and OMP_PLACES=cores on every CPU this
    process may use, on bounded samples of every workload (oracle/cpu_baseline.py)."""
    nth, caps = _cpu_threads()
    native = os.path.join(ROOT, "oracle", "_native", "liboracle_rms.so")
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "native"], capture_output=True,
                       text=True, timeout=300)
    lib, flags = (native, "-O3 -march=native -fopenmp -ffp-contract=off") if r.returncode == 0 else \
        (os.path.join(ROOT, "oracle", "liboracle_rms.so"), "-O3 -march=x86-64-v3 -fopenmp -ffp-contract=off "
         "(native build failed)")
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    t0 = time.perf_counter()
    p = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), "--threads", str(nth),
                        "--lib", lib, "--c2-periodic", args.periodic, "--configs", args.cpu_configs],
                       capture_output=True, text=True, timeout=900, env=env)
    if p.returncode != 0:
        return {"value": None, "unit": "walker-steps/s", "cores": nth, "kind": "port",
                "sample": "CPU baseline failed: " + p.stderr[-400:]}
    d = json.loads(p.stdout.strip().splitlines()[-1])
    pc = d["per_config"]
    gpu = {"c1": line["forms"].get("planes"), "c2": line["forms"]["plain"], "c3": line["forms"].get("batched"),
           "c4": line["forms"].get("gradient"), "c5": line["forms"].get("sweep")}
    ratios = {k: (gpu[k]["walker_steps_per_s"] / v["value"]) for k, v in pc.items() if gpu.get(k)}
    c2 = pc.get("c2")
    return {
        "value": None if c2 is None else c2["value"], "unit": "walker-steps/s", "cores": nth, "kind": "port",
        "sample": None if c2 is None else c2["sample"] + f", FP64, {c2['seconds']:.1f} s",
        "cpu_model": d["cpu_model"], "threads_limits": caps, "build_flags": flags,
        "omp": {"OMP_PROC_BIND": d["omp_proc_bind"], "OMP_PLACES": d["omp_places"]},
        "per_config": pc, "gpu_over_cpu": ratios, "target_gpu_over_cpu_c2": 100,
        "seconds_total": time.perf_counter() - t0,
    }


if __name__ == "__main__":
    main()
