"""Throughput benchmark: walker-steps/s of the RMS random walk on one MI355X.

Workload (one "step" of this benchmark = one full simulation run): a single
synthetic beaded, undulating axon (r0 = 0.8 um, caliber CV 0.2, undulation
0.5 um) voxelised at 0.1 um in a 64x64x1000 uint8 lattice; 1e6 intra-axonal
walkers, D0 = 2 um^2/ms, dt = 2e-4 ms, 1e5 steps; PGSE table of 30
directions x b in {0.5,1,2,3} ms/um^2 x Delta in {10,20,40} ms at delta = 5 ms
(360 rows); FP32 in-voxel offsets with FP64 moment accumulation; moments
recorded every 100 steps (1000 records, as the reference's main.cu).

Prints one JSON line with the throughput, the kernel's position against the
VALU roofline and the CPU baseline (the oracle, FP64 reference semantics, on
a bounded walker subset of the same workload, timed on this host).

Beside the plain form the line carries ("forms") the two other forms of the
metric, each with the standard errors of D(t_end), K(t_end) and S from 32
batch-means replicas:
  batched   256 axons with extra-axonal shells (gamma radii mean 0.6 um,
            CV 0.4, lengths 20-200 um, uneven uint16 lattices), kappa drawn in
            [0, 0.02] um/ms, T2 80/60 ms, D 2.0/1.5 um^2/ms, 1e7 walkers in
            total, 5e4 steps, the PGSE table above, one persistent launch with
            the autotuned launch shape and geometry order;
  gradient  the axon above with extra-axonal space (label 2 outside), kappa
            0.01 um/ms, 1e6 walkers, 5e4 steps, 20 records, narrow-pulse
            b-table of 64 directions x b in {0.5,1,2,3} ms/um^2: forward pass
            plus the adjoint pass of S(b, g, t) and D(t), K(t) along z.
On request (--forms sweep, not in the default run) the parameter sweep: 4096
axons with extra-axonal space in 70x70x1000 uint16 lattices (caliber CV,
undulation and kappa drawn per geometry, ~20e9 voxels, ~180 GB of cell words
and classes in HBM), 4e7 walkers, 2e4 steps, batched forward plus adjoint.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# VALU peak of MI355X: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz (non-packed lane ops/s)
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the walk kernel on this workload, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE passes scaled by a 1 GiB device copy (x2.0 reads, x1.0 writes); the summary with the
# commands is profiles/walk_kernel_axon_pgse_f32.md.  Keyed by (precision, walkers, steps).
MEASURED_TRAFFIC = {("f32", 1_000_000, 100_000): 2.007e10}


def ops_per_step(s, pgse=True):
    """Algorithmic lane-ops per walker-step: Philox block (98) + u01 (6) +
    direction (17) + reciprocals (12) + 28 per face iteration (1+1.5s) +
    6 per crossing (1.5s) + bookkeeping (3); PGSE window sums add 7."""
    return 136 + 28 * (1 + 1.5 * s) + 9 * s + (7 if pgse else 0)


def pgse_table_c2():
    """30 directions x 4 b-values x 3 Delta at delta = 5 ms (rows Delta,
    delta, |g|, gx, gy, gz; b = g^2 delta^2 (Delta - delta/3))."""
    n = 30
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    dirs = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    rows = []
    delta = 5.0
    for Delta in (10.0, 20.0, 40.0):
        for b in (0.5, 1.0, 2.0, 3.0):
            g = math.sqrt(b / (delta ** 2 * (Delta - delta / 3)))
            for d in dirs:
                rows.append([Delta, delta, g, *d])
    return np.array(rows)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--walkers", type=int, default=1_000_000)
    ap.add_argument("--tn", type=int, default=100_000)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--cpu-walkers", type=int, default=10_000)
    ap.add_argument("--cpu-tn", type=int, default=0, help="CPU baseline steps (0 = same as --tn)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true", help="replay the run as a captured hipGraph")
    ap.add_argument("--periodic", default="", help="axes of the single-axon lattice whose edges wrap around "
                    "(default: the reference's mirrors)")
    ap.add_argument("--forms", default="plain,batched,gradient",
                    help="comma list of forms to time beside the plain form (also: sweep)")
    ap.add_argument("--batched-walkers", type=int, default=10_000_000)
    ap.add_argument("--batched-tn", type=int, default=50_000)
    ap.add_argument("--grad-walkers", type=int, default=1_000_000)
    ap.add_argument("--grad-tn", type=int, default=50_000)
    ap.add_argument("--sweep-geoms", type=int, default=4096)
    ap.add_argument("--sweep-walkers", type=int, default=40_000_000)
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
              "ranges with one FP64 all-reduce at the end, which this build does not run.", file=sys.stderr)
        sys.exit(2)

    import torch

    import __graft_entry__ as ge

    rms = ge.load_package()
    from rms_amd import geometry as G

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dt, D0, res = 2e-4, 2.0, 0.1
    lab = G.axon(n=64, nz=1000, res=res, r0=0.8, cv=0.2, undulation=0.5, seed=1235)
    grad = pgse_table_c2()
    TN = args.tn
    timepoints = min(1000, TN)
    geom = rms.Geometry(lab, res, [0.0, D0], T2=[math.inf, math.inf], periodic=args.periodic)
    s_vox = math.sqrt(6 * dt * D0) / res
    ops = ops_per_step(s_vox, pgse=True)
    job = dict(geom=geom, nwalk=args.walkers, seed=0x5EED0001, grad=grad)
    eng = rms.Engine(0)
    plan = eng.plan([job], dt, TN, timepoints=timepoints, precision=args.precision, replicas=32,
                    use_graph=args.graph)
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
    traffic = MEASURED_TRAFFIC.get((args.precision, args.walkers, TN))
    res0 = r[0]
    counters = res0["counters"].tolist()
    # physics summary of the last run (SE from 32 batch-means replicas)
    from rms_amd import analysis as A
    TD = A.diff_times(dt, TN, timepoints)
    K, D = A.akc_mom(res0["sig0"], res0["dx2"], res0["dx4"], TD, [0, 0, 1])
    S = res0["sig"] / res0["rec"][-1, 0]
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
            "lattice_edges": f"periodic {args.periodic}" if args.periodic else "mirror (reference)",
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
                    "S_first_rows": [float(x) for x in S[:3]], "counters": counters},
    }
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    line["forms"] = {"plain": dict(walker_steps_per_s=value, kernel_ms=kern_ms,
                                   **_stat_summary(A, res0, dt, TN, timepoints, 2, grad.shape[0]))}
    plan.close()
    if "batched" in forms:
        line["forms"]["batched"] = form_batched(rms, eng, args, dt, grad)
    if "gradient" in forms:
        line["forms"]["gradient"] = form_gradient(rms, eng, args, dt)
    if "sweep" in forms:
        line["forms"]["sweep"] = form_sweep(rms, eng, args, dt)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lab, res, D0, dt, grad, args)
    print(json.dumps(line), flush=True)
    eng.close()


def _jackknife(fn, rec_rep, sig_rep=None):
    """Delete-one jackknife over the batch-means replicas of fn(rec, sig)."""
    R = rec_rep.shape[0]
    rt = rec_rep.sum(0)
    st = None if sig_rep is None else sig_rep.sum(0)
    est = np.asarray(fn(rt, st))
    loo = np.stack([np.asarray(fn(rt - rec_rep[r], None if st is None else st - sig_rep[r])) for r in range(R)])
    se = np.sqrt((R - 1) / R * np.sum((loo - loo.mean(0)) ** 2, axis=0))
    return est, se


def _stat_summary(A, r, dt, TN, T, nclass, ngrad):
    """D and K along z at the last record and the first PGSE signals S =
    sig / sig0(t=0), with jackknife standard errors."""
    TD = A.diff_times(dt, TN, T)

    def fn(rec, sig):
        s = A.split(rec, nclass)
        K, D = A.akc_mom(s["sig0"], s["dx2"], s["dx4"], TD, [0.0, 0.0, 1.0])
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


def form_batched(rms, eng, args, dt, grad):
    """Bundle of 256 axons in one persistent launch (autotuned shape/order)."""
    import torch
    from rms_amd import analysis as A
    from rms_amd import autotune
    from rms_amd import geometry as G

    labs = G.bundle(256, mean_r=0.6, cv_r=0.4, len_lo=20.0, len_hi=200.0, res=0.1, seed=1236)
    rng = np.random.default_rng(2024)
    kap = rng.uniform(0.0, 0.02, len(labs))
    walkable = np.array([np.count_nonzero(l) for l in labs], float)
    nw = np.maximum(64, np.round(args.batched_walkers * walkable / walkable.sum())).astype(np.int64)
    jobs = [dict(geom=rms.Geometry(l, 0.1, [0.0, 2.0, 1.5], T2=[math.inf, 80.0, 60.0], kappa=float(k)),
                 nwalk=int(w), seed=0x5EED0002 | (i << 32), grad=grad) for i, (l, k, w) in enumerate(zip(labs, kap, nw))]
    TN = args.batched_tn
    T = min(1000, TN)
    t0 = time.perf_counter()
    best = autotune.tune(eng, jobs, dt, timepoints=10, precision=args.precision, probe_steps=200, reps=1,
                         use_cache=False)
    tune_s = time.perf_counter() - t0
    ordered = [jobs[k] for k in best["order"]]
    plan = eng.plan(ordered, dt, TN, timepoints=T, precision=args.precision, replicas=32,
                    block_threads=best["block_threads"], grid_blocks=best["grid_blocks"],
                    xcd_split=best["xcd_split"])
    plan.execute(want_replicas=False)
    torch.cuda.synchronize()
    n = max(1, min(args.steps, 2))
    ks = []
    t0 = time.perf_counter()
    for k in range(n):
        r = plan.execute(want_replicas=(k == n - 1))
        ks.append(plan.timing["walk_ms"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    plan.close()
    total = float(nw.sum()) * TN
    stats = [_stat_summary(A, rr, dt, TN, T, 3, grad.shape[0]) for rr in r]
    rel = np.array([s["D_end_se"] / s["D_end"] for s in stats])
    return {
        "walker_steps_per_s": total * n / el, "kernel_ms": float(np.mean(ks)),
        "kernel_walker_steps_per_s": total / (np.mean(ks) / 1e3),
        "geometries": len(jobs), "walkers": int(nw.sum()), "TN": TN, "records": T,
        "voxels": int(sum(l.size for l in labs)), "label_bytes": 2,
        "autotune": {k: best[k] for k in ("block_threads", "grid_blocks", "xcd_split", "pack")},
        "autotune_s": tune_s,
        "geometry0": stats[best["order"].index(0)] if 0 in best["order"] else stats[0],
        "D_end_rel_se_median": float(np.median(rel)),
    }


def form_gradient(rms, eng, args, dt):
    """Forward plus adjoint pass on the axon with extra-axonal space."""
    import torch
    from rms_amd import analysis as A
    from rms_amd import geometry as G

    lab = G.axon(n=64, nz=1000, res=0.1, r0=0.8, cv=0.2, undulation=0.5, seed=1235, ecs=True)
    n = 64
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    dirs = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    bvals = np.repeat([0.5, 1.0, 2.0, 3.0], n)
    bvecs = np.tile(dirs, (4, 1))
    nb = bvals.size
    geom = rms.Geometry(lab, 0.1, [0.0, 2.0, 1.5], T2=[math.inf, 80.0, 60.0], kappa=0.01)
    TN, T, W = args.grad_tn, 20, args.grad_walkers
    TD = A.diff_times(dt, TN, T)
    plan = eng.plan([dict(geom=geom, nwalk=W, seed=0x5EED0003, bval=bvals, bvec=bvecs)], dt, TN, timepoints=T,
                    precision=args.precision, replicas=32)
    i_re = 1 + 3 + 21

    def once(want):
        out = plan.execute(want_replicas=want)[0]
        rec = out["rec"]
        # objective: sum_t [D_t(z) + K_t(z) + sum_j S_j(t)], S = sigRe / sig0
        cot = A.akc_cotangent(rec, 3, TD, [0.0, 0.0, 1.0])
        cot[:, i_re:i_re + nb] += (1.0 / rec[:, 0])[:, None]
        cot[:, 0] += -rec[:, i_re:i_re + nb].sum(1) / rec[:, 0] ** 2
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
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    plan.close()
    gr = adj["grad"]
    se = np.sqrt(32) * np.std(adj["grad_rep"], axis=0, ddof=1)
    names = {"D_intra": 1, "D_extra": 2, "T2_intra": 5, "T2_extra": 6, "kappa": 8}
    fin = lambda x: None if not np.isfinite(x) else float(x)  # noqa: E731
    TDs = A.diff_times(dt, TN, T)

    def fS(rec, sig):
        s = A.split(rec, 3, nb)
        K, D = A.akc_mom(s["sig0"], s["dx2"], s["dx4"], TDs, [0.0, 0.0, 1.0])
        S = A.narrow_pulse_signal(s["sig0"], s["sigRe"])
        return np.array([D[-1, 0], K[-1, 0], S[-1, n], S[-1, 2 * n]])

    est, ses = _jackknife(fS, out["rec_rep"])
    return {
        "walker_steps_per_s": float(W) * TN * nrun / el,
        "forward_kernel_ms": float(np.mean(fwd_ms)), "adjoint_kernel_ms": float(np.mean(adj_ms)),
        "walkers": W, "TN": TN, "records": T, "bvecs": int(nb), "kappa": 0.01,
        "objective": "sum_t [D_t(z) + K_t(z) + sum_j S_j(t)], narrow-pulse S = sigRe/sig0",
        "gradient": {k: fin(gr[v]) for k, v in names.items()},
        "gradient_se": {k: fin(se[v]) for k, v in names.items()},
        "note": "D entries are undefined (NaN) with membrane exchange: a class change at a crossing is a jump "
                "the pathwise derivative does not contain",
        "D_end": float(est[0]), "D_end_se": float(ses[0]), "K_end": float(est[1]), "K_end_se": float(ses[1]),
        "S_end_b1_first": float(est[2]), "S_end_b1_first_se": float(ses[2]),
        "S_end_b2_first": float(est[3]), "S_end_b2_first_se": float(ses[3]),
    }


def form_sweep(rms, eng, args, dt):
    """Parameter sweep: many axons with extra-axonal space, caliber CV in
    [0, 0.4], undulation amplitude in [0, 1] um and kappa in [0, 0.03] um/ms
    drawn per geometry, 70x70x1000 uint16 lattices, all in one batched forward
    launch plus its adjoint (D(t), K(t) along z and narrow-pulse S per
    geometry, 20 records, 16 directions x b in {1, 2} ms/um^2)."""
    import torch
    from rms_amd import analysis as A
    from rms_amd import geometry as G

    ng = args.sweep_geoms
    rng = np.random.default_rng(1238)
    cv, und, kap = rng.uniform(0.0, 0.4, ng), rng.uniform(0.0, 1.0, ng), rng.uniform(0.0, 0.03, ng)
    n = 16
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    dirs = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    bvals, bvecs = np.repeat([1.0, 2.0], n), np.tile(dirs, (2, 1))
    nb = bvals.size
    nw = max(64, args.sweep_walkers // ng)
    TN, T = args.sweep_tn, 20
    jobs, voxels = [], 0
    t0 = time.perf_counter()
    for k in range(ng):
        lab = G.axon(n=70, nz=1000, res=0.1, r0=0.8, cv=float(cv[k]), undulation=float(und[k]), seed=1238 + k,
                     ecs=True, dtype=np.uint16)
        geom = rms.Geometry(lab, 0.1, [0.0, 2.0, 1.5], T2=[math.inf, 80.0, 60.0], kappa=float(kap[k]))
        eng.add_geometry(geom)
        voxels += lab.size
        jobs.append(dict(geom=geom, nwalk=nw, seed=0x5EED0004 | (k << 32), bval=bvals, bvec=bvecs))
        if (k + 1) % 256 == 0:
            print(f"sweep: {k + 1}/{ng} geometries uploaded ({time.perf_counter() - t0:.0f} s)", file=sys.stderr,
                  flush=True)
    setup_s = time.perf_counter() - t0
    TD = A.diff_times(dt, TN, T)
    i_re = 1 + 3 + 21
    plan = eng.plan(jobs, dt, TN, timepoints=T, precision=args.precision, replicas=32)
    t0 = time.perf_counter()
    outs = plan.execute(want_replicas=False)
    cots = []
    for o in outs:
        rec = o["rec"]
        cot = A.akc_cotangent(rec, 3, TD, [0.0, 0.0, 1.0])
        cot[:, i_re:i_re + nb] += (1.0 / rec[:, 0])[:, None]
        cot[:, 0] += -rec[:, i_re:i_re + nb].sum(1) / rec[:, 0] ** 2
        cots.append(cot)
    adj = plan.adjoint(cots, None)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    fwd_ms, adj_ms = plan.timing["walk_ms"], plan.adj_timing["walk_ms"]
    plan.close()
    dk = np.array([a["grad"][8] for a in adj])
    dT2 = np.array([a["grad"][5] for a in adj])
    Dend = []
    for o in outs:
        s = A.split(o["rec"], 3, nb)
        K, D = A.akc_mom(s["sig0"], s["dx2"], s["dx4"], TD, [0.0, 0.0, 1.0])
        Dend.append(D[-1, 0])
    total = float(nw) * ng * TN
    return {
        "walker_steps_per_s": total / el, "forward_walker_steps_per_s": total / (fwd_ms / 1e3),
        "forward_kernel_ms": float(fwd_ms), "adjoint_kernel_ms": float(adj_ms),
        "geometries": ng, "walkers": int(nw * ng), "TN": TN, "records": T, "bvecs": int(nb),
        "voxels": int(voxels), "label_bytes": 2, "lattice_bytes_in_hbm": int(voxels * 9),
        "setup_s": setup_s,
        "D_end_median": float(np.median(Dend)), "D_end_range": [float(np.min(Dend)), float(np.max(Dend))],
        "dobj_dkappa_median": float(np.median(dk)), "dobj_dT2_intra_median": float(np.median(dT2)),
    }


def cpu_baseline(lab, res, D0, dt, grad, args):
    """The oracle (FP64 reference semantics, OpenMP over 64-walker batches) on
    a walker subset of the same workload, timed on this host."""
    from oracle import oracle as orc

    nthreads = min(16, os.cpu_count() or 1)
    tn = args.cpu_tn or args.tn
    g = orc.Geometry(lab, res, [0.0, D0], periodic=args.periodic)
    job = dict(geom=g, nwalk=args.cpu_walkers, seed=0x5EED0001, grad=grad)
    t0 = time.perf_counter()
    orc.run([job], dt, tn, min(1000, tn), prec="f64", replicas=32, nthreads=nthreads)
    el = time.perf_counter() - t0
    return {"value": args.cpu_walkers * tn / el, "unit": "walker-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"{args.cpu_walkers} walkers x {tn} steps of the same axon/PGSE workload, FP64, "
                      f"{el:.1f} s"}


if __name__ == "__main__":
    main()
