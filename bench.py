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
    geom = rms.Geometry(lab, res, [0.0, D0], T2=[math.inf, math.inf])
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
            "positions": args.precision + " in-voxel offsets + int32 voxel", "accumulation": "fp64",
            "graph": bool(args.graph), "grid": plan.timing["grid"], "block": plan.timing["block"],
        },
        "roofline": {
            "bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "Tlane-op/s",
            "frac": achieved / VALU_PEAK_TOPS, "traffic": None,
            "ops_per_walker_step": ops, "kernel_ms_avg": kern_ms, "kernel_walker_steps_per_s": kern_rate,
        },
        "physics": {"D_axial_end": float(D[-1, 0]), "K_axial_end": float(K[-1, 0]),
                    "S_first_rows": [float(x) for x in S[:3]], "counters": counters},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lab, res, D0, dt, grad, args)
    print(json.dumps(line), flush=True)
    plan.close()
    eng.close()


def cpu_baseline(lab, res, D0, dt, grad, args):
    """The oracle (FP64 reference semantics, OpenMP over 64-walker batches) on
    a walker subset of the same workload, timed on this host."""
    from oracle import oracle as orc

    nthreads = min(16, os.cpu_count() or 1)
    tn = args.cpu_tn or args.tn
    g = orc.Geometry(lab, res, [0.0, D0])
    job = dict(geom=g, nwalk=args.cpu_walkers, seed=0x5EED0001, grad=grad)
    t0 = time.perf_counter()
    orc.run([job], dt, tn, min(1000, tn), prec="f64", replicas=32, nthreads=nthreads)
    el = time.perf_counter() - t0
    return {"value": args.cpu_walkers * tn / el, "unit": "walker-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"{args.cpu_walkers} walkers x {tn} steps of the same axon/PGSE workload, FP64, "
                      f"{el:.1f} s"}


if __name__ == "__main__":
    main()
