"""The benchmark workloads on the GPU, against the oracle and closed forms.

Every test builds its geometry and tables with the same functions bench.py
uses (rms_amd.workloads), so the engine paths the throughput numbers are
quoted on are the paths checked here:

  C1  planes at full size (1e4 walkers x 1e4 steps), bit-exact against the
      oracle; D_perp(t) against the plane series, D_par = D0, and K_perp(t)
      against the eigen-expansion of the fourth moment
  C2  the single axon with its 360-row PGSE table on 1e4 walkers x 1e5 steps,
      mirror (reference) and periodic edges, bit-exact; a validity run at
      TN = 2.25e5 where every row has both lobes, bit-exact on 1024 walkers,
      and the periodic free-diffusion analogue against the Gaussian phase
  C3  16 of the 256 bundle geometries at 1 % of their walkers, launched with
      the autotuned shape and order, bit-exact, each job equal to itself run
      alone
  C4  the axon with extra-axonal space at kappa = 0.01, 1e4 walkers x 5e4
      steps, forward and adjoint bit-exact (the boundary-term estimator, and
      the default one: its central differences on walker subsets against the
      oracle's on its own marks)
  C5  8 of the 4096 sweep geometries (70x70x1000 uint16), 640 walkers each,
      forward and adjoint bit-exact, both estimators
"""
import math

import numpy as np
import pytest

from helpers import assert_same, gaussian_phase_signal, planes_moments

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def W(rms):
    return rms.workloads


def _jack(fn, rep):
    """Delete-one jackknife standard error over the replica axis."""
    R = rep.shape[0]
    tot = rep.sum(0)
    loo = np.stack([fn(tot - rep[r]) for r in range(R)])
    return np.sqrt((R - 1) / R * np.sum((loo - loo.mean(0)) ** 2, axis=0))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_c1_planes_full_size(rms, orc, engine, W, prec):
    c = W.C1
    print(f"case C1 planes 1e4 x 1e4 {prec}", flush=True)
    gpu = engine.run(W.c1_jobs(rms.Geometry), W.DT, c["TN"], c["timepoints"], precision=prec, replicas=32,
                     final_state=True)
    ref = orc.run(W.c1_jobs(orc.Geometry), W.DT, c["TN"], c["timepoints"], prec=prec, replicas=32, final_state=True)
    assert_same(gpu, ref, final_state=True)
    r = gpu[0]
    assert np.all(r["NPar_count"][:, 1] == c["nwalk"]) and np.all(r["NPar_count"][:, 0] == 0)
    assert r["counters"][1] > 0 and r["counters"][2] == 0  # wall reflections; x/y edges wrap, never mirror
    TD = orc.diff_times(W.DT, c["TN"], c["timepoints"])
    m2, m4 = planes_moments(c["gap_um"], c["D0"], TD)
    nc = 2
    i2, i4 = 1 + nc, 1 + nc + 6

    def stats(rec):
        s0 = rec[:, 0]
        dx2 = rec[:, i2:i2 + 6] / s0[:, None]
        dz4 = rec[:, i4 + 14] / s0
        Dpar = (dx2[:, 0] + dx2[:, 3]) / (4 * TD)
        return np.stack([dx2[:, 5], Dpar, dz4 / dx2[:, 5] ** 2 - 3], 1)

    est = stats(r["rec"])
    se = _jack(stats, r["rec_rep"])
    late = TD >= 0.2
    z_perp = (est[late, 0] - m2[late]) / se[late, 0]
    z_par = (est[late, 1] - c["D0"]) / se[late, 1]
    z_k = (est[late, 2] - (m4[late] / m2[late] ** 2 - 3)) / se[late, 2]
    assert np.max(np.abs(z_perp)) < 4.5, z_perp
    assert np.max(np.abs(z_par)) < 4.5, z_par
    assert np.max(np.abs(z_k)) < 4.5, z_k


@pytest.mark.parametrize("periodic", ["xyz", ""])
def test_c2_axon_pgse_subset(rms, orc, engine, W, periodic):
    c = W.C2
    print(f"case C2 axon PGSE 360 rows, 1e4 x 1e5, edges '{periodic or 'mirror'}'", flush=True)
    lab = W.c2_labels()
    nwalk = 10_000
    gpu = engine.run(W.c2_jobs(rms.Geometry, nwalk, periodic, lab), W.DT, c["TN"], c["timepoints"], precision="f32",
                     replicas=32)
    ref = orc.run(W.c2_jobs(orc.Geometry, nwalk, periodic, lab), W.DT, c["TN"], c["timepoints"], prec="f32",
                  replicas=32)
    assert_same(gpu, ref)
    assert gpu[0]["sig"].shape == (360,)
    assert np.all(gpu[0]["NPar_count"][:, 1] == nwalk)


def test_c2_validity_all_lobes(rms, orc, engine, W):
    """At TN = 1e5 (20 ms) the Delta = 20 and 40 ms rows lose lobe 2
    (main_PGSE.cu:341-345); at TN = 2.25e5 every row has both lobes."""
    TN, T = 225_000, 1000
    grad = W.pgse_table_c2()
    abe = orc.pgse_lobes(W.DT, TN, grad)
    assert np.all(abe[:, 2] < TN) and np.all(abe[:, 2] - abe[:, 1] >= abe[:, 0] - 1)
    abe_short = orc.pgse_lobes(W.DT, 100_000, grad)
    assert np.sum(abe_short[:, 2] - abe_short[:, 1] < abe_short[:, 0] - 1) == 240
    print("case C2 validity TN 2.25e5, 1024 walkers", flush=True)
    lab = W.c2_labels()
    gpu = engine.run(W.c2_jobs(rms.Geometry, 1024, "xyz", lab), W.DT, TN, T, precision="f32", replicas=32)
    ref = orc.run(W.c2_jobs(orc.Geometry, 1024, "xyz", lab), W.DT, TN, T, prec="f32", replicas=32)
    assert_same(gpu, ref)
    print("case C2 validity: periodic free-diffusion analogue, 65536 walkers", flush=True)
    free = np.ones((16, 16, 16), np.uint8)
    N = 65536
    g = rms.Geometry(free, W.RES, [0.0, W.C2["D0"]], T2=[math.inf, math.inf], periodic="xyz")
    r = engine.run([dict(geom=g, nwalk=N, seed=W.C2["seed"], grad=grad)], W.DT, TN, T, precision="f32",
                   replicas=32)[0]
    S = r["sig"] / N
    se = np.std(r["sig_rep"] / (N / 32), axis=0, ddof=1) / math.sqrt(32)
    want = gaussian_phase_signal(abe, grad, W.DT, TN, W.C2["D0"])
    z = (S - want) / (se + 1e-3 * want)
    assert np.max(np.abs(z)) < 4.5, (np.argmax(np.abs(z)), S[np.argmax(np.abs(z))], want[np.argmax(np.abs(z))])
    # the free walk: D = D0 on every axis at every record
    assert r["counters"][2] == 0


def test_c3_bundle_subset(rms, orc, engine, W):
    from rms_amd import autotune
    c = W.C3
    labs = W.c3_labels()
    which = list(range(0, 256, 16))
    print("case C3 bundle: 16 of 256 geometries at 1% walkers, autotuned", flush=True)
    gj = W.c3_jobs(rms.Geometry, c["nwalk"] // 100, labs, which)
    oj = W.c3_jobs(orc.Geometry, c["nwalk"] // 100, labs, which)
    best = autotune.tune(engine, gj, W.DT, timepoints=10, precision="f32", probe_steps=200, reps=1, use_cache=False)
    order = best["order"]
    print(f"autotuned {best['block_threads']} threads x {best['grid_blocks']} WGs, xcd {best['xcd_split']}, "
          f"pack {best['pack']}", flush=True)
    plan = engine.plan([gj[k] for k in order], W.DT, c["TN"], timepoints=c["timepoints"], precision="f32",
                       replicas=32, block_threads=best["block_threads"], grid_blocks=best["grid_blocks"],
                       xcd_split=best["xcd_split"])
    gpu_ordered = plan.execute()
    plan.close()
    gpu = [None] * len(gj)
    for pos, k in enumerate(order):
        gpu[k] = gpu_ordered[pos]
    ref = orc.run(oj, W.DT, c["TN"], c["timepoints"], prec="f32", replicas=32)
    assert_same(gpu, ref)
    assert sum(int(r["counters"][4]) for r in gpu) > 0  # exchange draws happened
    print("case C3 bundle: each job alone", flush=True)
    for k in range(len(gj)):
        alone = engine.run([gj[k]], W.DT, c["TN"], c["timepoints"], precision="f32", replicas=32)
        assert_same(alone, [gpu[k]])


def _adjoint_pair(rms, orc, engine, W, gj, oj, TN, T, nclass):
    from rms_amd import analysis as A
    plan = engine.plan(gj, W.DT, TN, timepoints=T, precision="f32", replicas=32)
    fwd = plan.execute()
    TD = A.diff_times(W.DT, TN, T)
    cots = [W.c4_cotangent(A, f["rec"], nclass, TD, int(j["bval"].size)) for f, j in zip(fwd, gj)]
    adj = plan.adjoint(cots, None, want_replicas=True, estimator="boundary")
    # the default estimator: central differences of D on walker subsets
    dflt = plan.adjoint(cots, None, want_replicas=True)
    plan.close()
    ref = orc.run(oj, W.DT, TN, T, prec="f32", replicas=32, cot_rec=cots)
    assert_same(fwd, ref)
    for a, b in zip(adj, ref):
        np.testing.assert_array_equal(a["grad"], b["grad"])
        np.testing.assert_array_equal(a["grad_rep"], b["grad_rep"])
    print("case central differences on walker subsets against the oracle", flush=True)
    g, gr = orc.central_differences(oj, W.DT, TN, T, cots, h=0.05, prec="f32", replicas=32, subsets=True)
    for k, a in enumerate(dflt):
        np.testing.assert_array_equal(a["grad"][1:nclass], g[k, 1:nclass])
        np.testing.assert_array_equal(a["grad_rep"][:, 1:nclass], gr[k, :, 1:nclass])
        np.testing.assert_array_equal(a["grad"][4:9], adj[k]["grad"][4:9])
    return fwd, adj


def test_c4_gradient_subset(rms, orc, engine, W):
    c = W.C4
    print("case C4 ECS axon kappa 0.01, 1e4 x 5e4, forward + adjoint", flush=True)
    lab = W.c4_labels()
    fwd, adj = _adjoint_pair(rms, orc, engine, W, W.c4_jobs(rms.Geometry, 10_000, lab),
                             W.c4_jobs(orc.Geometry, 10_000, lab), c["TN"], c["timepoints"], 3)
    g = adj[0]["grad"]
    assert np.all(np.isfinite(g[4:7])) and np.isfinite(g[8])
    assert fwd[0]["counters"][4] > 0


def test_c5_sweep_subset(rms, orc, engine, W):
    c = W.C5
    print("case C5 sweep: 8 of 4096 geometries, 640 walkers each, forward + adjoint", flush=True)
    gj = W.c5_jobs(rms.Geometry, range(8), 640)
    oj = W.c5_jobs(orc.Geometry, range(8), 640)
    assert all(j["geom"].labels.dtype == np.uint16 and j["geom"].labels.shape == (70, 70, 1000) for j in gj)
    _adjoint_pair(rms, orc, engine, W, gj, oj, c["TN"], c["timepoints"], 3)
