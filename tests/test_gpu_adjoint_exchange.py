"""The D derivative with its boundary terms on the GPU (the jump of the class
change at exchange faces with the score of the draws, the origin mirror at
lattice edges, the face order at voxel edges): bit-exact against the
oracle's adjoint on the same seeds, and the five C4 derivatives against
central differences of the forward pass with common random numbers."""
import numpy as np
import pytest

from helpers import fib_dirs, random_blobs

pytestmark = pytest.mark.gpu


def _pair(rms, orc, lab, D, T2, kappa, periodic=""):
    mk = lambda M: M.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic)  # noqa: E731
    return mk(rms), mk(orc)


@pytest.mark.parametrize("prec,nclass,periodic", [("f32", 3, ""), ("f64", 3, ""), ("f32", 3, "xyz"),
                                                  ("f32", 4, ""), ("f64", 4, "xz"), ("f32", 2, ""),
                                                  ("f64", 2, "y")])
def test_adjoint_with_exchange_bitexact(rms, orc, engine, prec, nclass, periodic):
    print(f"case adjoint jump {prec} nclass={nclass} periodic={periodic}", flush=True)
    lab = random_blobs((11, 10, 12), nclass, seed=30 + nclass)
    D = [0.0, 2.0, 0.7, 1.3][:nclass]
    T2 = [np.inf, 80.0, 30.0, 50.0][:nclass]
    g, og = _pair(rms, orc, lab, D, T2, 0.3, periodic)
    bval, bvec = np.array([0.5, 2.0, 3.0]), fib_dirs(3)
    job = dict(nwalk=1100, seed=61, bval=bval, bvec=bvec)
    dt, TN, T = 2e-4, 600, 12
    nval = rms.nval(nclass, 3)
    cot = [np.random.default_rng(5).standard_normal((T, nval))]
    plan = engine.plan([dict(geom=g, **job)], dt, TN, T, precision=prec, replicas=8)
    try:
        fwd = plan.execute()
        adj = plan.adjoint(cot, None, want_replicas=True)
    finally:
        plan.close()
    ref = orc.run([dict(geom=og, **job)], dt, TN, T, prec=prec, replicas=8, cot_rec=cot)
    np.testing.assert_array_equal(fwd[0]["rec"], ref[0]["rec"])
    np.testing.assert_array_equal(adj[0]["grad"], ref[0]["grad"])
    np.testing.assert_array_equal(adj[0]["grad_rep"], ref[0]["grad_rep"])
    gD = adj[0]["grad"][1:nclass]
    assert np.all(np.isfinite(gD)) and np.all(gD != 0.0)
    assert (fwd[0]["counters"][4] > 0) == (nclass > 2)  # exchange draws happened


def test_c4_gradient_against_central_differences(rms, engine):
    """C4 (the axon with extra-axonal space, kappa = 0.01, 5e4 steps, 20
    records, 256 narrow-pulse b-vectors), objective sum_t [D_t(z) + K_t(z) +
    sum_j S_j(t)] linearised at the base point (fixed cotangent).  Each of the
    five derivatives (D_intra, D_extra, T2_intra, T2_extra, kappa) against
    (F(theta + h) - F(theta - h)) / 2h with the same seed, within 1 % or 3
    combined standard errors (32 batch-means replicas on both sides)."""
    from rms_amd import analysis as A
    W = rms.workloads
    c = W.C4
    N, TN, T, R = 1_000_000, c["TN"], c["timepoints"], 32
    lab = W.c4_labels()
    base = dict(D=(0.0, 2.0, 1.5), T2=(np.inf, 80.0, 60.0), kappa=c["kappa"])
    print("case C4 gradient vs central differences: base forward + adjoint", flush=True)
    jobs = W.c4_jobs(rms.Geometry, N, lab, **base)
    plan = engine.plan(jobs, W.DT, TN, timepoints=T, precision="f32", replicas=R)
    try:
        fwd = plan.execute()[0]
        TD = A.diff_times(W.DT, TN, T)
        cot = W.c4_cotangent(A, fwd["rec"], 3, TD, int(jobs[0]["bval"].size))
        adj = plan.adjoint([cot], None, want_replicas=True)[0]
    finally:
        plan.close()

    def F_rep(**phys):
        p = dict(base, **phys)
        r = engine.run(W.c4_jobs(rms.Geometry, N, lab, **p), W.DT, TN, T, precision="f32", replicas=R)[0]
        return (cot[None] * r["rec_rep"]).reshape(R, -1).sum(1)

    checks = []
    for name, idx, h in (("D_intra", 1, 0.05 * 2.0), ("D_extra", 2, 0.05 * 1.5), ("T2_intra", 5, 1e-3 * 80.0),
                         ("T2_extra", 6, 1e-3 * 60.0), ("kappa", 8, 0.2 * c["kappa"])):
        print(f"case C4 central difference {name}", flush=True)
        if name.startswith("D_"):
            lo, hi = list(base["D"]), list(base["D"])
            lo[idx] -= h
            hi[idx] += h
            fd = (F_rep(D=tuple(hi)) - F_rep(D=tuple(lo))) / (2 * h)
        elif name.startswith("T2_"):
            k = idx - 4
            lo, hi = list(base["T2"]), list(base["T2"])
            lo[k] -= h
            hi[k] += h
            fd = (F_rep(T2=tuple(hi)) - F_rep(T2=tuple(lo))) / (2 * h)
        else:
            fd = (F_rep(kappa=c["kappa"] + h) - F_rep(kappa=c["kappa"] - h)) / (2 * h)
        g_adj, se_adj = adj["grad"][idx], np.sqrt(R) * np.std(adj["grad_rep"][:, idx], ddof=1)
        g_fd, se_fd = fd.sum(), np.sqrt(R) * np.std(fd, ddof=1)
        ok = abs(g_adj - g_fd) <= max(0.01 * abs(g_fd), 3 * np.hypot(se_adj, se_fd))
        print(f"{name}: adjoint {g_adj:.6g} +- {se_adj:.3g}, central difference {g_fd:.6g} +- {se_fd:.3g}", flush=True)
        checks.append((name, ok, g_adj, se_adj, g_fd, se_fd))
    assert all(c[1] for c in checks), checks
