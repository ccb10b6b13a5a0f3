"""The D derivative with its boundary terms on the GPU (the jump of the class
change at exchange faces with the score of the draws, the origin mirror at
lattice edges, the face order at voxel edges): bit-exact against the
oracle's adjoint on the same seeds, and the five C4 derivatives against
central differences of the forward pass with common random numbers."""
import numpy as np
import pytest

from helpers import fib_dirs, pgse_table, planes, random_blobs, slabs_z

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
        adj = plan.adjoint(cot, None, want_replicas=True, estimator="boundary")
    finally:
        plan.close()
    ref = orc.run([dict(geom=og, **job)], dt, TN, T, prec=prec, replicas=8, cot_rec=cot)
    np.testing.assert_array_equal(fwd[0]["rec"], ref[0]["rec"])
    np.testing.assert_array_equal(adj[0]["grad"], ref[0]["grad"])
    np.testing.assert_array_equal(adj[0]["grad_rep"], ref[0]["grad_rep"])
    gD = adj[0]["grad"][1:nclass]
    assert np.all(np.isfinite(gD)) and np.all(gD != 0.0)
    assert (fwd[0]["counters"][4] > 0) == (nclass > 2)  # exchange draws happened


def _contract(cot_rec, rec, cot_sig, sig):
    """sum(cot_rec * rec) + sum(cot_sig * sig), summed in row-major order."""
    f = 0.0
    for c, r in zip(np.ravel(cot_rec).tolist(), np.ravel(rec).tolist()):
        f = f + c * r
    if cot_sig is not None:
        for c, r in zip(np.ravel(cot_sig).tolist(), np.ravel(sig).tolist()):
            f = f + c * r
    return f


@pytest.mark.parametrize("prec,tables,periodic", [("f32", "narrow", ""), ("f64", "pgse", "xz"),
                                                  ("f32", "both", "xyz")])
def test_auto_adjoint_central_differences_match_oracle(rms, orc, engine, prec, tables, periodic):
    """The default D estimator on a geometry whose walk jumps in D (exchange
    between two non-null classes, voxel edges), its central differences
    walking every walker: T2 and kappa entries bit-identical to the oracle's
    adjoint; each D entry, total and per replica, equal to the central
    difference of the oracle's forward runs at D_c (1 +- h) with the same
    seeds, contracted in the same order."""
    print(f"case auto adjoint {prec} {tables} periodic={periodic}", flush=True)
    lab = random_blobs((10, 11, 9), 3, seed=41)
    D, T2, kappa, h = [0.0, 2.0, 0.8], [np.inf, 70.0, 40.0], 0.2, 0.03
    job = dict(nwalk=900, seed=5)
    if tables in ("narrow", "both"):
        job.update(bval=np.array([0.7, 2.5]), bvec=fib_dirs(2))
    if tables in ("pgse", "both"):
        job.update(grad=pgse_table(fib_dirs(2), [1.0], [0.03], 0.01))
    dt, TN, T, R = 2e-4, 400, 20, 4
    nb = len(job["bval"]) if "bval" in job else 0
    ng = job["grad"].shape[0] if "grad" in job else 0
    rng = np.random.default_rng(9)
    cot_rec = [rng.standard_normal((T, rms.nval(3, nb)))]
    cot_sig = [rng.standard_normal(ng)] if ng else None
    plan = engine.plan([dict(geom=rms.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic), **job)],
                       dt, TN, T, precision=prec, replicas=R)
    try:
        adj = plan.adjoint(cot_rec, cot_sig, want_replicas=True, fd_step=h, fd_all_walkers=True)[0]
    finally:
        plan.close()
    og = orc.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic)
    ref = orc.run([dict(geom=og, **job)], dt, TN, T, prec=prec, replicas=R, cot_rec=cot_rec, cot_sig=cot_sig)[0]
    np.testing.assert_array_equal(adj["grad"][4:9], ref["grad"][4:9])
    np.testing.assert_array_equal(adj["grad_rep"][:, 4:9], ref["grad_rep"][:, 4:9])
    cs = None if cot_sig is None else cot_sig[0]
    for c in (1, 2):
        runs = []
        for f in (1.0 + h, 1.0 - h):
            Dc = list(D)
            Dc[c] = D[c] * f
            gc = orc.Geometry(lab, 0.1, Dc, T2=T2, kappa=kappa, periodic=periodic)
            runs.append((Dc[c], orc.run([dict(geom=gc, **job)], dt, TN, T, prec=prec, replicas=R)[0]))
        (dp, rp), (dm, rm) = runs
        tot = (_contract(cot_rec[0], rp["rec"], cs, rp["sig"]) - _contract(cot_rec[0], rm["rec"], cs, rm["sig"]))
        assert adj["grad"][c] == tot / (dp - dm), (c, adj["grad"][c], tot / (dp - dm))
        for r in range(R):
            sp = None if cs is None else rp["sig_rep"][r]
            sm = None if cs is None else rm["sig_rep"][r]
            fr = _contract(cot_rec[0], rp["rec_rep"][r], cs, sp) - _contract(cot_rec[0], rm["rec_rep"][r], cs, sm)
            assert adj["grad_rep"][r, c] == fr / (dp - dm), (c, r)
    assert np.all(np.isfinite(adj["grad"][1:3])) and np.all(adj["grad"][1:3] != 0.0)


def _fd_scale(cot_rec, rec, cot_sig, sig, h, Dc):
    """Size of the rounding in a central difference of D_c: the objective's
    absolute terms over the difference of the two D values, times 1e-12."""
    f = np.sum(np.abs(cot_rec) * np.abs(rec))
    if cot_sig is not None:
        f += np.sum(np.abs(cot_sig) * np.abs(sig))
    return 1e-12 * f / (2 * h * Dc)


SUBSET_CASES = {
    # two non-null classes with a permeability: the draws at faces into the
    # class whose D moves mark their walkers
    "two": dict(labels=3, D=[0.0, 2.0, 0.8], kappa=0.2, prec="f32", tables="narrow"),
    # three non-null classes, FP64, PGSE signal in the objective
    "three": dict(labels=4, D=[0.0, 2.0, 0.7, 1.3], kappa=0.3, prec="f64", tables="pgse"),
    # the reference rule P = min(1, sqrt(Dj/Di)) without a permeability:
    # walkers cross into the faster class without a draw (their class change
    # marks them)
    "nokappa": dict(labels=3, D=[0.0, 2.0, 0.8], kappa=None, prec="f32", tables="narrow"),
    # D_1 and D_2 so close that D_c (1 +- h) moves P_12 off 1: a face that
    # crosses without a draw at the base point draws at another, so every
    # walker is walked and the entries equal those of the full walks exactly
    "kink": dict(labels=3, D=[0.0, 1.0, 1.02], kappa=None, prec="f32", tables="narrow"),
    # class 3 has no voxel: no walker can walk differently with D_3
    "absent": dict(labels=3, nclass=4, D=[0.0, 2.0, 0.8, 1.1], kappa=0.2, prec="f32", tables="narrow"),
}


@pytest.mark.parametrize("case", sorted(SUBSET_CASES))
def test_central_differences_on_walker_subsets(rms, engine, case):
    """The central differences of D walk, for class c, only the walkers that
    the tangent pass saw in class c or drawing at an exchange face whose P
    depends on D_c: each D entry equals that of the walks of every walker to
    rounding (the other walkers' walks are the same at both points), T2 and
    kappa entries are untouched, and the replicas sum to the total.  A plan
    with a second job whose D entries are pathwise (periodic planes) leaves
    that job's entries bit-identical."""
    cfg = SUBSET_CASES[case]
    print(f"case central differences on walker subsets: {case}", flush=True)
    lab = random_blobs((10, 11, 9), cfg["labels"], seed=52)
    nclass = cfg.get("nclass", cfg["labels"])
    D = cfg["D"]
    T2 = [np.inf, 70.0, 40.0, 55.0][:nclass]
    kw = {} if cfg["kappa"] is None else dict(kappa=cfg["kappa"])
    lc = np.arange(cfg["labels"], dtype=np.int32) if nclass != cfg["labels"] else None
    g = rms.Geometry(lab, 0.1, D, T2=T2, label_class=lc, **kw)
    job = dict(geom=g, nwalk=1500, seed=77)
    if cfg["tables"] == "narrow":
        job.update(bval=np.array([0.7, 2.5]), bvec=fib_dirs(2))
    else:
        job.update(grad=pgse_table(fib_dirs(2), [1.0], [0.03], 0.01))
    gp = rms.Geometry(planes(7, 6, 12), 0.1, [0.0, 1.8], T2=[50.0, 60.0], periodic="xyz")
    jobs = [job, dict(job, geom=gp, seed=78, nwalk=700)]
    dt, TN, T, R, h = 2e-4, 400, 16, 4, 0.05
    nb = len(job["bval"]) if "bval" in job else 0
    ng = job["grad"].shape[0] if "grad" in job else 0
    rng = np.random.default_rng(13)
    cot_rec = [rng.standard_normal((T, rms.nval(nclass, nb))), rng.standard_normal((T, rms.nval(2, nb)))]
    cot_sig = [rng.standard_normal(ng), rng.standard_normal(ng)] if ng else None
    plan = engine.plan(jobs, dt, TN, T, precision=cfg["prec"], replicas=R)
    try:
        fwd = plan.execute()
        assert plan.timing["walker_steps"] == sum(j["nwalk"] for j in jobs) * TN
        sub = plan.adjoint(cot_rec, cot_sig, want_replicas=True)
        t_sub = plan.adj_timing["walk_ms"]
        full = plan.adjoint(cot_rec, cot_sig, want_replicas=True, fd_all_walkers=True)
        t_full = plan.adj_timing["walk_ms"]
    finally:
        plan.close()
    print(f"adjoint walk ms: subsets {t_sub:.2f}, every walker {t_full:.2f}", flush=True)
    a, b = sub[0], full[0]
    np.testing.assert_array_equal(a["grad"][4:9], b["grad"][4:9])
    np.testing.assert_array_equal(a["grad_rep"][:, 4:9], b["grad_rep"][:, 4:9])
    cs = None if cot_sig is None else cot_sig[0]
    sig = fwd[0]["sig"]
    for c in range(1, nclass):
        tol = _fd_scale(cot_rec[0], fwd[0]["rec"], cs, sig, h, D[c])
        assert abs(a["grad"][c] - b["grad"][c]) <= tol, (c, a["grad"][c], b["grad"][c], tol)
        assert abs(a["grad_rep"][:, c].sum() - a["grad"][c]) <= tol + 1e-12 * np.abs(a["grad_rep"][:, c]).sum()
    if case == "kink":
        np.testing.assert_array_equal(a["grad"], b["grad"])
        np.testing.assert_array_equal(a["grad_rep"], b["grad_rep"])
    if case == "absent":
        assert a["grad"][3] == 0.0 and b["grad"][3] == 0.0
        assert np.all(a["grad_rep"][:, 3] == 0.0)
    assert np.all(np.isfinite(a["grad"][1:nclass])) and np.all(a["grad"][1:3] != 0.0)
    # the pathwise job: no central differences at all
    np.testing.assert_array_equal(sub[1]["grad"], full[1]["grad"])
    np.testing.assert_array_equal(sub[1]["grad_rep"], full[1]["grad_rep"])


@pytest.mark.parametrize("prec,case", [("f32", "strong"), ("f64", "three"), ("f32", "nokappa"), ("f32", "mixed"),
                                       ("f64", "pgse")])
def test_walker_subset_differences_bitexact_against_oracle(rms, orc, engine, prec, case):
    """The default central differences (walker subsets, h = 0.05) on slab
    geometries where most walkers never meet another class: every D entry,
    total and per replica, bit-identical to the oracle's central differences
    on the walkers its own tangent pass marks (same packing into batches:
    the marked walkers in index order), for two jobs in one plan, the second
    a shard (walker_base: Philox ids from 64 R 3).  "mixed": the first job's
    lattice has no voxel of class 1, so its walks for D_1 are left out while
    the second job's are not, and both walk for D_2.  "pgse": a PGSE table,
    its signal in the objective."""
    cfg = {"strong": dict(labels=3, D=[0.0, 2.0, 0.8], kappa=5.0),
           "three": dict(labels=4, D=[0.0, 2.0, 0.7, 1.3], kappa=0.3),
           "nokappa": dict(labels=3, D=[0.0, 2.0, 0.8], kappa=None),
           "mixed": dict(labels=3, D=[0.0, 2.0, 0.8], kappa=5.0),
           "pgse": dict(labels=3, D=[0.0, 2.0, 0.8], kappa=5.0)}[case]
    print(f"case walker-subset central differences vs oracle: {prec} {case}", flush=True)
    lab = slabs_z(cfg["labels"])
    labs = [lab, lab]
    if case == "mixed":
        labs[0] = np.where(lab == 1, 2, lab).astype(np.uint8)
    nc = len(cfg["D"])
    kw = {} if cfg["kappa"] is None else dict(kappa=cfg["kappa"])
    T2 = [np.inf, 70.0, 40.0, 55.0][:nc]
    mk = lambda M, L: M.Geometry(L, 0.1, cfg["D"], T2=T2, **kw)  # noqa: E731
    tab = dict(bval=np.array([0.8, 2.0]), bvec=fib_dirs(2))
    if case == "pgse":
        tab = dict(grad=pgse_table(fib_dirs(3), [1.0, 2.0], [0.03], 0.01))
    dt, TN, T, R = 2e-4, 500, 10, 4
    spec = [dict(nwalk=1500, seed=12, **tab), dict(nwalk=900, seed=13, walker_base=64 * R * 3, **tab)]
    rng = np.random.default_rng(4)
    nb = 0 if case == "pgse" else 2
    cot = [rng.standard_normal((T, rms.nval(nc, nb))) for _ in spec]
    cots = [rng.standard_normal(tab["grad"].shape[0]) for _ in spec] if case == "pgse" else None
    plan = engine.plan([dict(geom=mk(rms, L), **j) for L, j in zip(labs, spec)], dt, TN, T, precision=prec,
                       replicas=R)
    try:
        adj = plan.adjoint(cot, cots, want_replicas=True)
        walked = plan.adj_timing["walker_steps"]
    finally:
        plan.close()
    g, gr = orc.central_differences([dict(geom=mk(orc, L), **j) for L, j in zip(labs, spec)], dt, TN, T, cot,
                                    cots, h=0.05, prec=prec, replicas=R, subsets=True)
    if case == "mixed":
        assert adj[0]["grad"][1] == 0.0 and adj[0]["grad"][2] != 0.0 and adj[1]["grad"][1] != 0.0
    for k in range(len(spec)):
        for c in range(1, nc):
            assert adj[k]["grad"][c] == g[k, c], (k, c, adj[k]["grad"][c], g[k, c])
            np.testing.assert_array_equal(adj[k]["grad_rep"][:, c], gr[k, :, c])
    # the tangent pass walks every walker, the central differences fewer
    full = sum(j["nwalk"] for j in spec) * TN
    assert full < walked < full * (1 + 2 * (nc - 1)), (walked, full)


def test_central_differences_nan_where_a_perturbation_needs_another_variant(rms, engine):
    """Three non-null classes with equal D under the reference rule
    P = min(1, sqrt(Dj/Di)): every pair crosses (P = 1) and the plan runs the
    variant without exchange draws.  Moving D_c by +-h makes some pairs draw,
    which that variant cannot walk, so the central differences of D are NaN
    while T2 stays finite; with two non-null classes the plan draws anyway
    and the differences are finite."""
    print("case central differences outside the plan's variant", flush=True)
    lab = random_blobs((8, 9, 10), 4, seed=12)
    g = rms.Geometry(lab, 0.1, [0.0, 1.5, 1.5, 1.5], T2=[np.inf, 60.0, 50.0, 40.0])
    job = dict(geom=g, nwalk=500, seed=4, bval=np.array([1.0]), bvec=fib_dirs(1))
    plan = engine.plan([job], 2e-4, 200, 10, precision="f32", replicas=4)
    try:
        a = plan.adjoint([np.ones((10, rms.nval(4, 1)))], None, want_replicas=True)[0]
    finally:
        plan.close()
    assert np.all(np.isnan(a["grad"][1:4])) and np.all(np.isnan(a["grad_rep"][:, 1:4]))
    assert np.all(np.isfinite(a["grad"][5:8]))
    lab3 = random_blobs((8, 9, 10), 3, seed=12)
    g3 = rms.Geometry(lab3, 0.1, [0.0, 1.5, 1.5], T2=[np.inf, 60.0, 50.0])
    plan = engine.plan([dict(job, geom=g3)], 2e-4, 200, 10, precision="f32", replicas=4)
    try:
        a3 = plan.adjoint([np.ones((10, rms.nval(3, 1)))], None)[0]
    finally:
        plan.close()
    assert np.all(np.isfinite(a3["grad"][1:3]))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_auto_adjoint_pathwise_where_continuous(rms, orc, engine, prec):
    """Periodic impermeable planes (no exchange, no lattice-edge mirror, no
    voxel edge whose face order matters): the default estimator is the
    pathwise derivative, every entry bit-identical to the oracle's adjoint
    and D finite, and asking for central differences gives another, nearby
    value."""
    print(f"case auto adjoint pathwise {prec}", flush=True)
    lab = planes(7, 6, 12)
    geom = dict(labels=lab, res=0.1, D=[0.0, 1.8], T2=[50.0, 60.0], periodic="xyz")
    job = dict(nwalk=2000, seed=23, bval=np.array([1.5, 3.0]), bvec=fib_dirs(2))
    dt, TN, T, R = 2e-4, 500, 25, 8
    rng = np.random.default_rng(3)
    cot = [rng.standard_normal((T, rms.nval(2, 2)))]
    g = rms.Geometry(lab, 0.1, geom["D"], T2=geom["T2"], periodic="xyz")
    plan = engine.plan([dict(geom=g, **job)], dt, TN, T, precision=prec, replicas=R)
    try:
        a = plan.adjoint(cot, None, want_replicas=True)[0]
        f = plan.adjoint(cot, None, want_replicas=True, estimator="fd", fd_step=0.02)[0]
    finally:
        plan.close()
    og = orc.Geometry(lab, 0.1, geom["D"], T2=geom["T2"], periodic="xyz")
    ref = orc.run([dict(geom=og, **job)], dt, TN, T, prec=prec, replicas=R, cot_rec=cot)[0]
    np.testing.assert_array_equal(a["grad"], ref["grad"])
    np.testing.assert_array_equal(a["grad_rep"], ref["grad_rep"])
    assert np.isfinite(a["grad"][1]) and a["grad"][1] != 0.0
    np.testing.assert_array_equal(f["grad"][4:9], a["grad"][4:9])
    se = np.sqrt(R) * np.std(f["grad_rep"][:, 1] - a["grad_rep"][:, 1], ddof=1)
    assert f["grad"][1] != a["grad"][1] and abs(f["grad"][1] - a["grad"][1]) <= 4 * se + 1e-3 * abs(a["grad"][1])


def test_c4_gradient_against_central_differences(rms, engine):
    """C4 (the axon with extra-axonal space, kappa = 0.01, 5e4 steps, 20
    records, 256 narrow-pulse b-vectors), objective sum_t [D_t(z) + K_t(z) +
    sum_j S_j(t)] linearised at the base point (fixed cotangent).  The five
    derivatives of the default adjoint (D_intra and D_extra by the library's
    central differences at h = 5 %, T2 analytic, kappa by its score) against
    this test's central differences with the same seeds (D at h = 2.5 %, T2
    at 0.1 %, kappa at 20 %), within 1 % or 3 combined standard errors (32
    batch-means replicas on both sides); the boundary-term estimator's D
    entries within 3 of its (large) standard errors of the same."""
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
        print("case C4 boundary-term estimator", flush=True)
        bnd = plan.adjoint([cot], None, want_replicas=True, estimator="boundary")[0]
    finally:
        plan.close()

    def F_rep(**phys):
        p = dict(base, **phys)
        r = engine.run(W.c4_jobs(rms.Geometry, N, lab, **p), W.DT, TN, T, precision="f32", replicas=R)[0]
        return (cot[None] * r["rec_rep"]).reshape(R, -1).sum(1)

    def se(x):
        return np.sqrt(R) * np.std(x, ddof=1)

    checks = []
    for name, idx, h in (("D_intra", 1, 0.025 * 2.0), ("D_extra", 2, 0.025 * 1.5), ("T2_intra", 5, 1e-3 * 80.0),
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
        g_adj, se_adj = adj["grad"][idx], se(adj["grad_rep"][:, idx])
        g_fd, se_fd = fd.sum(), se(fd)
        ok = abs(g_adj - g_fd) <= max(0.01 * abs(g_fd), 3 * np.hypot(se_adj, se_fd))
        print(f"{name}: adjoint {g_adj:.6g} +- {se_adj:.3g}, central difference {g_fd:.6g} +- {se_fd:.3g}", flush=True)
        checks.append((name, ok, g_adj, se_adj, g_fd, se_fd))
        if name.startswith("D_"):
            g_b, se_b = bnd["grad"][idx], se(bnd["grad_rep"][:, idx])
            ok_b = np.isfinite(g_b) and abs(g_b - g_fd) <= 3 * np.hypot(se_b, se_fd)
            print(f"{name}: boundary terms {g_b:.6g} +- {se_b:.3g}", flush=True)
            checks.append((name + " boundary", ok_b, g_b, se_b, g_fd, se_fd))
    np.testing.assert_array_equal(bnd["grad"][4:9], adj["grad"][4:9])
    assert all(c[1] for c in checks), checks


def test_central_differences_in_chunked_graph_plans(rms, engine):
    """The central differences walk the plan's own launch sequence (chunks
    here) at D (1 +- h) and restore its descriptors: the D entries equal
    those of a single-launch plan, and execute() afterwards (a captured
    graph) still returns the unperturbed records."""
    print("case central differences in a chunked graph plan", flush=True)
    lab = random_blobs((10, 9, 11), 3, seed=19)
    g = rms.Geometry(lab, 0.1, [0.0, 1.8, 0.9], T2=[np.inf, 60.0, 40.0], kappa=0.15)
    job = dict(geom=g, nwalk=900, seed=31, bval=np.array([0.5, 2.0]), bvec=fib_dirs(2))
    dt, TN, T = 2e-4, 480, 12
    cot = [np.random.default_rng(2).standard_normal((T, rms.nval(3, 2)))]
    one = engine.plan([job], dt, TN, T, precision="f32", replicas=4)
    ch = engine.plan([job], dt, TN, T, precision="f32", replicas=4, chunk_steps=100, use_graph=True)
    try:
        r0 = ch.execute()
        a1 = one.adjoint(cot, None, want_replicas=True)[0]
        a2 = ch.adjoint(cot, None, want_replicas=True)[0]
        r1 = ch.execute()
    finally:
        one.close()
        ch.close()
    np.testing.assert_array_equal(a1["grad"], a2["grad"])
    np.testing.assert_array_equal(a1["grad_rep"], a2["grad_rep"])
    np.testing.assert_array_equal(r0[0]["rec"], r1[0]["rec"])
    assert np.all(np.isfinite(a2["grad"][1:3]))
