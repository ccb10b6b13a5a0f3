"""Brick storage of the cell words (rms_geometry.storage): the same walk as
dense storage, bit for bit, on lattices whose sides are not multiples of the
2x2x4 brick, with periodic axes, 2-4 classes, both precisions, the adjoint's
estimators and plans that mix both storages; and the footprint it saves."""
import numpy as np
import pytest

from helpers import fib_dirs, pgse_table, random_blobs

pytestmark = pytest.mark.gpu


def _same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x["rec"], y["rec"])
        np.testing.assert_array_equal(x["counters"], y["counters"])
        if x.get("sig") is not None:
            np.testing.assert_array_equal(x["sig"], y["sig"])
        for k in ("fin_g", "fin_o", "fin_gi", "fin_oi", "fin_c"):
            if k in x:
                np.testing.assert_array_equal(x[k], y[k])


CASES = [((9, 7, 13), 2, "", "f32", "narrow"), ((5, 11, 6), 3, "xyz", "f32", "pgse"),
         ((10, 10, 10), 3, "", "f64", "both"), ((7, 6, 17), 4, "xz", "f32", "narrow"),
         ((3, 3, 3), 3, "y", "f64", "pgse")]


@pytest.mark.parametrize("shape,nclass,periodic,prec,tables", CASES)
def test_bricks_equal_dense_and_oracle(rms, orc, engine, shape, nclass, periodic, prec, tables):
    print(f"case bricks {shape} nclass={nclass} periodic={periodic} {prec} {tables}", flush=True)
    lab = random_blobs(shape, nclass, seed=sum(shape) + nclass)
    D = [0.0, 2.0, 0.7, 1.3][:nclass]
    T2 = [np.inf, 80.0, 30.0, 50.0][:nclass]
    job = dict(nwalk=1500, seed=int(np.prod(shape)))
    if tables in ("narrow", "both"):
        job.update(bval=np.array([0.5, 2.0]), bvec=fib_dirs(2))
    if tables in ("pgse", "both"):
        job.update(grad=pgse_table(fib_dirs(2), [1.0], [0.03], 0.01))
    kw = dict(dt=2e-4, TN=600, timepoints=20, precision=prec, replicas=4)
    out = {}
    for st in ("dense", "bricks"):
        g = rms.Geometry(lab, 0.1, D, T2=T2, kappa=0.2, periodic=periodic, storage=st)
        out[st] = engine.run([dict(geom=g, **job)], final_state=True, **kw)
        assert engine.geometry_info(g)[1] == st
    _same(out["bricks"], out["dense"])
    og = orc.Geometry(lab, 0.1, D, T2=T2, kappa=0.2, periodic=periodic)
    ref = orc.run([dict(geom=og, **job)], kw["dt"], kw["TN"], kw["timepoints"], prec=prec, replicas=4)
    np.testing.assert_array_equal(out["bricks"][0]["rec"], ref[0]["rec"])
    np.testing.assert_array_equal(out["bricks"][0]["counters"], ref[0]["counters"])


@pytest.mark.parametrize("estimator", ["auto", "boundary"])
def test_bricks_adjoint_equal_dense(rms, engine, estimator):
    """The tangent pass (pathwise D, T2, kappa), the central differences and
    the boundary-term estimator (whose continuations read cell words too)."""
    print(f"case bricks adjoint {estimator}", flush=True)
    lab = random_blobs((11, 9, 10), 3, seed=77)
    job = dict(nwalk=1300, seed=3, bval=np.array([0.8, 2.5]), bvec=fib_dirs(2))
    dt, TN, T = 2e-4, 500, 10
    cot = [np.random.default_rng(1).standard_normal((T, rms.nval(3, 2)))]
    res = {}
    for st in ("dense", "bricks"):
        g = rms.Geometry(lab, 0.1, [0.0, 2.0, 0.9], T2=[np.inf, 70.0, 40.0], kappa=0.3, storage=st)
        plan = engine.plan([dict(geom=g, **job)], dt, TN, T, precision="f32", replicas=4)
        try:
            fwd = plan.execute()
            adj = plan.adjoint(cot, None, want_replicas=True, estimator=estimator)
        finally:
            plan.close()
        res[st] = (fwd, adj)
    _same(res["bricks"][0], res["dense"][0])
    np.testing.assert_array_equal(res["bricks"][1][0]["grad"], res["dense"][1][0]["grad"])
    np.testing.assert_array_equal(res["bricks"][1][0]["grad_rep"], res["dense"][1][0]["grad_rep"])


def test_bricks_and_dense_jobs_in_one_chunked_plan(rms, engine):
    """A plan whose jobs use either storage, in chunks: each job equal to its
    dense twin run alone."""
    print("case bricks mixed plan", flush=True)
    rng = np.random.default_rng(5)
    jobs_b, jobs_d = [], []
    for k, (shape, nc) in enumerate((((8, 9, 10), 3), ((6, 5, 21), 2), ((12, 7, 9), 3), ((5, 5, 5), 4))):
        lab = random_blobs(shape, nc, seed=100 + k)
        D = [0.0, 2.0, 0.6, 1.1][:nc]
        st = "bricks" if k % 2 == 0 else "dense"
        per = "xyz" if k == 2 else ""
        job = dict(nwalk=int(rng.integers(200, 900)), seed=k + 1, bval=np.array([1.0]), bvec=fib_dirs(1))
        jobs_b.append(dict(geom=rms.Geometry(lab, 0.1, D, kappa=0.1, periodic=per, storage=st), **job))
        jobs_d.append(dict(geom=rms.Geometry(lab, 0.1, D, kappa=0.1, periodic=per, storage="dense"), **job))
    kw = dict(dt=2e-4, TN=400, timepoints=10, precision="f32", replicas=4)
    mixed = engine.run(jobs_b, chunk_steps=96, **kw)
    for k in range(len(jobs_d)):
        _same([mixed[k]], engine.run([jobs_d[k]], **kw))


def test_bricks_footprint_and_auto_choice(rms, engine):
    """A sweep geometry (70 x 70 x 1000 uint16 axon with extra-axonal
    space): bricks hold under 2x the label bytes, dense storage 9 bytes per
    voxel; with the device's memory nearly free, auto keeps dense words."""
    W = rms.workloads
    print("case bricks footprint (sweep geometry 0)", flush=True)
    gb = W.c5_jobs(lambda *a, **k: rms.Geometry(*a, storage="bricks", **k), [0], 64)[0]["geom"]
    gd = W.c5_jobs(lambda *a, **k: rms.Geometry(*a, storage="dense", **k), [0], 64)[0]["geom"]
    ga = W.c5_jobs(rms.Geometry, [0], 64)[0]["geom"]
    nb, _ = engine.geometry_info(gb)
    ndn, _ = engine.geometry_info(gd)
    nvox = gb.labels.size
    print(f"bricks {nb / nvox:.3f} B/voxel, dense {ndn / nvox:.3f} B/voxel", flush=True)
    assert ndn == 9 * nvox and nb <= 2 * gb.labels.nbytes
    assert engine.geometry_info(ga)[1] == "dense"
    kw = dict(dt=2e-4, TN=300, timepoints=10, precision="f32", replicas=4)
    jb = W.c5_jobs(lambda *a, **k: rms.Geometry(*a, storage="bricks", **k), [0], 4096)
    jd = W.c5_jobs(lambda *a, **k: rms.Geometry(*a, storage="dense", **k), [0], 4096)
    _same(engine.run(jb, **kw), engine.run(jd, **kw))


def test_bricks_allocation_free_run_equals_run(rms, engine):
    """rms_run_noalloc on a brick-stored geometry (the brick kernel variant
    through the static scratch array) equals rms_run on it and on its dense
    twin."""
    print("case bricks allocation-free run", flush=True)
    lab = random_blobs((9, 10, 11), 3, seed=90)
    job = dict(nwalk=700, seed=17, bval=np.array([1.0, 2.0]), bvec=fib_dirs(2))
    kw = dict(dt=2e-4, TN=300, timepoints=10, precision="f32", replicas=4)
    gb = rms.Geometry(lab, 0.1, [0.0, 2.0, 0.8], T2=[np.inf, 70.0, 40.0], kappa=0.2, storage="bricks")
    gd = rms.Geometry(lab, 0.1, [0.0, 2.0, 0.8], T2=[np.inf, 70.0, 40.0], kappa=0.2, storage="dense")
    na = engine.run_noalloc(dict(geom=gb, **job), **kw)
    _same([na], engine.run([dict(geom=gb, **job)], **kw))
    _same([na], engine.run([dict(geom=gd, **job)], **kw))
