"""HIP engine vs the CPU oracle on the same seeds: bit-exact records, counters
and final walker states (the reduction order is fixed on both sides)."""
import numpy as np
import pytest

from helpers import cylinder, fib_dirs, pgse_table, planes, random_blobs

pytestmark = pytest.mark.gpu


def _both(rms, orc, engine, jobs, prec, final_state=True, **kw):
    gjobs, ojobs = [], []
    for jb in jobs:
        g = jb["geom"]
        gg = rms.Geometry(g["labels"], g["res"], g["D"], T2=g.get("T2"), kappa=g.get("kappa"),
                          label_class=g.get("label_class"), periodic=g.get("periodic", ""))
        og = orc.Geometry(g["labels"], g["res"], g["D"], T2=g.get("T2"), kappa=g.get("kappa"),
                          label_class=g.get("label_class"), periodic=g.get("periodic", ""))
        rest = {k: v for k, v in jb.items() if k != "geom"}
        gjobs.append(dict(geom=gg, **rest))
        ojobs.append(dict(geom=og, **rest))
    run_kw = {k: v for k, v in kw.items() if k not in ("dt", "TN", "timepoints", "replicas")}
    gpu = engine.run(gjobs, kw["dt"], kw["TN"], kw["timepoints"], precision=prec, replicas=kw.get("replicas", 32),
                     final_state=final_state, **run_kw)
    ref = orc.run(ojobs, kw["dt"], kw["TN"], kw["timepoints"], prec=prec, replicas=kw.get("replicas", 32),
                  final_state=final_state)
    return gpu, ref


def _assert_same(gpu, ref, final_state=True):
    for a, b in zip(gpu, ref):
        np.testing.assert_array_equal(a["counters"], b["counters"])
        if final_state:
            for k in ("fin_g", "fin_gi", "fin_c", "fin_cnt"):
                np.testing.assert_array_equal(a[k], b[k], err_msg=k)
            for k in ("fin_o", "fin_oi"):
                np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        np.testing.assert_array_equal(a["rec"], b["rec"])
        np.testing.assert_array_equal(a["rec_rep"], b["rec_rep"])
        if b["sig"] is not None:
            np.testing.assert_array_equal(a["sig"], b["sig"])
            np.testing.assert_array_equal(a["sig_rep"], b["sig_rep"])


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_planes_narrow_pulse_bitexact(rms, orc, engine, prec):
    print(f"case planes narrow-pulse {prec}", flush=True)
    lab = planes(8, 8, 14)
    bval = np.array([0.5, 1.0, 3.0])
    bvec = fib_dirs(3)
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0], T2=[80.0, 60.0])
    jobs = [dict(geom=geom, nwalk=1000, seed=12345, bval=bval, bvec=bvec)]
    gpu, ref = _both(rms, orc, engine, jobs, prec, dt=2e-4, TN=500, timepoints=50)
    _assert_same(gpu, ref)
    assert gpu[0]["counters"][1] > 0  # reflections happened


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_exchange_three_classes_bitexact(rms, orc, engine, prec):
    print(f"case exchange 3 classes {prec}", flush=True)
    lab = random_blobs((10, 9, 11), 3, seed=3)
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0, 0.6], T2=[50.0, 80.0, 20.0])
    jobs = [dict(geom=geom, nwalk=777, seed=99)]
    gpu, ref = _both(rms, orc, engine, jobs, prec, dt=2e-4, TN=400, timepoints=40)
    _assert_same(gpu, ref)
    assert gpu[0]["counters"][4] > 0  # exchange draws happened


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_exchange_four_classes_bitexact(rms, orc, engine, prec):
    """Three non-null classes: a neighbour of another class is not one fixed
    class, so every exchange face goes through the march and its draws."""
    print(f"case exchange 4 classes {prec}", flush=True)
    lab = random_blobs((10, 9, 11), 4, seed=8)
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0, 0.6, 1.2], T2=[50.0, 80.0, 20.0, 40.0])
    jobs = [dict(geom=geom, nwalk=640, seed=31)]
    gpu, ref = _both(rms, orc, engine, jobs, prec, dt=2e-4, TN=400, timepoints=40)
    _assert_same(gpu, ref)
    assert gpu[0]["counters"][4] > 0


def test_exchange_draws_split_uniform_bitexact(rms, orc, engine):
    """f32 exchange draw 0 takes its high 16 bits from the direction words and
    its low half from the exchange block only when P lies inside the high
    half's 2^-16 interval.  With kappa = 0.05 both directions draw (P ~ 1e-3),
    6.5e6 draws give ~100 such steps, and noisy blobs give steps with two or
    three exchange faces (draws 1, 2)."""
    print("case exchange draws split uniform f32", flush=True)
    lab = random_blobs((12, 12, 12), 3, seed=7)
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0, 0.6], T2=[50.0, 80.0, 20.0], kappa=0.05)
    jobs = [dict(geom=geom, nwalk=16384, seed=77)]
    gpu, ref = _both(rms, orc, engine, jobs, "f32", dt=2e-4, TN=2000, timepoints=20)
    _assert_same(gpu, ref)
    assert gpu[0]["counters"][4] > 6_000_000


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_pgse_bitexact(rms, orc, engine, prec):
    print(f"case pgse {prec}", flush=True)
    lab = cylinder(16, 6, 6.0)
    grad = pgse_table(fib_dirs(4), [0.5, 2.0], [0.04, 0.06], 0.02)
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0], T2=[80.0, 80.0])
    jobs = [dict(geom=geom, nwalk=640, seed=5, grad=grad)]
    gpu, ref = _both(rms, orc, engine, jobs, prec, dt=2e-4, TN=450, timepoints=45)
    _assert_same(gpu, ref)


def test_batched_jobs_bitexact_and_match_single(rms, orc, engine):
    print("case batched jobs", flush=True)
    lab1 = planes(6, 7, 10)
    lab2 = random_blobs((9, 8, 7), 3, seed=8, dtype=np.uint16)
    lab3 = (cylinder(14, 5, 5.0) * 7).astype(np.uint16)  # label 7 -> class 1 via map
    cmap = np.zeros(8, np.uint8)
    cmap[7] = 1
    g1 = dict(labels=lab1, res=0.1, D=[0.0, 2.0])
    g2 = dict(labels=lab2, res=0.1, D=[0.0, 1.5, 2.0], T2=[60.0, 60.0, 80.0], kappa=0.05)
    g3 = dict(labels=lab3, res=0.1, D=[0.0, 2.0], label_class=cmap)
    jobs = [dict(geom=g1, nwalk=130, seed=1), dict(geom=g2, nwalk=257, seed=2), dict(geom=g3, nwalk=64, seed=3)]
    gpu, ref = _both(rms, orc, engine, jobs, "f32", dt=2e-4, TN=300, timepoints=30)
    _assert_same(gpu, ref)
    single, _ = _both(rms, orc, engine, jobs[1:2], "f32", dt=2e-4, TN=300, timepoints=30)
    np.testing.assert_array_equal(single[0]["rec"], gpu[1]["rec"])


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_mixed_tables_in_one_plan_bitexact(rms, orc, engine, prec):
    """One plan with a PGSE job, a job without tables and a narrow-pulse job:
    the plan runs the PGSE kernel for all three, and a job without a PGSE
    table has no lobe boundaries to read.  Single launch and chunks."""
    print(f"case mixed tables {prec}", flush=True)
    grad = pgse_table(fib_dirs(3), [0.5, 1.5], [0.03], 0.01)
    g1 = dict(labels=cylinder(10, 5, 4.0), res=0.1, D=[0.0, 2.0], T2=[80.0, 80.0], periodic="z")
    g2 = dict(labels=random_blobs((5, 3, 12), 3, seed=9), res=0.1, D=[0.0, 2.0, 0.7], T2=[60.0, 60.0, 80.0],
              kappa=0.5)
    g3 = dict(labels=random_blobs((3, 3, 8), 2, seed=10), res=0.1, D=[0.0, 1.2], T2=[50.0, 50.0])
    jobs = [dict(geom=g1, nwalk=112, seed=1, grad=grad), dict(geom=g2, nwalk=16, seed=2),
            dict(geom=g3, nwalk=100, seed=3, bval=np.array([1.0, 2.0]), bvec=fib_dirs(2))]
    base, ref = _both(rms, orc, engine, jobs, prec, dt=2e-4, TN=242, timepoints=8)
    _assert_same(base, ref)
    ch, _ = _both(rms, orc, engine, jobs, prec, dt=2e-4, TN=242, timepoints=8, chunk_steps=60)
    _assert_same(ch, ref)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_chunked_and_graph_equal_single_launch(rms, orc, engine, prec):
    print(f"case chunked/graph {prec}", flush=True)
    lab = random_blobs((8, 8, 8), 3, seed=4)
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0, 1.0], T2=[80.0, 70.0, 40.0])
    grad = pgse_table(fib_dirs(3), [1.0], [0.03], 0.01)
    bval = np.array([1.0])
    jobs = [dict(geom=geom, nwalk=500, seed=21, grad=grad, bval=bval, bvec=fib_dirs(1))]
    base, ref = _both(rms, orc, engine, jobs, prec, dt=2e-4, TN=301, timepoints=30)
    _assert_same(base, ref)
    ch, _ = _both(rms, orc, engine, jobs, prec, dt=2e-4, TN=301, timepoints=30, chunk_steps=77)
    _assert_same(ch, ref)
    gr, _ = _both(rms, orc, engine, jobs, prec, final_state=False, dt=2e-4, TN=301, timepoints=30,
                  chunk_steps=100, use_graph=True)
    _assert_same(gr, ref, final_state=False)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_periodic_edges_bitexact(rms, orc, engine, prec):
    """Periodic lattice edges: a crossing wraps to the voxel on the other
    side (its class decides cross / reflect / draw) and moves the origin by
    one lattice length, which the PGSE sums follow; x and z periodic, y a
    mirror.  Single launch and chunks of 250 steps."""
    print(f"case periodic {prec}", flush=True)
    lab = random_blobs((10, 9, 11), 3, seed=5)
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0, 0.6], T2=[50.0, 80.0, 20.0], kappa=0.05, periodic="xz")
    grad = pgse_table(fib_dirs(3), [0.5, 1.5], [0.03], 0.01)
    jobs = [dict(geom=geom, nwalk=640, seed=52, grad=grad, bval=np.array([1.0]), bvec=fib_dirs(1))]
    base, ref = _both(rms, orc, engine, jobs, prec, dt=2e-4, TN=700, timepoints=35)
    _assert_same(base, ref)
    assert base[0]["counters"][2] > 0  # mirrors on y
    gi = base[0]["fin_gi"]
    assert np.any((gi[0] < 0) | (gi[0] >= lab.shape[0]))  # origins moved by wraps in x
    ch, _ = _both(rms, orc, engine, jobs, prec, dt=2e-4, TN=700, timepoints=35, chunk_steps=250)
    _assert_same(ch, ref)


def test_pgse_sums_flush_period_bitexact(rms, orc, engine):
    """The PGSE sums keep sum(g) in int32 between flushes at lobe boundaries
    and every 2^16 steps.  70000 steps cross the 2^16 flush; the z lattice
    edges of the cylinder mirror the origin between flushes; chunks of 30000
    steps carry the unflushed sums across launches."""
    print("case pgse flush period f32", flush=True)
    lab = cylinder(16, 6, 6.0)
    grad = pgse_table(fib_dirs(3), [1.0], [0.03], 0.01)
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0], T2=[80.0, 80.0])
    jobs = [dict(geom=geom, nwalk=64, seed=606, grad=grad)]
    base, ref = _both(rms, orc, engine, jobs, "f32", dt=2e-4, TN=70000, timepoints=7)
    _assert_same(base, ref)
    assert base[0]["counters"][2] > 1000  # mirrors happened
    ch, _ = _both(rms, orc, engine, jobs, "f32", dt=2e-4, TN=70000, timepoints=7, chunk_steps=30000)
    _assert_same(ch, ref)


def test_pgse_flush_periods_differ_between_jobs_bitexact(rms, orc, engine):
    """A 3 x 3 x 40000 lattice flushes its int32 sums every 2^15 steps (2^16
    would let sum(g) reach 2^31), a small cylinder every 2^16: both in one
    plan over 70000 steps, single launch and chunks of 25000."""
    print("case pgse flush periods per job f32", flush=True)
    tube = np.ones((3, 3, 40000), np.uint8)
    tube[1, 1, ::7] = 0
    grad = pgse_table(fib_dirs(2), [1.0], [5.0], 2.0)
    g1 = dict(labels=tube, res=0.1, D=[0.0, 2.0], T2=[80.0, 80.0], periodic="z")
    g2 = dict(labels=cylinder(10, 6, 4.0), res=0.1, D=[0.0, 2.0], T2=[80.0, 80.0])
    jobs = [dict(geom=g1, nwalk=64, seed=71, grad=grad), dict(geom=g2, nwalk=64, seed=72, grad=grad)]
    base, ref = _both(rms, orc, engine, jobs, "f32", dt=2e-4, TN=70000, timepoints=7)
    _assert_same(base, ref)
    ch, _ = _both(rms, orc, engine, jobs, "f32", dt=2e-4, TN=70000, timepoints=7, chunk_steps=25000)
    _assert_same(ch, ref)


def _adj_jobs(rms, orc, jobs):
    gj, oj = [], []
    for jb in jobs:
        g = jb["geom"]
        mk = lambda M: M.Geometry(g["labels"], g["res"], g["D"], T2=g.get("T2"), kappa=g.get("kappa"),  # noqa: E731
                                  label_class=g.get("label_class"), periodic=g.get("periodic", ""))
        rest = {k: v for k, v in jb.items() if k != "geom"}
        gj.append(dict(geom=mk(rms), **rest))
        oj.append(dict(geom=mk(orc), **rest))
    return gj, oj


@pytest.mark.parametrize("prec,kappa,periodic", [("f32", 0.08, ""), ("f64", 0.08, ""), ("f32", 0.0, ""),
                                                 ("f64", 0.0, ""), ("f32", 0.08, "xyz")])
def test_adjoint_bitexact(rms, orc, engine, prec, kappa, periodic):
    """Tangent pass (pathwise D, score-function kappa, analytic T2) against
    the oracle's adjoint on the same seeds: gradients and their replica
    partials are bit-identical.  The plan has a PGSE table and the voxelised
    blobs have voxel-edge order effects (and mirror edges unless periodic), so
    the D entries are NaN on both sides; kappa = 0 closes the membranes (no
    draw, kappa NaN)."""
    print(f"case adjoint {prec} kappa={kappa} periodic={periodic}", flush=True)
    lab = random_blobs((10, 9, 11), 3, seed=3)
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0, 0.6], T2=[50.0, 80.0, 20.0], kappa=kappa, periodic=periodic)
    bval = np.array([0.5, 2.0])
    bvec = fib_dirs(2)
    grad = pgse_table(fib_dirs(3), [0.5, 1.5], [0.03], 0.01)
    jobs = [dict(geom=geom, nwalk=700, seed=77, bval=bval, bvec=bvec, grad=grad)]
    dt, TN, T = 2e-4, 300, 30
    nval = rms.nval(3, 2)
    rng = np.random.default_rng(0)
    cot_rec = [rng.standard_normal((T, nval))]
    cot_sig = [rng.standard_normal(grad.shape[0])]
    gj, oj = _adj_jobs(rms, orc, jobs)
    plan = engine.plan(gj, dt, TN, T, precision=prec, replicas=8)
    try:
        fwd = plan.execute(final_state=False)
        out = plan.adjoint(cot_rec, cot_sig, want_replicas=True, estimator="boundary")
        out2 = plan.adjoint(cot_rec, None, want_replicas=True, estimator="boundary")
    finally:
        plan.close()
    ref = orc.run(oj, dt, TN, T, prec=prec, replicas=8, cot_rec=cot_rec, cot_sig=cot_sig)
    ref2 = orc.run(oj, dt, TN, T, prec=prec, replicas=8, cot_rec=cot_rec, cot_sig=None)
    np.testing.assert_array_equal(fwd[0]["rec"], ref[0]["rec"])
    np.testing.assert_array_equal(out[0]["grad"], ref[0]["grad"])
    np.testing.assert_array_equal(out[0]["grad_rep"], ref[0]["grad_rep"])
    np.testing.assert_array_equal(out2[0]["grad"], ref2[0]["grad"])
    g = out[0]["grad"]
    assert np.all(np.isfinite(g[4:8])) and g[5] != 0.0 and g[6] != 0.0
    assert np.all(np.isnan(g[1:4]))
    if kappa > 0:
        assert np.isfinite(g[8]) and g[8] != 0.0
    else:
        assert np.isnan(g[8])


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_adjoint_pgse_pathwise_bitexact(rms, orc, engine, prec):
    """PGSE tangent pass where the pathwise D derivative is the whole
    derivative (periodic planes: no mirror, no voxel-edge order effect): D,
    T2 and their replica partials bit-identical to the oracle, D finite."""
    print(f"case adjoint PGSE pathwise {prec}", flush=True)
    lab = planes(7, 6, 12)
    geom = dict(labels=lab, res=0.1, D=[0.0, 1.8], T2=[50.0, 60.0], periodic="xyz")
    bval, bvec = np.array([1.5]), fib_dirs(1)
    grad = pgse_table(fib_dirs(3), [0.5, 1.5], [0.03], 0.01)
    jobs = [dict(geom=geom, nwalk=600, seed=21, bval=bval, bvec=bvec, grad=grad)]
    dt, TN, T = 2e-4, 300, 30
    rng = np.random.default_rng(4)
    cot_rec = [rng.standard_normal((T, rms.nval(2, 1)))]
    cot_sig = [rng.standard_normal(grad.shape[0])]
    gj, oj = _adj_jobs(rms, orc, jobs)
    plan = engine.plan(gj, dt, TN, T, precision=prec, replicas=8)
    try:
        out = plan.adjoint(cot_rec, cot_sig, want_replicas=True, estimator="boundary")
    finally:
        plan.close()
    ref = orc.run(oj, dt, TN, T, prec=prec, replicas=8, cot_rec=cot_rec, cot_sig=cot_sig)
    np.testing.assert_array_equal(out[0]["grad"], ref[0]["grad"])
    np.testing.assert_array_equal(out[0]["grad_rep"], ref[0]["grad_rep"])
    g = out[0]["grad"]
    assert np.isfinite(g[1]) and g[1] != 0.0 and np.isfinite(g[5]) and g[5] != 0.0


def test_adjoint_batched_jobs_bitexact(rms, orc, engine):
    print("case adjoint batched jobs", flush=True)
    g1 = dict(labels=planes(6, 7, 10), res=0.1, D=[0.0, 2.0], T2=[70.0, 70.0])
    g2 = dict(labels=random_blobs((9, 8, 7), 3, seed=8, dtype=np.uint16), res=0.1, D=[0.0, 1.5, 2.0],
              T2=[60.0, 60.0, 80.0], kappa=0.05)
    jobs = [dict(geom=g1, nwalk=130, seed=1), dict(geom=g2, nwalk=257, seed=2)]
    dt, TN, T = 2e-4, 240, 24
    rng = np.random.default_rng(1)
    cot = [rng.standard_normal((T, rms.nval(2, 0))), rng.standard_normal((T, rms.nval(3, 0)))]
    gj, oj = _adj_jobs(rms, orc, jobs)
    plan = engine.plan(gj, dt, TN, T, precision="f32", replicas=4)
    try:
        out = plan.adjoint(cot, None, want_replicas=True, estimator="boundary")
    finally:
        plan.close()
    ref = orc.run(oj, dt, TN, T, prec="f32", replicas=4, cot_rec=cot)
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a["grad"], b["grad"])
        np.testing.assert_array_equal(a["grad_rep"], b["grad_rep"])


def test_graph_replays_match_eager(rms, orc, engine):
    """A captured plan replayed several times, with eager runs of another plan
    in between, gives the eager result on every replay."""
    print("case graph replays", flush=True)
    lab = random_blobs((8, 8, 8), 3, seed=6)
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0, 1.0], T2=[80.0, 70.0, 40.0])
    grad = pgse_table(fib_dirs(2), [1.0], [0.03], 0.01)
    jobs = [dict(geom=geom, nwalk=300, seed=8, grad=grad)]
    gj, _ = _adj_jobs(rms, orc, jobs)
    eager = engine.run(gj, 2e-4, 200, 20, precision="f32", replicas=4)
    gp = engine.plan(gj, 2e-4, 200, 20, precision="f32", replicas=4, chunk_steps=64, use_graph=True)
    try:
        for k in range(4):
            r = gp.execute(final_state=False)
            np.testing.assert_array_equal(r[0]["rec"], eager[0]["rec"], err_msg=f"replay {k}")
            np.testing.assert_array_equal(r[0]["sig"], eager[0]["sig"], err_msg=f"replay {k}")
            np.testing.assert_array_equal(r[0]["counters"], eager[0]["counters"], err_msg=f"replay {k}")
            other = engine.run(gj, 2e-4, 100, 10, precision="f64", replicas=2)
            assert np.isfinite(other[0]["rec"]).all()
    finally:
        gp.close()


def test_autograd_matches_adjoint(rms, orc, engine):
    """torch.autograd through autograd.simulate: the backward pass is the
    adjoint pass, bit-identical to the oracle's gradient."""
    import importlib
    import torch
    ag = importlib.import_module(rms.__name__ + ".autograd")
    print("case autograd", flush=True)
    lab = random_blobs((9, 9, 9), 3, seed=12)
    grad = pgse_table(fib_dirs(2), [1.0], [0.03], 0.01)
    dt, TN, T = 2e-4, 200, 10
    D = torch.tensor([0.0, 2.0, 0.9], dtype=torch.float64, requires_grad=True)
    T2 = torch.tensor([40.0, 70.0, 30.0], dtype=torch.float64, requires_grad=True)
    kap = torch.tensor(0.0, dtype=torch.float64, requires_grad=True)
    rec, sig = ag.simulate(engine, lab, 0.1, D, T2, nwalk=300, dt=dt, TN=TN, timepoints=T, kappa=kap, seed=4,
                           grad=grad, precision="f32", replicas=4, d_estimator="boundary")
    rng = np.random.default_rng(2)
    wr = torch.from_numpy(rng.standard_normal(tuple(rec.shape)))
    ws = torch.from_numpy(rng.standard_normal(tuple(sig.shape)))
    loss = (wr * rec).sum() + (ws * sig).sum()
    loss.backward()
    og = orc.Geometry(lab, 0.1, [0.0, 2.0, 0.9], T2=[40.0, 70.0, 30.0], kappa=0.0)
    ref = orc.run([dict(geom=og, nwalk=300, seed=4, grad=grad)], dt, TN, T, prec="f32", replicas=4,
                  cot_rec=[wr.numpy()], cot_sig=[ws.numpy()])[0]
    np.testing.assert_array_equal(rec.detach().numpy(), ref["rec"])
    np.testing.assert_array_equal(D.grad.numpy(), ref["grad"][:3])
    np.testing.assert_array_equal(T2.grad.numpy(), ref["grad"][4:7])
    # kappa = 0 closes the membranes: no kappa estimate on either side
    assert np.isnan(float(kap.grad)) and np.isnan(ref["grad"][8])


def test_autotuned_launch_gives_identical_results(rms, engine):
    """The autotuner's launch shape and job order change the schedule only:
    every job's records are bit-identical to the default launch."""
    import importlib
    at = importlib.import_module(rms.__name__ + ".autotune")
    print("case autotune", flush=True)
    labs = [random_blobs((9, 8, 10), 3, seed=s, dtype=np.uint16) for s in (1, 2, 3)] + [planes(6, 6, 20)]
    jobs = []
    for k, lab in enumerate(labs):
        nc = 3 if lab.dtype == np.uint16 else 2
        g = rms.Geometry(lab, 0.1, [0.0, 2.0, 1.5][:nc], T2=[np.inf, 80.0, 60.0][:nc],
                         kappa=0.02 if nc == 3 else None)
        jobs.append(dict(geom=g, nwalk=200 + 97 * k, seed=100 + k))
    best = at.tune(engine, jobs, 2e-4, timepoints=10, precision="f32", probe_steps=50, reps=1, use_cache=False)
    assert best["block_threads"] in (64, 128, 256) and best["grid_blocks"] >= 8
    assert sorted(best["order"]) == list(range(len(jobs)))
    base = engine.run(jobs, 2e-4, 300, 30, precision="f32", replicas=4)
    ordered = [jobs[k] for k in best["order"]]
    tuned = engine.run(ordered, 2e-4, 300, 30, precision="f32", replicas=4, block_threads=best["block_threads"],
                       grid_blocks=best["grid_blocks"], xcd_split=best["xcd_split"])
    for pos, k in enumerate(best["order"]):
        np.testing.assert_array_equal(tuned[pos]["rec"], base[k]["rec"])
        np.testing.assert_array_equal(tuned[pos]["rec_rep"], base[k]["rec_rep"])
        np.testing.assert_array_equal(tuned[pos]["counters"], base[k]["counters"])


def test_near_max_lattice_bitexact(rms, orc, engine):
    """The largest lattice the host accepts is below 2^31 voxels (the
    reference's int index): 1200 x 1200 x 1491 = 2.147e9 voxels (2 GB of
    labels, 17 GB of cell words), label-0 planes every 100 slices, periodic
    x.  Voxel indices near 2^31 - 1 in the walk, the cell build and the
    march stay in range: bit-exact against the oracle."""
    print("case near-max lattice 1200x1200x1491 f32", flush=True)
    lab = np.ones((1200, 1200, 1491), np.uint8)
    lab[:, :, ::100] = 0
    assert lab.size < 2 ** 31 and lab.size > 2 ** 31 - 2 ** 21
    geom = dict(labels=lab, res=0.1, D=[0.0, 2.0], T2=[80.0, 80.0], periodic="x")
    jobs = [dict(geom=geom, nwalk=256, seed=2147, bval=np.array([1.0]), bvec=fib_dirs(1))]
    gpu, ref = _both(rms, orc, engine, jobs, "f32", dt=2e-4, TN=300, timepoints=10)
    _assert_same(gpu, ref)
    assert int(np.max(gpu[0]["fin_g"][0])) > 1000 and int(np.max(gpu[0]["fin_g"][2])) > 1400


def test_autograd_many_geometries_equals_one_at_a_time(rms, engine):
    """autograd.simulate_many: three geometries with their own D, T2 and
    kappa in one plan; the records and the gradients of each (default D
    estimator: central differences where exchange makes the walk jump) are
    bit-identical to autograd.simulate on that geometry alone."""
    import importlib
    import torch
    ag = importlib.import_module(rms.__name__ + ".autograd")
    print("case autograd many geometries", flush=True)
    labs = [random_blobs((9, 8, 10), 3, seed=40 + k) for k in range(3)]
    bval, bvec = np.array([0.7, 2.0]), fib_dirs(2)
    jobs = [dict(labels=labs[k], nwalk=400 + 100 * k, seed=7 + k, bval=bval, bvec=bvec) for k in range(3)]
    D0 = np.array([[0.0, 2.0, 0.9], [0.0, 1.6, 1.1], [0.0, 2.2, 0.5]])
    T20 = np.array([[np.inf, 70.0, 30.0], [np.inf, 60.0, 40.0], [np.inf, 90.0, 25.0]])
    k0 = np.array([0.05, 0.2, 0.1])
    dt, TN, T = 2e-4, 300, 10
    D = torch.tensor(D0, requires_grad=True)
    T2 = torch.tensor(T20, requires_grad=True)
    kap = torch.tensor(k0, requires_grad=True)
    rec, sig = ag.simulate_many(engine, jobs, D, T2, dt, TN, timepoints=T, kappa=kap, replicas=4)
    assert tuple(rec.shape) == (3, T, rms.nval(3, 2)) and tuple(sig.shape) == (3, 0)
    wr = torch.from_numpy(np.random.default_rng(6).standard_normal(tuple(rec.shape)))
    (wr * rec).sum().backward()
    for k in range(3):
        Dk = torch.tensor(D0[k], requires_grad=True)
        Tk = torch.tensor(T20[k], requires_grad=True)
        kk = torch.tensor(k0[k], requires_grad=True)
        r1, _ = ag.simulate(engine, labs[k], 0.1, Dk, Tk, nwalk=jobs[k]["nwalk"], dt=dt, TN=TN, timepoints=T,
                            kappa=kk, seed=jobs[k]["seed"], bval=bval, bvec=bvec, replicas=4)
        (wr[k] * r1).sum().backward()
        np.testing.assert_array_equal(rec[k].detach().numpy(), r1.detach().numpy())
        np.testing.assert_array_equal(D.grad[k].numpy(), Dk.grad.numpy())
        np.testing.assert_array_equal(T2.grad[k, 1:].numpy(), Tk.grad[1:].numpy())
        assert float(kap.grad[k]) == float(kk.grad)
        assert np.all(np.isfinite(D.grad[k, 1:].numpy())) and np.isfinite(float(kap.grad[k]))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_periodic_sides_both_encodings_bitexact(rms, orc, engine, prec):
    """Steps through a periodic side are walked by the fast path when more
    than 1/256 of the non-null voxels lie on a periodic side (wrapped cell
    words), by the march otherwise: the library's choice equals the oracle's
    rule, and both lattices are bit-exact against the oracle."""
    print(f"case periodic sides, both encodings, {prec}", flush=True)
    # a tube along z through a 20 x 20 x 300 lattice: 2 of 300 slices on the
    # z sides (0.7 % > 1/256, wrapped); a 30 x 30 x 600 one with the tube
    # ending one voxel short of the z sides and touching them through a
    # narrow neck (under 1/256, marched)
    def tube(n, nz, r, neck):
        x, y = np.meshgrid(np.arange(n) - n / 2 + 0.5, np.arange(n) - n / 2 + 0.5, indexing="ij")
        lab = np.zeros((n, n, nz), np.uint8)
        lab[x ** 2 + y ** 2 <= r * r] = 1
        if neck:
            lab[:, :, 0] = 0
            lab[:, :, -1] = 0
            lab[n // 2, n // 2, 0] = lab[n // 2, n // 2, -1] = 1
        return lab
    for lab, want in ((tube(20, 300, 5.0, False), True), (tube(30, 600, 8.0, True), False)):
        g = rms.Geometry(lab, 0.1, [0.0, 2.0], T2=[np.inf, 70.0], periodic="xyz")
        og = orc.Geometry(lab, 0.1, [0.0, 2.0], T2=[np.inf, 70.0], periodic="xyz")
        assert engine.geometry_info(g, wrap_cells=True)[2] is want
        assert orc.lib().or_wrap_cells_rule(og._c()) == int(want)
        job = dict(nwalk=3000, seed=5, bval=np.array([1.0]), bvec=fib_dirs(1),
                   grad=pgse_table(fib_dirs(2), [1.0], [0.03], 0.01))
        kw = dict(dt=2e-4, TN=4000, timepoints=20)
        out = engine.run([dict(geom=g, **job)], precision=prec, replicas=4, final_state=True, **kw)[0]
        ref = orc.run([dict(geom=og, **job)], kw["dt"], kw["TN"], kw["timepoints"], prec=prec, replicas=4,
                      final_state=True)[0]
        np.testing.assert_array_equal(out["rec"], ref["rec"])
        np.testing.assert_array_equal(out["sig"], ref["sig"])
        np.testing.assert_array_equal(out["fin_gi"], ref["fin_gi"])
        assert np.any(np.abs(out["fin_gi"][2]) >= lab.shape[2] - 1) or not want  # walkers wrapped in z


@pytest.mark.parametrize("prec,exch,chunk,graph", [("f32", False, 0, False), ("f32", True, 77, False),
                                                  ("f64", True, 0, False), ("f32", True, 100, True)])
def test_skipped_counters_leave_every_other_output_unchanged(rms, engine, prec, exch, chunk, graph):
    """counters=False (rms_run_params.skip_counters): the diagnostic counters
    are zero and records, replica sums and the PGSE signal bit-identical to
    the counted run, single launch, chunks or a captured graph; asking for
    the final walker state counts again (its per-walker counters)."""
    print(f"case skipped counters {prec} exchange={exch} chunk={chunk} graph={graph}", flush=True)
    lab = random_blobs((9, 8, 10), 3 if exch else 2, seed=8)
    g = rms.Geometry(lab, 0.1, [0.0, 2.0, 0.9][:3 if exch else 2], T2=[np.inf, 70.0, 40.0][:3 if exch else 2],
                     kappa=0.2 if exch else None, periodic="x")
    jobs = [dict(geom=g, nwalk=700, seed=4, grad=pgse_table(fib_dirs(2), [1.0], [0.03], 0.01),
                 bval=np.array([1.0, 2.0]), bvec=fib_dirs(2))]
    kw = dict(precision=prec, replicas=4, chunk_steps=chunk, use_graph=graph)
    on = engine.plan(jobs, 2e-4, 300, 12, **kw)
    off = engine.plan(jobs, 2e-4, 300, 12, counters=False, **kw)
    try:
        a = on.execute()[0]
        b = off.execute()[0]
        b2 = off.execute()[0]  # a second (graph) replay
        c = off.execute(final_state=True)[0]
        d = on.execute(final_state=True)[0]
    finally:
        on.close()
        off.close()
    assert a["counters"][0] > 0 and a["counters"][1] > 0
    for r in (b, b2):
        assert np.all(r["counters"] == 0)
        for key in ("rec", "rec_rep", "sig", "sig_rep"):
            np.testing.assert_array_equal(r[key], a[key])
    np.testing.assert_array_equal(c["counters"], a["counters"])
    for key in ("rec", "fin_g", "fin_o", "fin_cnt", "fin_c"):
        np.testing.assert_array_equal(c[key], d[key])


def test_skipped_counters_allocation_free(rms, engine):
    """The allocation-free run with counters=False: counters zero, records
    equal to the counted allocation-free run."""
    print("case skipped counters, allocation-free run", flush=True)
    lab = random_blobs((8, 9, 7), 3, seed=12)
    g = rms.Geometry(lab, 0.1, [0.0, 2.0, 0.9], T2=[np.inf, 70.0, 40.0], kappa=0.1)
    job = dict(geom=g, nwalk=600, seed=2, bval=np.array([1.5]), bvec=fib_dirs(1))
    a = engine.run_noalloc(job, 2e-4, 200, 10, replicas=4)
    b = engine.run_noalloc(job, 2e-4, 200, 10, replicas=4, counters=False)
    assert a["counters"][0] > 0 and np.all(b["counters"] == 0)
    np.testing.assert_array_equal(a["rec"], b["rec"])
    np.testing.assert_array_equal(a["rec_rep"], b["rec_rep"])


@pytest.mark.parametrize("prec,case", [("f32", "axon"), ("f64", "axon"), ("f32", "cross"), ("f32", "closed"),
                                       ("f32", "bricks")])
def test_lean_walks_equal_the_full_variant(rms, orc, engine, prec, case):
    """Plans whose classes all have infinite T2 and whose pairs never draw
    take, without counters, the forward variants that keep no time-in-class
    sums and no draw bits (weight exp(-0) = 1): records, replica sums and
    the PGSE signal bit-identical to the counted run (the full variant) and
    to the oracle.  "cross": two classes with equal D and no permeability
    (P = 1 both ways: class changes without draws); "closed": kappa = 0."""
    print(f"case lean walk {prec} {case}", flush=True)
    if case == "axon":
        lab = cylinder(20, 12, 6.0)
        D, kw = [0.0, 2.0], {}
    else:
        lab = random_blobs((10, 9, 11), 3, seed=31)
        D, kw = [0.0, 1.5, 1.5 if case != "closed" else 0.8], ({} if case == "cross" else dict(kappa=0.0))
    T2 = [np.inf] * len(D)
    st = "bricks" if case == "bricks" else "auto"
    if case == "bricks":
        lab = random_blobs((10, 9, 11), 3, seed=31)
        D, kw = [0.0, 1.5, 1.5], {}
    grad = pgse_table(fib_dirs(2), [1.0], [0.03], 0.01)
    mk = lambda M, **x: M.Geometry(lab, 0.1, D, T2=T2, periodic="xz", **kw, **x)  # noqa: E731
    job = dict(nwalk=900, seed=6, grad=grad, bval=np.array([1.0, 2.0]), bvec=fib_dirs(2))
    full = engine.run([dict(geom=mk(rms, storage=st), **job)], 2e-4, 400, 20, precision=prec, replicas=4)[0]
    lean = engine.run([dict(geom=mk(rms, storage=st), **job)], 2e-4, 400, 20, precision=prec, replicas=4,
                      counters=False)[0]
    ref = orc.run([dict(geom=mk(orc), **job)], 2e-4, 400, 20, prec=prec, replicas=4)[0]
    # the same lean walk in launches of 130 steps (walker state saved between
    # them), replayed as a captured graph
    chunked = engine.run([dict(geom=mk(rms, storage=st), **job)], 2e-4, 400, 20, precision=prec, replicas=4,
                         counters=False, chunk_steps=130, use_graph=True)[0]
    for key in ("rec", "rec_rep", "sig", "sig_rep"):
        np.testing.assert_array_equal(lean[key], full[key], err_msg=key)
        np.testing.assert_array_equal(lean[key], ref[key], err_msg=key)
        np.testing.assert_array_equal(chunked[key], ref[key], err_msg="chunked " + key)
    assert np.all(lean["counters"] == 0) and full["counters"][1] > 0
