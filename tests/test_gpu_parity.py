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
                          label_class=g.get("label_class"))
        og = orc.Geometry(g["labels"], g["res"], g["D"], T2=g.get("T2"), kappa=g.get("kappa"),
                          label_class=g.get("label_class"))
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
