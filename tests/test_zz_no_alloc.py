"""rms_run_noalloc: the allocation-free run (every device buffer in the
library's static scratch array) gives the same results as rms_run on the
same job, bit for bit, on two calls in a row (the second smaller, so it
reuses the scratch the first left behind), and refuses a run that does not
fit the array before launching anything."""
import numpy as np
import pytest

from helpers import fib_dirs, pgse_table, random_blobs

pytestmark = pytest.mark.gpu


def _same(a, b):
    np.testing.assert_array_equal(a["counters"], b["counters"])
    np.testing.assert_array_equal(a["rec"], b["rec"])
    np.testing.assert_array_equal(a["rec_rep"], b["rec_rep"])
    if b["sig"] is not None:
        np.testing.assert_array_equal(a["sig"], b["sig"])
        np.testing.assert_array_equal(a["sig_rep"], b["sig_rep"])


def test_noalloc_equals_run_twice_second_smaller(rms, engine):
    lab = random_blobs((12, 11, 13), 3, seed=40)
    geom = rms.Geometry(lab, 0.1, [0.0, 2.0, 1.1], T2=[np.inf, 70.0, 35.0], kappa=0.05)
    grad = pgse_table(fib_dirs(5), [0.5, 2.0], [0.03, 0.05], 0.012)
    bval, bvec = np.array([0.5, 1.5, 3.0]), fib_dirs(3)
    big = dict(geom=geom, nwalk=3000, seed=17, grad=grad, bval=bval, bvec=bvec)
    print("case noalloc call 1: 3000 walkers, PGSE + b-table, f32, 60 records", flush=True)
    a = engine.run_noalloc(big, 2e-4, 600, 60, precision="f32", replicas=32)
    b = engine.run([big], 2e-4, 600, 60, precision="f32", replicas=32)[0]
    _same(a, b)
    assert a["counters"][4] > 0
    print("case noalloc call 2: 500 walkers, b-table only, f64, 20 records", flush=True)
    small = dict(geom=geom, nwalk=500, seed=3, bval=bval[:2], bvec=bvec[:2])
    a = engine.run_noalloc(small, 2e-4, 200, 20, precision="f64", replicas=8)
    b = engine.run([small], 2e-4, 200, 20, precision="f64", replicas=8)[0]
    _same(a, b)
    print("case noalloc call 3: 300 walkers, no tables, f32", flush=True)
    plain = dict(geom=geom, nwalk=300, seed=5)
    a = engine.run_noalloc(plain, 2e-4, 100, 10, precision="f32", replicas=4)
    b = engine.run([plain], 2e-4, 100, 10, precision="f32", replicas=4)[0]
    _same(a, b)
    print("case noalloc call 4: periodic lattice with wrapped cell words, 700 walkers, f32", flush=True)
    gp = rms.Geometry(lab, 0.1, [0.0, 2.0, 1.1], T2=[np.inf, 70.0, 35.0], kappa=0.05, periodic="xyz")
    assert engine.geometry_info(gp, wrap_cells=True)[2]
    per = dict(geom=gp, nwalk=700, seed=9, bval=bval, bvec=bvec)
    a = engine.run_noalloc(per, 2e-4, 400, 20, precision="f32", replicas=4)
    b = engine.run([per], 2e-4, 400, 20, precision="f32", replicas=4)[0]
    _same(a, b)


def test_noalloc_refuses_runs_beyond_the_scratch(rms, engine):
    print("case noalloc size limit", flush=True)
    capi = __import__(rms.__name__ + "._capi", fromlist=["RmsError"])
    geom = rms.Geometry(np.ones((4, 4, 4), np.uint8), 0.1, [0.0, 2.0])
    # 1e6 walkers x 1000 records x 24 values of partials = 3 GB > 256 MiB
    with pytest.raises(capi.RmsError, match="RMS_SCRATCH_BYTES") as ei:
        engine.run_noalloc(dict(geom=geom, nwalk=1_000_000, seed=1), 2e-4, 1000, 1000)
    assert ei.value.status == 1
    # the array still serves a run that fits
    r = engine.run_noalloc(dict(geom=geom, nwalk=64, seed=1), 2e-4, 100, 10)
    assert r["NPar_count"].sum(1).tolist() == [64.0] * 10
