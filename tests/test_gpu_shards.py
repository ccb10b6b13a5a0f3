"""Walker-range shards (rms_job.walker_base): the walk of walker w uses
Philox walker id w whatever the plan it runs in, so a run split into shards
follows the same trajectories as the whole run, and its batches land in the
same replicas.  This is the single-GPU surface of the multi-GPU plan (one
walker range per GPU, one reduction of the replica sums at the end)."""
import numpy as np
import pytest

from helpers import planes, random_blobs

pytestmark = pytest.mark.gpu

R = 32


def _geoms(rms, orc):
    lab = random_blobs((10, 9, 12), 3, seed=21)
    kw = dict(T2=[np.inf, 70.0, 40.0], kappa=0.05)
    return (rms.Geometry(lab, 0.1, [0.0, 2.0, 1.2], **kw), orc.Geometry(lab, 0.1, [0.0, 2.0, 1.2], **kw))


def test_two_shards_reproduce_the_whole_run_bitexact(rms, orc, engine):
    """2R batches split into two shards of R batches: every replica holds one
    batch of each shard, so shard replica sums add to the whole run's replica
    sums exactly (0 + a = a), and the totals follow in the same order."""
    print("case shards: 4096 walkers as 2 x 2048", flush=True)
    g, og = _geoms(rms, orc)
    bval, bvec = np.array([1.0, 2.0]), np.array([[0, 0, 1.0], [1.0, 0, 0]])
    kw = dict(dt=2e-4, TN=600, timepoints=30)
    whole = engine.run([dict(geom=g, nwalk=4096, seed=5, bval=bval, bvec=bvec)], kw["dt"], kw["TN"],
                       kw["timepoints"], replicas=R, final_state=True)[0]
    shards = [engine.run([dict(geom=g, nwalk=2048, seed=5, bval=bval, bvec=bvec, walker_base=b)], kw["dt"], kw["TN"],
                         kw["timepoints"], replicas=R, final_state=True)[0] for b in (0, 2048)]
    rep = shards[0]["rec_rep"] + shards[1]["rec_rep"]
    np.testing.assert_array_equal(rep, whole["rec_rep"])
    tot = np.zeros_like(whole["rec"])
    for r in range(R):
        tot = tot + rep[r]
    np.testing.assert_array_equal(tot, whole["rec"])
    for k in ("fin_g", "fin_o", "fin_gi", "fin_oi", "fin_cnt"):
        np.testing.assert_array_equal(np.concatenate([s[k] for s in shards], axis=-1), whole[k], err_msg=k)
    np.testing.assert_array_equal(np.concatenate([s["fin_c"] for s in shards]), whole["fin_c"])
    np.testing.assert_array_equal(shards[0]["counters"] + shards[1]["counters"], whole["counters"])
    # the second shard against the oracle with the same walker base
    ref = orc.run([dict(geom=og, nwalk=2048, seed=5, bval=bval, bvec=bvec, walker_base=2048)], kw["dt"], kw["TN"],
                  kw["timepoints"], prec="f32", replicas=R, final_state=True)[0]
    np.testing.assert_array_equal(shards[1]["rec_rep"], ref["rec_rep"])
    np.testing.assert_array_equal(shards[1]["fin_o"], ref["fin_o"])


def test_uneven_shards_reproduce_trajectories(rms, orc, engine):
    """Shards of 2048 + 2952 walkers (PGSE): identical final states, replica
    sums equal up to the order of the additions."""
    from helpers import fib_dirs, pgse_table
    print("case shards: 5000 walkers as 2048 + 2952, PGSE", flush=True)
    lab = planes(8, 8, 16)
    g = rms.Geometry(lab, 0.1, [0.0, 2.0])
    grad = pgse_table(fib_dirs(3), [1.0], [0.05], 0.02)
    kw = dict(dt=2e-4, TN=400, timepoints=20)
    whole = engine.run([dict(geom=g, nwalk=5000, seed=9, grad=grad)], kw["dt"], kw["TN"], kw["timepoints"],
                       replicas=R, final_state=True)[0]
    parts = [engine.run([dict(geom=g, nwalk=n, seed=9, grad=grad, walker_base=b)], kw["dt"], kw["TN"],
                        kw["timepoints"], replicas=R, final_state=True)[0] for b, n in ((0, 2048), (2048, 2952))]
    for k in ("fin_g", "fin_o", "fin_gi", "fin_oi"):
        np.testing.assert_array_equal(np.concatenate([s[k] for s in parts], axis=-1), whole[k], err_msg=k)
    np.testing.assert_allclose(parts[0]["rec_rep"] + parts[1]["rec_rep"], whole["rec_rep"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(parts[0]["sig_rep"] + parts[1]["sig_rep"], whole["sig_rep"], rtol=1e-12, atol=1e-12)
    # a base that is not a multiple of 64 * replicas is refused
    capi = __import__(rms.__name__ + "._capi", fromlist=["RmsError"])
    with pytest.raises(capi.RmsError, match="walker_base"):
        engine.run([dict(geom=g, nwalk=64, seed=9, walker_base=64)], kw["dt"], kw["TN"], kw["timepoints"], replicas=R)
