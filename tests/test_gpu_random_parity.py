"""Randomised parity: 300 seeded random valid inputs -- lattice shapes 1..12
per side, 2..4 classes (uint8 or uint16 labels), random D, T2, kappa (or the
reference rule), periodic axes, FP32/FP64, narrow-pulse or PGSE tables,
chunked launches -- each run by the HIP engine and the oracle and compared
bit for bit (records, replicas, counters, final walker states, PGSE signal);
the adjoint on 100 of them, random multi-job plans, and the central
differences of D on walker subsets in 100 random plans."""
import numpy as np
import pytest

from helpers import fib_dirs, pgse_table, random_blobs

pytestmark = pytest.mark.gpu


def _case(rng):
    shape = tuple(int(x) for x in rng.integers(1, 13, 3))
    nclass = int(rng.integers(2, 5))
    lab = random_blobs(shape, nclass, seed=int(rng.integers(0, 1 << 30)), p_null=float(rng.uniform(0, 0.4)),
                       dtype=np.uint16 if rng.random() < 0.3 else np.uint8)
    if not np.any(lab != 0):
        lab.flat[0] = 1
    D = [0.0] + [float(x) for x in rng.uniform(0.3, 3.0, nclass - 1)]
    T2 = [float(x) for x in rng.uniform(20, 100, nclass)]
    kappa = None if rng.random() < 0.4 else float(rng.choice([0.0, 0.01, 0.05, 0.5]))
    periodic = "".join(a for a in "xyz" if rng.random() < 0.3)
    prec = "f32" if rng.random() < 0.6 else "f64"
    TN = int(rng.integers(20, 500))
    T = int(rng.integers(1, min(TN, 60) + 1))
    job = dict(nwalk=int(rng.integers(1, 400)), seed=int(rng.integers(0, 1 << 62)))
    r = rng.random()
    if r < 0.35:
        job["grad"] = pgse_table(fib_dirs(int(rng.integers(1, 4))), [float(rng.uniform(0.2, 2))], [TN * 2e-4 * 0.6],
                                 TN * 2e-4 * 0.2)
    elif r < 0.7:
        nb = int(rng.integers(1, 5))
        job["bval"] = rng.uniform(0.1, 3, nb)
        job["bvec"] = fib_dirs(nb)
    chunk = int(rng.choice([0, 0, int(rng.integers(7, 200))]))
    return lab, D, T2, kappa, periodic, prec, TN, T, job, chunk


def test_random_inputs_bitexact(rms, orc, engine):
    rng = np.random.default_rng(20261204)
    for k in range(300):
        lab, D, T2, kappa, periodic, prec, TN, T, job, chunk = _case(rng)
        desc = (f"case {k}: shape={lab.shape} D={np.round(D, 2).tolist()} kappa={kappa} periodic='{periodic}' "
                f"{prec} nwalk={job['nwalk']} TN={TN} T={T} chunk={chunk} tables={sorted(job)}")
        print(desc, flush=True)
        gg = rms.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic)
        og = orc.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic)
        a = engine.run([dict(geom=gg, **job)], 2e-4, TN, T, precision=prec, replicas=4, final_state=True,
                       chunk_steps=chunk)[0]
        b = orc.run([dict(geom=og, **job)], 2e-4, TN, T, prec=prec, replicas=4, final_state=True)[0]
        for key in ("counters", "fin_g", "fin_gi", "fin_c", "fin_cnt", "fin_o", "fin_oi", "rec", "rec_rep"):
            np.testing.assert_array_equal(np.asarray(a[key]), np.asarray(b[key]), err_msg=f"{key}: {desc}")
        if b["sig"] is not None:
            np.testing.assert_array_equal(a["sig"], b["sig"], err_msg=desc)
            np.testing.assert_array_equal(a["sig_rep"], b["sig_rep"], err_msg=desc)


def test_random_inputs_adjoint_bitexact(rms, orc, engine):
    """The adjoint pass on 100 seeded random inputs with random cotangents:
    gradients and their replica partials bit-identical to the oracle's (NaN
    D entries where exchange makes them undefined, on both sides)."""
    rng = np.random.default_rng(20261205)
    for k in range(100):
        lab, D, T2, kappa, periodic, prec, TN, T, job, _ = _case(rng)
        desc = (f"adjoint case {k}: shape={lab.shape} D={np.round(D, 2).tolist()} kappa={kappa} "
                f"periodic='{periodic}' {prec} nwalk={job['nwalk']} TN={TN} T={T} tables={sorted(job)}")
        print(desc, flush=True)
        gg = rms.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic)
        og = orc.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic)
        nb = len(job["bval"]) if "bval" in job else 0
        ng = job["grad"].shape[0] if "grad" in job else 0
        cot_rec = [rng.standard_normal((T, rms.nval(len(D), nb)))]
        cot_sig = [rng.standard_normal(ng)] if ng else None
        plan = engine.plan([dict(geom=gg, **job)], 2e-4, TN, T, precision=prec, replicas=4)
        try:
            plan.execute(final_state=False)
            a = plan.adjoint(cot_rec, cot_sig, want_replicas=True, estimator="boundary")[0]
        finally:
            plan.close()
        b = orc.run([dict(geom=og, **job)], 2e-4, TN, T, prec=prec, replicas=4, cot_rec=cot_rec, cot_sig=cot_sig)[0]
        np.testing.assert_array_equal(a["grad"], b["grad"], err_msg=desc)
        np.testing.assert_array_equal(a["grad_rep"], b["grad_rep"], err_msg=desc)


def test_random_multi_job_plans_bitexact(rms, orc, engine):
    """60 seeded random plans of 2-5 random jobs each (mixed classes, tables,
    periodic axes and lattice sizes in one launch, random launch shapes and
    chunks, with or without the diagnostic counters): every job bit-exact
    against the oracle and its records equal to the job run alone."""
    rng = np.random.default_rng(20261206)
    for k in range(60):
        nj = int(rng.integers(2, 6))
        prec = "f32" if rng.random() < 0.6 else "f64"
        TN = int(rng.integers(20, 400))
        T = int(rng.integers(1, min(TN, 40) + 1))
        chunk = int(rng.choice([0, int(rng.integers(7, 200))]))
        shape = dict(block_threads=int(rng.choice([0, 64, 128, 192, 256])),
                     grid_blocks=int(rng.choice([0, int(rng.integers(1, 3000))])), xcd_split=int(rng.choice([-1, 0, 1])))
        gj, oj = [], []
        for _ in range(nj):
            lab, D, T2, kappa, periodic, _, _, _, job, _ = _case(rng)
            gj.append(dict(geom=rms.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic), **job))
            oj.append(dict(geom=orc.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic), **job))
        counted = bool(rng.random() < 0.5)
        desc = (f"plan {k}: {nj} jobs {prec} TN={TN} T={T} chunk={chunk} launch={shape} counters={counted} "
                f"shapes={[j['geom'].labels.shape for j in gj]} tables={[sorted(j) for j in gj]}")
        print(desc, flush=True)
        a = engine.run(gj, 2e-4, TN, T, precision=prec, replicas=4, final_state=counted, chunk_steps=chunk,
                       counters=counted, **shape)
        b = orc.run(oj, 2e-4, TN, T, prec=prec, replicas=4, final_state=True)
        keys = ("counters", "fin_g", "fin_gi", "fin_c", "fin_o", "fin_oi", "rec", "rec_rep") if counted else \
            ("rec", "rec_rep")
        for j in range(nj):
            for key in keys:
                np.testing.assert_array_equal(np.asarray(a[j][key]), np.asarray(b[j][key]),
                                              err_msg=f"job {j} {key}: {desc}")
            if not counted:
                assert np.all(a[j]["counters"] == 0), desc
            if b[j]["sig"] is not None:
                np.testing.assert_array_equal(a[j]["sig"], b[j]["sig"], err_msg=f"job {j}: {desc}")
        j = int(rng.integers(0, nj))
        alone = engine.run([gj[j]], 2e-4, TN, T, precision=prec, replicas=4)[0]
        np.testing.assert_array_equal(alone["rec"], a[j]["rec"], err_msg=f"alone job {j}: {desc}")


def test_random_plans_subset_differences_bitexact(rms, orc, engine):
    """100 seeded random plans of 1-4 random jobs (random launch shapes and
    chunks), the central differences of D for every job (estimator "fd") on
    the walkers the tangent pass marks: every D entry, total and per
    replica, bit-identical to the oracle's differences on its own marks.
    Entries the engine leaves NaN (a perturbation that needs another kernel
    variant) are skipped; most are compared."""
    rng = np.random.default_rng(20261207)
    compared = skipped = 0
    for k in range(100):
        nj = int(rng.integers(1, 5))
        prec = "f32" if rng.random() < 0.6 else "f64"
        TN = int(rng.integers(20, 400))
        T = int(rng.integers(1, min(TN, 40) + 1))
        chunk = int(rng.choice([0, int(rng.integers(7, 200))]))
        shape = dict(block_threads=int(rng.choice([0, 64, 128, 256])),
                     grid_blocks=int(rng.choice([0, int(rng.integers(1, 3000))])), xcd_split=int(rng.choice([-1, 0, 1])))
        gj, oj, cot_rec, cot_sig = [], [], [], []
        for _ in range(nj):
            lab, D, T2, kappa, periodic, _, _, _, job, _ = _case(rng)
            gj.append(dict(geom=rms.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic), **job))
            oj.append(dict(geom=orc.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic), **job))
        nbs = [len(j["bval"]) if "bval" in j else 0 for j in gj]
        ngs = [j["grad"].shape[0] if "grad" in j else 0 for j in gj]
        cot_rec = [rng.standard_normal((T, rms.nval(len(j["geom"].D), n))) for j, n in zip(gj, nbs)]
        cot_sig = [rng.standard_normal(ng) for ng in ngs] if any(ngs) else None
        desc = (f"subset plan {k}: {nj} jobs {prec} TN={TN} T={T} chunk={chunk} launch={shape} "
                f"shapes={[j['geom'].labels.shape for j in gj]} D={[np.round(j['geom'].D, 2).tolist() for j in gj]}")
        print(desc, flush=True)
        plan = engine.plan(gj, 2e-4, TN, T, precision=prec, replicas=4, chunk_steps=chunk, **shape)
        try:
            a = plan.adjoint(cot_rec, cot_sig, want_replicas=True, estimator="fd")
        finally:
            plan.close()
        g, gr = orc.central_differences(oj, 2e-4, TN, T, cot_rec, cot_sig, h=0.05, prec=prec, replicas=4,
                                        subsets=True)
        for j in range(nj):
            for c in range(1, len(gj[j]["geom"].D)):
                if np.isnan(a[j]["grad"][c]):
                    skipped += 1
                    continue
                compared += 1
                assert a[j]["grad"][c] == g[j, c], f"job {j} D_{c}: {a[j]['grad'][c]} vs {g[j, c]}: {desc}"
                np.testing.assert_array_equal(a[j]["grad_rep"][:, c], gr[j, :, c], err_msg=f"job {j} D_{c}: {desc}")
    print(f"compared {compared} entries, skipped {skipped}", flush=True)
    assert compared > 5 * skipped
