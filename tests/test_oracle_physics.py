"""Statistical pins of the oracle against closed forms (free diffusion,
impermeable planes, a voxelised cylinder, T2 decay, detailed balance,
permeable membranes, PGSE and narrow-pulse signals) and its invariants.
Tolerances are z-scores against batch-means standard errors."""
import math

import numpy as np
import pytest

from helpers import cylinder, fib_dirs, pgse_table, planes, random_blobs

DT = 2e-4


def _D_axes(r, TD):
    dx2 = r["dx2"] / r["sig0"][:, None]
    return np.stack([dx2[:, 0], dx2[:, 3], dx2[:, 5]], 1) / (2 * TD[:, None])


def _rep_D_axes(r, TD, nclass):
    """per-replica D per axis: (R, T, 3)"""
    rep = r["rec_rep"]
    s0 = rep[:, :, 0]
    dx2 = rep[:, :, 1 + nclass:7 + nclass]
    return np.stack([dx2[..., 0], dx2[..., 3], dx2[..., 5]], -1) / s0[..., None] / (2 * TD[None, :, None])


def _se(rep_vals):
    R = rep_vals.shape[0]
    return rep_vals.std(axis=0, ddof=1) / math.sqrt(R)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_free_diffusion(orc, prec):
    lab = np.ones((16, 16, 16), np.uint8)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    TN, T = 2000, 20
    r = orc.run([dict(geom=g, nwalk=4096, seed=7)], DT, TN, T, prec=prec)[0]
    TD = orc.diff_times(DT, TN, T)
    D = _D_axes(r, TD)
    se = _se(_rep_D_axes(r, TD, 2))
    z = (D - 2.0) / se
    assert np.max(np.abs(z[-5:])) < 4.5, z[-5:]
    # kurtosis of a sum of N isotropic steps: -1.2/N per axis
    dx2 = r["dx2"] / r["sig0"][:, None]
    dx4 = r["dx4"] / r["sig0"][:, None]
    K = dx4[-1, 0] / dx2[-1, 0] ** 2 - 3
    assert abs(K) < 0.15
    # invariants
    assert np.all(r["NPar_count"].sum(1) == 4096) and np.all(r["NPar_count"][:, 0] == 0)
    assert r["counters"][3] == 0
    assert np.all(r["sig0"] == 4096.0)


def _planes_x2(a, D0, t, nterm=400):
    n = np.arange(1, 2 * nterm, 2)[None, :]
    t = np.asarray(t)[:, None]
    return a * a / 6 - 16 * a * a / np.pi ** 4 * np.sum(n ** -4.0 * np.exp(-(n ** 2) * np.pi ** 2 * D0 * t / (a * a)), 1)


def test_impermeable_planes_series(orc):
    nz = 12  # label-0 slabs at z=0 and z=11: gap a = 1.0 um
    lab = planes(8, 8, nz)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    TN, T = 2000, 40
    r = orc.run([dict(geom=g, nwalk=8192, seed=3)], DT, TN, T, prec="f32")[0]
    TD = orc.diff_times(DT, TN, T)
    a = (nz - 2) * 0.1
    dz2 = r["dx2"][:, 5] / r["sig0"]
    rep = r["rec_rep"]
    dz2_rep = rep[:, :, 1 + 2 + 5] / rep[:, :, 0]
    se = _se(dz2_rep)
    want = _planes_x2(a, 2.0, TD)
    z = (dz2 - want) / np.maximum(se, 1e-12)
    assert np.max(np.abs(z[5:])) < 4.5, z
    # parallel diffusion is free (no bias for elastic collisions)
    D = _D_axes(r, TD)
    sexy = _se(_rep_D_axes(r, TD, 2))
    assert np.max(np.abs((D[-10:, :2] - 2.0) / sexy[-10:, :2])) < 4.5
    # long-time kurtosis: <dz^4> = a^4/15, <dz^2> = a^2/6  ->  K = -0.6
    dz4 = r["dx4"][:, 14] / r["sig0"]
    K = dz4[-1] / dz2[-1] ** 2 - 3
    assert abs(K + 0.6) < 0.08


def test_cylinder_long_time_variance(orc):
    lab = cylinder(20, 6, 7.0)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    TN, T = 3000, 30
    r = orc.run([dict(geom=g, nwalk=4096, seed=5)], DT, TN, T, prec="f32")[0]
    # 2 Var(x) over the voxel set (uniform density in each voxel)
    idx = np.argwhere(lab[:, :, 0] == 1) + 0.5
    var = (idx[:, 0] ** 2).mean() - idx[:, 0].mean() ** 2 + 1 / 12
    want = 2 * var * 0.01
    dxx = r["dx2"][:, 0] / r["sig0"]
    rep = r["rec_rep"]
    se = _se(rep[:, :, 3] / rep[:, :, 0])
    assert abs(dxx[-1] - want) < 4.5 * se[-1] + 0.01 * want


def test_t2_weight_closed_form(orc):
    lab = planes(6, 6, 8)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0], T2=[1.0, 30.0])
    TN, T = 500, 50
    r = orc.run([dict(geom=g, nwalk=300, seed=9)], DT, TN, T, prec="f64")[0]
    steps = np.arange(T) * (TN // T) + 1
    want = 300 * np.exp(-steps * DT / 30.0)
    np.testing.assert_allclose(r["sig0"], want, rtol=1e-12)
    r32 = orc.run([dict(geom=g, nwalk=300, seed=9)], DT, TN, T, prec="f32")[0]
    np.testing.assert_allclose(r32["sig0"], want, rtol=1e-6)


def test_detailed_balance_volume_fractions(orc):
    lab = random_blobs((12, 12, 12), 3, seed=11)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0, 0.5])
    P = orc.derive(g, DT)["P"]
    assert 0 < P[1, 2] < 1  # probabilistic exchange
    r = orc.run([dict(geom=g, nwalk=6000, seed=2)], DT, 1500, 15, prec="f32")[0]
    vol = np.array([(lab == c).sum() for c in (1, 2)], float)
    frac = vol / vol.sum()
    got = r["NPar_count"][:, 1:] / 6000.0
    se = np.sqrt(frac * (1 - frac) / 6000.0)
    assert np.max(np.abs((got - frac) / se)) < 4.5
    assert r["counters"][4] > 0 and np.all(r["NPar_count"][:, 0] == 0)


def test_permeable_membranes_effective_diffusivity(orc):
    """Slabs of width a alternating between two classes with permeability
    kappa: long-time D_perp = 1/(1/D0 + 1/(kappa a)), D_parallel = D0.  Half
    slabs at both lattice ends keep the mirror images periodic."""
    w, nper = 6, 8
    nz = 2 * w * nper
    z = np.arange(nz)
    cls = np.where(((z + w // 2) // w) % 2 == 0, 1, 2).astype(np.uint8)
    lab = np.ascontiguousarray(np.broadcast_to(cls, (4, 4, nz)))
    a, D0, kap = w * 0.1, 2.0, 1.0
    g = orc.Geometry(lab, 0.1, [0.0, D0, D0], kappa=kap)
    TN, T = 12000, 24
    r = orc.run([dict(geom=g, nwalk=8192, seed=5)], DT, TN, T, prec="f32")[0]
    TD = orc.diff_times(DT, TN, T)
    dz2 = r["dx2"][:, 5] / r["sig0"]
    slope = np.polyfit(TD[T // 2:], dz2[T // 2:], 1)[0] / 2
    Deff = 1 / (1 / D0 + 1 / (kap * a))
    assert abs(slope / Deff - 1) < 0.06, (slope, Deff)
    D = _D_axes(r, TD)
    assert abs(D[-1, :2].mean() / D0 - 1) < 0.05
    # kappa = 0 makes the interfaces impermeable: nobody changes class
    g0 = orc.Geometry(lab, 0.1, [0.0, D0, D0], kappa=0.0)
    r0 = orc.run([dict(geom=g0, nwalk=512, seed=5)], DT, 400, 4, prec="f32")[0]
    assert r0["counters"][4] == 0
    np.testing.assert_array_equal(r0["NPar_count"][0], r0["NPar_count"][-1])


def test_pgse_window_sums_equal_literal_loop(orc):
    lab = cylinder(12, 6, 4.5)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0], T2=[70.0, 70.0])
    grad = pgse_table(fib_dirs(5), [0.3, 1.5], [0.05, 0.07], 0.02)
    grad = np.concatenate([grad, [[0.03, 0.05, 2.0, 0, 0, 1.0]]])  # Delta < delta row
    kw = dict(dt=DT, TN=600, timepoints=60)
    jobs = [dict(geom=g, nwalk=200, seed=8, grad=grad)]
    ws = orc.run(jobs, prec="f64", **kw)[0]
    bf = orc.run(jobs, prec="f64", pgse_bruteforce=True, **kw)[0]
    np.testing.assert_allclose(ws["sig"], bf["sig"], rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_pgse_split_sums_long_run_equal_literal_loop(orc, prec):
    """The window sums keep sum(g) in int32 and flush every 2^16 steps; over
    80000 steps the flush falls inside the second lobe, and the z lattice
    edges mirror the origin between flushes.  Equal to the literal per-step
    phase loop of main_PGSE.cu:340-346."""
    lab = cylinder(12, 6, 4.5)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0], T2=[70.0, 70.0])
    grad = pgse_table(fib_dirs(3), [0.5, 2.0], [10.0], 5.0)
    kw = dict(dt=DT, TN=80000, timepoints=8)
    jobs = [dict(geom=g, nwalk=64, seed=18, grad=grad)]
    ws = orc.run(jobs, prec=prec, **kw)[0]
    bf = orc.run(jobs, prec=prec, pgse_bruteforce=True, **kw)[0]
    assert ws["counters"][2] > 1000  # mirrors
    np.testing.assert_allclose(ws["sig"], bf["sig"], rtol=1e-10, atol=1e-9)


def test_pgse_free_diffusion_gaussian_phase(orc):
    """Gradients along z in a lattice 40 um long in z: no walker reaches a z
    edge, whose mirror would flip the sign of later displacements
    (main.cu:302-310) and scramble the phase."""
    lab = np.ones((4, 4, 400), np.uint8)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    TN = 400
    delta, Delta = 0.02, 0.05
    grad = pgse_table([[0, 0, 1.0]], [0.5, 1.0, 3.0], [Delta], delta)
    r = orc.run([dict(geom=g, nwalk=20000, seed=13, grad=grad)], DT, TN, 40, prec="f32")[0]
    S = r["sig"] / 20000.0
    # exact discrete variance of the phase: g^2 dt^2 * 2 D dt * sum_k W_k^2
    abe = orc.pgse_lobes(DT, TN, grad)
    for j, (a, b, e) in enumerate(abe):
        w = np.zeros(TN)
        w[:a] = 1
        w[b:e] = -1
        W = np.cumsum(w[::-1])[::-1]
        var = grad[j, 2] ** 2 * DT ** 2 * 2 * 2.0 * DT * np.sum(W ** 2)
        want = math.exp(-var / 2)
        se = np.std(r["sig_rep"][:, j] / (20000.0 / 32), ddof=1) / math.sqrt(32)
        assert abs(S[j] - want) < 4.5 * se + 0.01 * want, (j, S[j], want, se)


def test_narrow_pulse_free_diffusion(orc):
    lab = np.ones((12, 12, 12), np.uint8)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    bval = np.array([0.5, 1.0, 2.0])
    bvec = fib_dirs(3)
    r = orc.run([dict(geom=g, nwalk=20000, seed=17, bval=bval, bvec=bvec)], DT, 500, 10, prec="f32")[0]
    S = r["sigRe"] / r["sig0"][:, None]
    want = np.exp(-bval * 2.0)[None, :]
    rep = r["rec_rep"]
    Srep = rep[:, :, 24:27] / rep[:, :, :1]
    se = _se(Srep)
    assert np.max(np.abs((S[-3:] - want) / (se[-3:] + 1e-3 * want))) < 4.5


def test_reduction_order_is_thread_count_independent(orc):
    lab = random_blobs((8, 8, 8), 3, seed=1)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0, 1.0])
    jobs = [dict(geom=g, nwalk=1000, seed=3, bval=[1.0], bvec=[[0, 0, 1.0]])]
    a = orc.run(jobs, DT, 200, 20, nthreads=1)[0]
    b = orc.run(jobs, DT, 200, 20, nthreads=5)[0]
    np.testing.assert_array_equal(a["rec"], b["rec"])
    np.testing.assert_array_equal(a["rec_rep"], b["rec_rep"])
    # totals are the in-order sum of the replicas
    tot = np.zeros_like(a["rec"])
    for rr in a["rec_rep"]:
        tot = tot + rr
    np.testing.assert_array_equal(tot, a["rec"])


def test_f32_and_f64_agree_statistically(orc):
    lab = cylinder(16, 8, 6.0)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    TN, T = 1500, 15
    TD = orc.diff_times(DT, TN, T)
    out = {}
    for prec in ("f32", "f64"):
        r = orc.run([dict(geom=g, nwalk=6000, seed=19 if prec == "f32" else 23)], DT, TN, T, prec=prec)[0]
        out[prec] = (_D_axes(r, TD), _se(_rep_D_axes(r, TD, 2)))
    d = out["f32"][0] - out["f64"][0]
    se = np.hypot(out["f32"][1], out["f64"][1])
    assert np.max(np.abs(d / se)) < 4.5


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_initial_positions_uniform_over_walkable_voxels(orc, prec):
    """main.cu:117-132: start points are uniform over the non-null voxels.
    A border of null voxels keeps walkers off the lattice edges, so after one
    step the origin (gi, oi) is still the start point: every start voxel is
    walkable, the voxel counts pass a chi-square test against the uniform
    distribution, and the in-voxel offsets are uniform on [0, 1)."""
    from scipy import stats
    rng = np.random.default_rng(21)
    lab = rng.integers(0, 3, size=(12, 10, 8)).astype(np.uint8)
    lab[[0, -1], :, :] = 0
    lab[:, [0, -1], :] = 0
    lab[:, :, [0, -1]] = 0
    g = orc.Geometry(lab, 0.1, [0.0, 2.0, 0.5])
    n = 40000
    r = orc.run([dict(geom=g, nwalk=n, seed=4242)], DT, 1, 1, prec=prec, final_state=True)[0]
    gi, oi = r["fin_gi"], r["fin_oi"]
    assert np.all(lab[gi[0], gi[1], gi[2]] != 0)
    walk = np.flatnonzero(lab.ravel() != 0)
    lin = np.ravel_multi_index((gi[0], gi[1], gi[2]), lab.shape)
    counts = np.bincount(lin, minlength=lab.size)[walk]
    assert stats.chisquare(counts).pvalue > 1e-3
    assert np.all((oi >= 0) & (oi < 1))
    for k in range(3):
        assert stats.kstest(oi[k], "uniform").pvalue > 1e-3


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_periodic_wall_series(orc, prec):
    """A single label-0 plane at z = 0 with z periodic: the walkers live
    between the plane and its periodic image, a gap a = (nz - 1) voxels, so
    <dz^2>(t) follows the impermeable-plane series for that gap.  With the
    reference's mirror edge the same lattice is a gap of 2 (nz - 1) voxels
    (checked too).  x and y are periodic and free: D = D0 along them."""
    nz = 12
    lab = np.ones((6, 6, nz), np.uint8)
    lab[:, :, 0] = 0
    TN, T = 3000, 30
    TD = orc.diff_times(DT, TN, T)
    for periodic, a in (("xyz", (nz - 1) * 0.1), ("", 2 * (nz - 1) * 0.1)):
        g = orc.Geometry(lab, 0.1, [0.0, 2.0], periodic=periodic)
        r = orc.run([dict(geom=g, nwalk=8192, seed=11)], DT, TN, T, prec=prec)[0]
        dz2 = r["dx2"][:, 5] / r["sig0"]
        rep = r["rec_rep"]
        se = _se(rep[:, :, 1 + 2 + 5] / rep[:, :, 0])
        z = (dz2 - _planes_x2(a, 2.0, TD)) / np.maximum(se, 1e-12)
        assert np.max(np.abs(z[5:])) < 4.5, (periodic, z)
        D = _D_axes(r, TD)
        sexy = _se(_rep_D_axes(r, TD, 2))
        assert np.max(np.abs((D[-10:, :2] - 2.0) / sexy[-10:, :2])) < 4.5
        assert np.all(r["NPar_count"][:, 0] == 0)
        if periodic:
            assert r["counters"][2] == 0  # no mirrors: every edge wraps


def test_periodic_free_diffusion_unwrapped(orc):
    """All-1 lattice, periodic in x, y, z: the displacement keeps growing past
    the lattice size (the origin moves with each wrap), <x^2> = 2 D0 t."""
    lab = np.ones((4, 4, 4), np.uint8)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0], periodic="xyz")
    TN, T = 3000, 30
    r = orc.run([dict(geom=g, nwalk=4096, seed=12)], DT, TN, T, prec="f32")[0]
    TD = orc.diff_times(DT, TN, T)
    D = _D_axes(r, TD)
    se = _se(_rep_D_axes(r, TD, 2))
    assert np.max(np.abs((D[-5:] - 2.0) / se[-5:])) < 4.5
    assert 2 * 2.0 * TD[-1] > (4 * 0.1) ** 2 * 4  # displacements well beyond the 0.4 um lattice
    assert r["counters"][2] == 0


def test_start_failure_is_an_error(orc):
    """A walker's start is rejection sampling with a cap of 2^26 trials
    (main.cu:117-132 has no cap and would loop for ever).  One walkable voxel
    in 1e8 gives a failure probability exp(-2^26 / 1e8) ~ 0.51 per walker:
    seed 2 exhausts the trials and the run is refused instead of reporting a
    walker parked in voxel (0, 0, 0); seed 1 finds the voxel."""
    lab = np.zeros((500, 500, 400), np.uint8)
    lab[123, 321, 77] = 1
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    r = orc.run([dict(geom=g, nwalk=1, seed=1)], 2e-4, 10, 10, prec="f32", replicas=1, final_state=True)[0]
    assert r["fin_g"].ravel().tolist() == [123, 321, 77]
    with pytest.raises(RuntimeError, match="no non-null voxel"):
        orc.run([dict(geom=g, nwalk=1, seed=2)], 2e-4, 10, 10, prec="f32", replicas=1)
