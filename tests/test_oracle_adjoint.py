"""The oracle's adjoint pass against finite differences of its own forward
pass (common random numbers: the same seed gives the same direction and
exchange draws, so F(theta + h) - F(theta - h) isolates the parameter).

Pins, per estimator:
  T2      analytic weight derivative; paths do not depend on T2, so the
          central difference matches to rounding.
  D       pathwise derivative (faces fixed, step length ~ sqrt(D)) on a
          homogeneous lattice, between impermeable planes (reflections are
          continuous in D) and in impermeable compartments, records and PGSE
          signal; free diffusion dD(t)/dD0 = 1 and dK(t)/dD0 = 0.  With
          exchange, plus the score of the draws and the jump of the class
          change at an exchange face (statistical check against central
          differences on permeable slabs); NaN with a PGSE cotangent.
  kappa   score function of the exchange draws (the forward map is
          piecewise constant in P, so the check is statistical: batch-means
          error bars on both sides, large-step central differences).
The GPU engine reproduces these gradients bit for bit (test_gpu_parity).
"""
import numpy as np
import pytest

from helpers import fib_dirs, pgse_table, planes, random_blobs

DT = 2e-4


def _F(res, cot, cots):
    v = float(np.sum(cot * res["rec"]))
    if cots is not None:
        v += float(np.sum(cots * res["sig"]))
    return v


def _job(orc, lab, D, T2, kappa=None, nwalk=512, seed=7, periodic="", **kw):
    g = orc.Geometry(lab, 0.1, D, T2=T2, kappa=kappa, periodic=periodic)
    return [dict(geom=g, nwalk=nwalk, seed=seed, **kw)]


def _scaled_cot(rng, T, nclass, nbvec):
    # comparable magnitudes for sig0, counts, dx2 (um^2) and dx4 (um^4)
    sc = np.r_[1.0, np.ones(nclass), np.full(6, 1e2), np.full(15, 1e4), np.ones(nbvec)]
    return rng.standard_normal((T, sc.size)) * sc


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_t2_gradient_matches_central_difference(orc, prec):
    lab = random_blobs((9, 9, 9), 3, seed=2)
    D, T2 = [0.0, 2.0, 1.0], [np.inf, 60.0, 25.0]
    bval, bvec = np.array([1.0, 2.5]), fib_dirs(2)
    grad = pgse_table(fib_dirs(2), [1.0], [0.03], 0.01)
    TN, T = 200, 10
    rng = np.random.default_rng(3)
    cot = _scaled_cot(rng, T, 3, 2)
    cots = rng.standard_normal(grad.shape[0])
    kw = dict(bval=bval, bvec=bvec, grad=grad)
    ra = orc.run(_job(orc, lab, D, T2, 0.2, **kw), DT, TN, T, prec=prec, cot_rec=[cot], cot_sig=[cots])[0]
    g = ra["grad"]
    assert g[4] == 0.0  # class 0 has T2 = inf
    for c in (1, 2):
        h = 1e-4 * T2[c]
        tp, tm = list(T2), list(T2)
        tp[c] += h
        tm[c] -= h
        fp = _F(orc.run(_job(orc, lab, D, tp, 0.2, **kw), DT, TN, T, prec=prec)[0], cot, cots)
        fm = _F(orc.run(_job(orc, lab, D, tm, 0.2, **kw), DT, TN, T, prec=prec)[0], cot, cots)
        fd = (fp - fm) / (2 * h)
        assert g[4 + c] == pytest.approx(fd, rel=1e-6), (c, g[4 + c], fd)
        assert abs(fd) > 0


@pytest.mark.parametrize("pgse", [True, False])
@pytest.mark.parametrize("geom", ["homogeneous", "planes"])
def test_d_pathwise_gradient_matches_central_difference(orc, geom, pgse):
    """Periodic lattices without voxel-edge order effects: every face a walk
    meets is crossed or reflects continuously in D, so the pathwise derivative
    is the whole derivative and equals a central difference with common random
    numbers to rounding (with and without a PGSE table; without one the
    boundary-event search runs and finds nothing)."""
    if geom == "homogeneous":
        lab = np.ones((12, 12, 12), np.uint8)
    else:
        lab = planes(6, 6, 9)  # walls at z = 0 and z = 8: many reflections
    D, T2 = [0.0, 1.7], [np.inf, 70.0]
    bval, bvec = np.array([2.0, 4.0]), fib_dirs(2)
    grad = pgse_table(fib_dirs(3), [1.0, 3.0], [0.03], 0.012) if pgse else None
    TN, T = 240, 12
    rng = np.random.default_rng(5)
    cot = _scaled_cot(rng, T, 2, 2)
    cots = rng.standard_normal(grad.shape[0]) if pgse else None
    kw = dict(bval=bval, bvec=bvec, grad=grad, periodic="xyz")
    ra = orc.run(_job(orc, lab, D, T2, **kw), DT, TN, T, prec="f64", cot_rec=[cot], cot_sig=[cots])[0]
    g = ra["grad"][1]
    h = 1e-6 * D[1]
    fp = _F(orc.run(_job(orc, lab, [0.0, D[1] + h], T2, **kw), DT, TN, T, prec="f64")[0], cot, cots)
    fm = _F(orc.run(_job(orc, lab, [0.0, D[1] - h], T2, **kw), DT, TN, T, prec="f64")[0], cot, cots)
    fd = (fp - fm) / (2 * h)
    assert g == pytest.approx(fd, rel=2e-4), (g, fd)
    if geom == "planes":
        r = orc.run(_job(orc, lab, D, T2, **kw), DT, TN, T, prec="f64")[0]
        assert r["counters"][1] > 1000  # the reflections were exercised


def test_d_gradient_nan_with_pgse_where_boundary_terms_exist(orc):
    """With a PGSE table the D entries are NaN wherever the walk's end jumps
    with D: lattice-edge mirrors (a homogeneous lattice with mirror edges) and
    voxel-edge order effects (a voxelised sphere, periodic edges)."""
    grad = pgse_table(fib_dirs(2), [1.0], [0.004], 0.002)
    cot = np.ones((5, orc.nval(2, 0)))
    cases = (("", np.ones((6, 6, 6), np.uint8)),
             ("xyz", (np.sum((np.indices((9, 9, 9)) - 4.0) ** 2, 0) <= 12.0).astype(np.uint8)))
    for per, lab in cases:
        r = orc.run(_job(orc, lab, [0.0, 2.0], [np.inf, 50.0], grad=grad, periodic=per), DT, 50, 5, prec="f64",
                    replicas=4, cot_rec=[cot], cot_sig=[np.ones(grad.shape[0])])[0]
        assert np.isnan(r["grad"][1]) and np.all(np.isnan(r["grad_rep"][:, 1])), per
        assert np.all(np.isfinite(r["grad"][4:6]))
        # without the PGSE table the boundary terms are there
        r = orc.run(_job(orc, lab, [0.0, 2.0], [np.inf, 50.0], periodic=per), DT, 50, 5, prec="f64", replicas=4,
                    cot_rec=[cot])[0]
        assert np.isfinite(r["grad"][1])


def test_d_gradient_impermeable_compartments(orc):
    """Two non-null classes in periodic slabs separated by impermeable
    membranes (kappa = 0): no walker changes class and no step's end jumps
    with D, so the pathwise D derivative is exact for both."""
    lab = _slabs(w=4, nper=2, nx=5)
    D, T2 = [0.0, 2.0, 0.8], [np.inf, 70.0, 35.0]
    TN, T = 200, 10
    rng = np.random.default_rng(9)
    cot = _scaled_cot(rng, T, 3, 0)
    kw = dict(periodic="xyz")
    ra = orc.run(_job(orc, lab, D, T2, 0.0, **kw), DT, TN, T, prec="f64", cot_rec=[cot])[0]
    for c in (1, 2):
        h = 1e-6 * D[c]
        dp, dm = list(D), list(D)
        dp[c] += h
        dm[c] -= h
        fd = (_F(orc.run(_job(orc, lab, dp, T2, 0.0, **kw), DT, TN, T, prec="f64")[0], cot, None)
              - _F(orc.run(_job(orc, lab, dm, T2, 0.0, **kw), DT, TN, T, prec="f64")[0], cot, None)) / (2 * h)
        assert ra["grad"][c] == pytest.approx(fd, rel=2e-4), (c, ra["grad"][c], fd)


def test_free_diffusion_dD_dD0_is_one_and_dK_dD0_is_zero(orc):
    """Free diffusion: the apparent D(t) = <dz^2>/(2t) has dD/dD0 = 1 and the
    kurtosis K(t) = <dz^4>/<dz^2>^2 - 3 has dK/dD0 = 0 (scale invariance of
    a path whose step length is sqrt(6 D0 dt)).  The lattice edges are
    periodic, so the walk is the free walk unwrapped."""
    n = 48
    lab = np.ones((n, n, n), np.uint8)
    D0 = 1.5
    TN, T, N = 200, 4, 4096
    TD = orc.diff_times(DT, TN, T)
    res = orc.run(_job(orc, lab, [0.0, D0], [np.inf, np.inf], nwalk=N, periodic="xyz"), DT, TN, T, prec="f64")[0]
    rec = res["rec"]
    iz2, iz4 = 1 + 2 + 5, 1 + 2 + 6 + 14  # dx2 zz, dx4 zzzz columns
    # D_est = sum_t dz2_t / (N 2 TD_t): cotangent on the dz2 column
    cot = np.zeros_like(rec)
    cot[:, iz2] = 1.0 / (N * 2.0 * TD)
    gD = orc.run(_job(orc, lab, [0.0, D0], [np.inf, np.inf], nwalk=N, periodic="xyz"), DT, TN, T, prec="f64",
                 cot_rec=[cot])[0]["grad"][1]
    Dest = rec[:, iz2] / (N * 2.0 * TD)
    # each path is the unit path scaled by sqrt(D0): dD_est/dD0 = D_est/D0 exactly
    assert gD == pytest.approx(Dest.sum() / D0, rel=1e-9)
    assert gD / T == pytest.approx(1.0, abs=0.05)  # and D_est estimates D0
    # K_t = (dz4/N) / (dz2/N)^2 - 3, summed over t: chain rule for the cotangent
    m2, m4 = rec[:, iz2] / N, rec[:, iz4] / N
    cotK = np.zeros_like(rec)
    cotK[:, iz4] = 1.0 / (N * m2 ** 2)
    cotK[:, iz2] = -2.0 * m4 / (N * m2 ** 3)
    gK = orc.run(_job(orc, lab, [0.0, D0], [np.inf, np.inf], nwalk=N, periodic="xyz"), DT, TN, T, prec="f64",
                 cot_rec=[cotK])[0]["grad"][1]
    assert abs(gK) < 1e-9 * T / D0


def test_d_gradient_with_exchange_nan_only_with_pgse_tables(orc):
    """With exchange the D entries carry the jump term of the class change;
    the continuations that term needs carry no PGSE prefix sums, so in a run
    with a PGSE table the D entries stay NaN."""
    lab = random_blobs((8, 8, 8), 3, seed=4)
    cot = np.ones((5, orc.nval(3, 0)))
    grad = pgse_table(fib_dirs(2), [1.0], [0.004], 0.002)
    for kappa in (0.3, None):  # permeable membranes, or the sqrt(Dj/Di) rule
        r = orc.run(_job(orc, lab, [0.0, 2.0, 1.0], [np.inf, 50.0, 30.0], kappa), DT, 50, 5, prec="f64",
                    replicas=4, cot_rec=[cot])[0]
        assert np.all(np.isfinite(r["grad"][1:3])) and np.all(np.isfinite(r["grad"][4:9]))
        for cs in (np.ones(grad.shape[0]), None):
            r = orc.run(_job(orc, lab, [0.0, 2.0, 1.0], [np.inf, 50.0, 30.0], kappa, grad=grad), DT, 50, 5,
                        prec="f64", replicas=4, cot_rec=[cot], cot_sig=[cs])[0]
            assert np.all(np.isnan(r["grad"][1:4])) and np.all(np.isnan(r["grad_rep"][:, 1:4]))
            assert np.all(np.isfinite(r["grad"][4:9]))


def _slabs(w=5, nper=3, nx=4):
    """Alternating slabs of classes 1 and 2, w voxels thick, along z."""
    z = np.arange(2 * w * nper)
    cls = np.where(((z + w // 2) // w) % 2 == 0, 1, 2).astype(np.uint8)
    return np.ascontiguousarray(np.broadcast_to(cls, (nx, nx, z.size)))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_d_gradient_with_exchange_matches_central_difference(orc, prec):
    """Permeable slabs (kappa = 0.5): pathwise + score of the draws + the jump
    of the class change at an exchange face (smoothed perturbation analysis:
    boundary weight times the difference of the walks continued on either
    side), against central differences with common random numbers.  Without
    the jump term the D_2 estimate falls about 5 standard errors short of the
    differences at this size (and 10 at 4x the walkers)."""
    lab = _slabs()
    D, T2, kap = [0.0, 2.0, 1.2], [np.inf, 80.0, 40.0], 0.5
    TN, T, R, N = 800, 8, 32, 1 << 16
    cot = np.zeros((T, orc.nval(3, 0)))
    cot[:, 0] = 0.3          # sig0
    cot[:, 2] = 1.0          # walkers in class 1
    cot[:, 1 + 3 + 0] = 30.0  # dx^2
    cot[:, 1 + 3 + 5] = 50.0  # dz^2
    ra = orc.run(_job(orc, lab, D, T2, kap, nwalk=N, seed=11), DT, TN, T, prec=prec, replicas=R, cot_rec=[cot])[0]

    def L(d):
        r = orc.run(_job(orc, lab, d, T2, kap, nwalk=N, seed=11), DT, TN, T, prec=prec, replicas=R)[0]
        return (cot[None] * r["rec_rep"]).reshape(R, -1).sum(1)

    for c in (1, 2):
        adj, adj_se = _batch_stats(ra["grad_rep"][:, c])
        h = 0.1 * D[c]
        dp, dm = list(D), list(D)
        dp[c] += h
        dm[c] -= h
        fd, fd_se = _batch_stats((L(dp) - L(dm)) / (2 * h))
        assert abs(fd) > 10 * fd_se
        # 3 % for the h^2 curvature of a 10 % central difference
        assert abs(adj - fd) < 4 * np.hypot(adj_se, fd_se) + 0.03 * abs(fd), (c, adj, adj_se, fd, fd_se)


def _batch_stats(x):
    R = x.shape[0]
    return x.sum(), x.std(ddof=1) * np.sqrt(R)


def test_kappa_score_gradient_matches_finite_difference(orc):
    """Membrane exchange (P between 0 and 1): the score-function gradient and a
    central difference of the expectation agree within their error bars."""
    w, nper = 5, 3
    z = np.arange(2 * w * nper)
    cls = np.where(((z + w // 2) // w) % 2 == 0, 1, 2).astype(np.uint8)
    lab = np.ascontiguousarray(np.broadcast_to(cls, (4, 4, z.size)))
    D, T2, kap = [0.0, 2.0, 1.2], [np.inf, 80.0, 40.0], 0.5
    TN, T, R, N = 800, 8, 32, 1 << 13
    cot = np.zeros((T, orc.nval(3, 0)))
    cot[:, 2] = 1.0          # walkers in class 1
    cot[:, 1 + 3 + 5] = 50.0  # sum of dz^2 (um^2)
    ra = orc.run(_job(orc, lab, D, T2, kap, nwalk=N, seed=11), DT, TN, T, prec="f64", replicas=R,
                 cot_rec=[cot])[0]
    adj, adj_se = _batch_stats(ra["grad_rep"][:, 8])

    def L(d, k):
        r = orc.run(_job(orc, lab, d, T2, k, nwalk=N, seed=11), DT, TN, T, prec="f64", replicas=R)[0]
        return (cot[None] * r["rec_rep"]).reshape(R, -1).sum(1)

    h = 0.2 * kap
    fdr = (L(D, kap + h) - L(D, kap - h)) / (2 * h)
    fd, fd_se = _batch_stats(fdr)
    assert abs(adj) > 4 * adj_se  # a real signal, not noise
    assert abs(adj - fd) < 4 * np.hypot(adj_se, fd_se) + 0.03 * abs(fd), (adj, adj_se, fd, fd_se)


def test_kappa_gradient_is_nan_where_kappa_closes_the_membrane(orc):
    """At kappa = 0 (and for a kappa so small that P < 1e-6) the membrane
    reflects without a draw, so the score function has nothing to sum; yet the
    expectation moves with kappa (a forward difference from 0 is clearly
    nonzero), so a zero would be wrong: the kappa entry is NaN there."""
    w, nper = 5, 3
    z = np.arange(2 * w * nper)
    cls = np.where(((z + w // 2) // w) % 2 == 0, 1, 2).astype(np.uint8)
    lab = np.ascontiguousarray(np.broadcast_to(cls, (4, 4, z.size)))
    D, T2 = [0.0, 2.0, 1.2], [np.inf, 80.0, 40.0]
    TN, T, R, N = 800, 8, 32, 1 << 12
    cot = np.zeros((T, orc.nval(3, 0)))
    cot[:, 1 + 3 + 5] = 50.0  # sum of dz^2 (um^2): walkers leave their slab only through the membranes
    for k0 in (0.0, 1e-9):
        r0 = orc.run(_job(orc, lab, D, T2, k0, nwalk=N, seed=11), DT, TN, T, prec="f64", replicas=R,
                     cot_rec=[cot])[0]
        assert np.isnan(r0["grad"][8]) and np.all(np.isnan(r0["grad_rep"][:, 8]))
        assert np.all(np.isfinite(r0["grad"][4:8]))
    # finite for a membrane that draws
    r1 = orc.run(_job(orc, lab, D, T2, 0.5, nwalk=N, seed=11), DT, TN, T, prec="f64", replicas=R, cot_rec=[cot])[0]
    assert np.isfinite(r1["grad"][8])

    def L(k):
        r = orc.run(_job(orc, lab, D, T2, k, nwalk=N, seed=11), DT, TN, T, prec="f64", replicas=R)[0]
        return (cot[None] * r["rec_rep"]).reshape(R, -1).sum(1)

    h = 0.1
    fd, fd_se = _batch_stats((L(h) - L(0.0)) / h)
    assert abs(fd) > 5 * fd_se, (fd, fd_se)


def test_akc_cotangent_matches_numeric_derivative(rms):
    """analysis.akc_cotangent (the chain rule from D(t), K(t) to the record
    sums) against central differences of analysis.akc_mom."""
    import importlib
    A = importlib.import_module(rms.__name__ + ".analysis")
    rng = np.random.default_rng(4)
    T, nclass = 5, 2
    nv = rms.nval(nclass, 0)
    rec = rng.uniform(0.5, 2.0, (T, nv))
    rec[:, 0] = rng.uniform(900, 1000, T)
    TD = np.linspace(0.1, 0.5, T)
    n = np.array([0.3, -0.5, 0.8])
    n /= np.linalg.norm(n)
    wD, wK = rng.standard_normal(T), rng.standard_normal(T)

    def L(r):
        s = A.split(r, nclass)
        K, D = A.akc_mom(s["sig0"], s["dx2"], s["dx4"], TD, n)
        return float(np.sum(wD * D[:, 0]) + np.sum(wK * K[:, 0]))

    cot = A.akc_cotangent(rec, nclass, TD, n, wD, wK)
    num = np.zeros_like(rec)
    for idx in np.ndindex(rec.shape):
        h = 1e-6 * abs(rec[idx])
        rp, rm = rec.copy(), rec.copy()
        rp[idx] += h
        rm[idx] -= h
        num[idx] = (L(rp) - L(rm)) / (2 * h)
    np.testing.assert_allclose(cot, num, rtol=1e-6, atol=1e-9 * np.abs(num).max())
