"""Central differences of D on walker subsets, pinned on the CPU oracle.

The adjoint's central differences for class c walk only the walkers whose
walk can change when D_c moves: those the tangent pass saw in class c, and
those that drew at an exchange face whose P differs between D_c (1 - h),
D_c and D_c (1 + h).  The claim behind it is checked walker by walker: every
other walker ends in the same state, bit for bit, at all three points (a
missing mark -- a class change or a draw not recorded, a dependence of P
missed -- leaves some such walker with another final state).  The
differences on the subsets then equal those of every walker to rounding."""
import numpy as np
import pytest

from helpers import fib_dirs, random_blobs, slabs_z

CASES = {
    # permeability: P_21 = pi0/(1+pi0) with pi0 from class 2's smaller step
    # does not move with D_1; P_12 does (the square root)
    "kappa": dict(labels=3, D=[0.0, 2.0, 0.8], kappa=0.2),
    # a large permeability (P of a few per cent): draws at faces into class 2
    # decide differently at D_2 (1 +- h) for some walkers that stay outside it
    "strong": dict(labels=3, D=[0.0, 2.0, 0.8], kappa=5.0),
    # three non-null classes, every pair drawing
    "three": dict(labels=4, D=[0.0, 2.0, 0.7, 1.3], kappa=0.3),
    # the reference rule without a permeability: class 2 -> 1 crosses
    # without a draw (P = 1), the class change marks those walkers
    "nokappa": dict(labels=3, D=[0.0, 2.0, 0.8], kappa=None),
}


def _geom(orc, cfg, D=None):
    lab = slabs_z(cfg["labels"])
    kw = {} if cfg["kappa"] is None else dict(kappa=cfg["kappa"])
    nc = len(cfg["D"])
    return orc.Geometry(lab, 0.1, cfg["D"] if D is None else D, T2=[np.inf, 70.0, 40.0, 55.0][:nc], **kw)


@pytest.mark.parametrize("case", sorted(CASES))
def test_walkers_outside_the_subset_walk_identically(orc, case):
    cfg = CASES[case]
    g = _geom(orc, cfg)
    nc = g.nclass
    job = dict(geom=g, nwalk=1280, seed=9, bval=np.array([1.0]), bvec=fib_dirs(1))
    # a wide step (h = 0.3) widens the windows of P in which a draw decides
    # differently, so a missing mark shows in a small run
    dt, TN, T, h = 2e-4, 600, 6, 0.3
    cot = [np.random.default_rng(1).standard_normal((T, orc.nval(nc, 1)))]
    vis = orc.run([job], dt, TN, T, cot_rec=cot, boundary=False, marks=True)[0]["vis"]
    base = orc.run([job], dt, TN, T, final_state=True)[0]
    assert np.all(vis & (1 << np.arange(nc)).sum() != 0)  # every walker visited its start class
    if cfg["labels"] >= 3:
        assert np.count_nonzero(vis >> 4) > 0  # draws happened
    sizes = []
    for c in range(1, nc):
        mask = orc.fd_walker_mask(g, dt, c, h)
        assert mask is not None
        member = (vis & np.uint32(mask)) != 0
        sizes.append(np.count_nonzero(member))
        moved = np.zeros(job["nwalk"], bool)
        for f in (1.0 + h, 1.0 - h):
            Dc = list(cfg["D"])
            Dc[c] = cfg["D"][c] * f
            r = orc.run([dict(job, geom=_geom(orc, cfg, Dc))], dt, TN, T, final_state=True)[0]
            same = np.ones(job["nwalk"], bool)
            for k in ("fin_g", "fin_o", "fin_gi", "fin_oi", "fin_cnt"):
                same &= np.all(r[k] == base[k], axis=0)
            same &= r["fin_c"] == base["fin_c"]
            assert np.all(same[~member]), (case, c, f, np.nonzero(~same & ~member)[0][:8])
            moved |= ~same
        assert np.count_nonzero(moved[member]) > 0  # the walks that can change do
    assert 0 < min(sizes) < job["nwalk"], sizes  # some class leaves walkers out


@pytest.mark.parametrize("case", sorted(CASES))
def test_subset_differences_equal_every_walker(orc, case):
    cfg = CASES[case]
    g = _geom(orc, cfg)
    nc = g.nclass
    jobs = [dict(geom=g, nwalk=700, seed=12, bval=np.array([0.8, 2.0]), bvec=fib_dirs(2))]
    dt, TN, T, h, R = 2e-4, 400, 8, 0.05, 4
    rng = np.random.default_rng(2)
    cot = [rng.standard_normal((T, orc.nval(nc, 2)))]
    gs, gsr = orc.central_differences(jobs, dt, TN, T, cot, h=h, replicas=R, subsets=True)
    ga, _ = orc.central_differences(jobs, dt, TN, T, cot, h=h, replicas=R, subsets=False)
    rec = orc.run(jobs, dt, TN, T, replicas=R)[0]["rec"]
    for c in range(1, nc):
        tol = 1e-12 * np.sum(np.abs(cot[0]) * np.abs(rec)) / (2 * h * cfg["D"][c])
        assert abs(gs[0, c] - ga[0, c]) <= tol, (c, gs[0, c], ga[0, c], tol)
        assert gs[0, c] != 0.0
        assert abs(gsr[0, :, c].sum() - gs[0, c]) <= tol + 1e-12 * np.abs(gsr[0, :, c]).sum()


def test_fd_mask_follows_the_dependence_of_P(orc):
    """With a permeability, class 2's step sets pi0 (D_2 < D_1), so D_1 moves
    only P_12 (the square root) and D_2 moves both; without one P_21 = 1 at
    every point; with D_1 = 1, D_2 = 1.02 and no permeability P_12 = 1 leaves
    1 at D_1 (1 + h): no trustworthy mask."""
    lab = random_blobs((6, 6, 6), 3, seed=1)
    b12, b21 = 1 << (4 + 4 + 2), 1 << (4 + 8 + 1)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0, 0.8], kappa=0.2)
    assert orc.fd_walker_mask(g, 2e-4, 1, 0.05) == 2 | b12
    assert orc.fd_walker_mask(g, 2e-4, 2, 0.05) == 4 | b12 | b21
    g = orc.Geometry(lab, 0.1, [0.0, 2.0, 0.8])
    assert orc.fd_walker_mask(g, 2e-4, 1, 0.05) == 2 | b12
    assert orc.fd_walker_mask(g, 2e-4, 2, 0.05) == 4 | b12
    g = orc.Geometry(lab, 0.1, [0.0, 1.0, 1.02])
    assert orc.fd_walker_mask(g, 2e-4, 1, 0.05) is None
