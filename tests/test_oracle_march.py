"""Pins of the oracle's face march (main.cu:153-300) against closed-form ray
geometry: straight segments in free lattices, triangle-wave folding at
reflecting walls and lattice-edge mirrors, corner ties, and the step-length
change when a walker crosses into a class with another diffusivity."""
import math

import numpy as np
import pytest

DT = 2e-4


def _step(orc, D):
    return math.sqrt(6.0 * DT * D) / 0.1


def _unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def _pos(r):
    return r["g"] + r["o"]


def _fold(x, lo, hi):
    """Position of a point reflected between walls at lo and hi."""
    L = hi - lo
    y = (x - lo) % (2 * L)
    return lo + (y if y <= L else 2 * L - y)


def test_free_segment_matches_straight_line(orc):
    lab = np.ones((9, 9, 9), np.uint8)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    s = _step(orc, 2.0)
    rng = np.random.default_rng(0)
    for _ in range(400):
        x0 = 3.0 + rng.random(3) * 3.0
        v = _unit(rng.normal(size=3))
        g0 = np.floor(x0).astype(int)
        r = orc.march_f64(g, DT, g0, x0 - g0, v.copy(), 1)
        end = x0 + s * v
        np.testing.assert_allclose(_pos(r), end, rtol=0, atol=1e-12)
        np.testing.assert_array_equal(r["g"], np.floor(end).astype(int))
        crossings = int(np.sum(np.abs(np.floor(end) - np.floor(x0))))
        assert r["cnt"][0] == crossings and r["cnt"][1] == 0 and r["cnt"][3] == 0


def test_reflecting_walls_fold_the_path(orc):
    """Label-0 slabs at z=0 and z=5: the walker reflects between z=1 and z=5."""
    lab = np.ones((6, 6, 6), np.uint8)
    lab[:, :, 0] = 0
    lab[:, :, 5] = 0
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    s = _step(orc, 2.0)
    rng = np.random.default_rng(1)
    nrefl = 0
    for _ in range(400):
        x0 = np.array([2.5, 2.5, 1.0]) + rng.random(3) * np.array([1.0, 1.0, 4.0])
        v = _unit(rng.normal(size=3))
        g0 = np.floor(x0).astype(int)
        r = orc.march_f64(g, DT, g0, x0 - g0, v.copy(), 1)
        end = x0 + s * v
        want = np.array([end[0], end[1], _fold(end[2], 1.0, 5.0)])
        np.testing.assert_allclose(_pos(r), want, atol=1e-12)
        hits = int(end[2] < 1.0) + int(end[2] > 5.0)
        assert r["cnt"][1] == hits
        assert np.sign(r["v"][2]) == (np.sign(v[2]) * (-1) ** hits)
        nrefl += hits
    assert nrefl > 20


def test_lattice_edge_mirror_and_origin_unfolding(orc):
    """Edges of a 2x1x1 all-walkable lattice are mirrors (main.cu:249-265) and
    the origin is mirrored with the walker (main.cu:302-310): the recorded
    displacement is the free displacement with its sign flipped per hit."""
    lab = np.ones((2, 1, 1), np.uint8)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    s = _step(orc, 2.0)
    N = np.array([2.0, 1.0, 1.0])
    rng = np.random.default_rng(2)
    for _ in range(400):
        x0 = rng.random(3) * N
        v = _unit(rng.normal(size=3))
        g0 = np.floor(x0).astype(int)
        r = orc.march_f64(g, DT, g0, x0 - g0, v.copy(), 1)
        end = x0 + s * v
        want = np.array([_fold(end[k], 0.0, N[k]) for k in range(3)])
        np.testing.assert_allclose(_pos(r), want, atol=1e-12)
        hits = np.array([int(end[k] < 0) + int(end[k] > N[k]) for k in range(3)])
        assert r["cnt"][2] == hits.sum()
        disp = (r["g"] - r["gi"]) + (r["o"] - r["oi"])
        np.testing.assert_allclose(disp, (-1.0) ** hits * (s * v), atol=1e-12)


def test_corner_tie_crosses_all_faces(orc):
    lab = np.ones((4, 4, 4), np.uint8)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0])
    s = _step(orc, 2.0)
    v = _unit([1.0, 1.0, 1.0])
    # start so that the diagonal meets the corner (2,2,2) within the step
    x0 = np.array([2.0, 2.0, 2.0]) - 0.2 * v
    g0 = np.floor(x0).astype(int)
    r = orc.march_f64(g, DT, g0, x0 - g0, v.copy(), 1)
    np.testing.assert_array_equal(r["g"], [2, 2, 2])
    np.testing.assert_allclose(_pos(r), x0 + s * v, atol=1e-12)
    assert r["cnt"][0] == 3 and r["cnt"][3] == 0


def test_class_change_rescales_the_remaining_step(orc):
    """Crossing into a class with D2 continues with the remaining fraction of
    the step at step[c2] (main.cu:231-236 with ai=aj, :274-276)."""
    lab = np.ones((6, 4, 4), np.uint8)
    lab[3:, :, :] = 2
    g = orc.Geometry(lab, 0.1, [0.0, 2.0, 0.5], kappa=None)
    s1, s2 = _step(orc, 2.0), _step(orc, 0.5)
    P = orc.derive(g, DT)["P"]
    assert P[1, 2] == 0.5 and P[2, 1] == 1.0
    rng = np.random.default_rng(3)
    for _ in range(200):
        v = _unit([1.0, 0.3 * rng.normal(), 0.3 * rng.normal()])
        x0 = np.array([3.0 - 0.3 * rng.random(), 1.5, 1.5])
        g0 = np.floor(x0).astype(int)
        # u=0.1 < P(1->2)=0.5: the exchange draw accepts the crossing
        r = orc.march_f64(g, DT, g0, x0 - g0, v.copy(), 1, u_exch=(0.1, 0.1, 0.1, 0.1))
        d1 = (3.0 - x0[0]) / v[0]
        if d1 <= s1:
            want = x0 + d1 * v + s2 * (1.0 - d1 / s1) * v
            assert r["cls"] == 2 and r["draws"] == 1
        else:
            want = x0 + s1 * v
            assert r["cls"] == 1
        np.testing.assert_allclose(_pos(r), want, atol=1e-12)
        # u=0.9 > 0.5: rejected, elastic reflection off the face x=3
        r = orc.march_f64(g, DT, g0, x0 - g0, v.copy(), 1, u_exch=(0.9, 0.9, 0.9, 0.9))
        end = x0 + s1 * v
        want = np.array([3.0 - (end[0] - 3.0) if end[0] > 3.0 else end[0], end[1], end[2]])
        np.testing.assert_allclose(_pos(r), want, atol=1e-12)
        assert r["cls"] == 1


@pytest.mark.parametrize("D", [0.5, 2.0, 3.0])
def test_step_length_is_conserved(orc, D):
    """Path length of a step equals step[c] (no truncation for step < 1)."""
    lab = np.ones((5, 5, 5), np.uint8)
    lab[2, 2, 2] = 0
    g = orc.Geometry(lab, 0.1, [0.0, D])
    s = _step(orc, D)
    rng = np.random.default_rng(4)
    for _ in range(300):
        x0 = 1.0 + rng.random(3) * 3.0
        g0 = np.floor(x0).astype(int)
        if lab[tuple(g0)] == 0:
            continue
        v = _unit(rng.normal(size=3))
        r = orc.march_f64(g, DT, g0, x0 - g0, v.copy(), 1)
        assert r["cnt"][3] == 0
        # unfolded path: |v| stays 1 and the walker never sits in the null voxel
        assert abs(np.linalg.norm(r["v"]) - 1.0) < 1e-14
        assert lab[tuple(r["g"])] != 0
        assert np.linalg.norm(_pos(r) - x0) <= s + 1e-12


def test_fast_path_equals_march(orc):
    """The straight-move fast path against the face march on random states in
    a lattice with null voxels, a second class and lattice edges: whenever the
    fast path applies it ends in the same voxel and class with the same
    crossing/reflection/mirror counts and the same mirrored origin, at the
    same position up to rounding; and it applies to most steps."""
    rng = np.random.default_rng(11)
    lab = rng.integers(0, 3, (7, 6, 8)).astype(np.uint8)
    lab[2:5, 2:5, 2:6] = 1
    g = orc.Geometry(lab, 0.1, [0.0, 2.0, 2.0], T2=[np.inf] * 3, kappa=0.0)  # classes 1, 2 impermeable
    dt = 2e-4
    step = orc.derive(g, dt)["step"][1]
    applied = mirrored = 0
    n = 4000
    for _ in range(n):
        while True:
            gv = rng.integers(0, [7, 6, 8])
            cls = int(lab[tuple(gv)])
            if cls != 0:
                break
        o = rng.random(3)
        u1, u2 = rng.random(2)
        cth = 1 - 2 * u1
        sth = 2 * np.sqrt(u1 * (1 - u1))
        v = np.array([sth * np.cos(2 * np.pi * u2), sth * np.sin(2 * np.pi * u2), cth])
        f = orc.fast_step_f64(g, dt, gv, o, v, cls)
        m = orc.march_f64(g, dt, gv, o, v, cls)
        if f is None:
            continue
        applied += 1
        np.testing.assert_array_equal(f["g"], m["g"])
        assert m["cls"] == cls
        np.testing.assert_allclose(f["o"], m["o"], atol=1e-12)
        np.testing.assert_array_equal(f["cnt"][:3], m["cnt"][:3])  # crossings, reflections, mirrors
        np.testing.assert_array_equal(f["gi"], m["gi"])
        np.testing.assert_array_equal(f["oi"], m["oi"])
    assert applied > 0.3 * n
    # lattice edges of a voxel whose other neighbours are all outside: mirrors
    # on the fast path, same voxel, counts and mirrored origin as the march
    g1 = orc.Geometry(np.ones((4, 5, 6), np.uint8), 0.1, [0.0, 2.0], T2=[np.inf] * 2)
    for _ in range(2000):
        gv = rng.integers(0, [4, 5, 6])
        o = rng.random(3)
        u1, u2 = rng.random(2)
        sth = 2 * np.sqrt(u1 * (1 - u1))
        v = np.array([sth * np.cos(2 * np.pi * u2), sth * np.sin(2 * np.pi * u2), 1 - 2 * u1])
        gi, oi = rng.integers(-3, 9, 3), rng.random(3)
        f = orc.fast_step_f64(g1, dt, gv, o, v, 1, gi=gi, oi=oi)
        m = orc.march_f64(g1, dt, gv, o, v, 1, gi=gi, oi=oi)
        assert f is not None
        np.testing.assert_array_equal(f["g"], m["g"])
        np.testing.assert_allclose(f["o"], m["o"], atol=1e-12)
        np.testing.assert_array_equal(f["cnt"][:3], m["cnt"][:3])
        np.testing.assert_array_equal(f["gi"], m["gi"])
        np.testing.assert_array_equal(f["oi"], m["oi"])
        mirrored += f["cnt"][2] > 0
    assert mirrored > 300
    # pure interior: always applies and is the straight move
    f = orc.fast_step_f64(g, dt, [3, 3, 3], [0.5, 0.5, 0.5], [0.6, 0.0, 0.8], 1)
    np.testing.assert_allclose(f["o"], [0.5 + step * 0.6, 0.5, 0.5 + step * 0.8], atol=1e-15)


def test_fast_path_equals_march_periodic(orc):
    """Periodic x and z, mirror y: a face on a periodic lattice side leads to
    the voxel on the other side (the voxel index wraps, the origin moves by
    one lattice length), on the fast path as in the march; the fast path
    equals the march on random states, wrapped steps included: same end
    voxel, counts and origin, same position up to rounding."""
    rng = np.random.default_rng(12)
    lab = rng.integers(0, 3, (5, 6, 4)).astype(np.uint8)
    lab[rng.random(lab.shape) < 0.5] = 1
    g = orc.Geometry(lab, 0.1, [0.0, 2.0, 2.0], T2=[np.inf] * 3, kappa=0.0, periodic="xz")
    dt = 2e-4
    applied = wrapped = fast_wrapped = 0
    for _ in range(4000):
        while True:
            gv = rng.integers(0, [5, 6, 4])
            cls = int(lab[tuple(gv)])
            if cls != 0:
                break
        o = rng.random(3)
        u1, u2 = rng.random(2)
        sth = 2 * np.sqrt(u1 * (1 - u1))
        v = np.array([sth * np.cos(2 * np.pi * u2), sth * np.sin(2 * np.pi * u2), 1 - 2 * u1])
        gi, oi = rng.integers(-3, 9, 3), rng.random(3)
        f = orc.fast_step_f64(g, dt, gv, o, v, cls, gi=gi, oi=oi)
        m = orc.march_f64(g, dt, gv, o, v, cls, gi=gi, oi=oi)
        assert m["gi"][0] - gi[0] in (-5, 0, 5) and m["gi"][2] - gi[2] in (-4, 0, 4)  # wraps, not mirrors
        assert np.all((m["g"] >= 0) & (m["g"] < [5, 6, 4]))
        wrapped += int(m["gi"][0] != gi[0] or m["gi"][2] != gi[2])
        if f is None:
            continue
        applied += 1
        fast_wrapped += int(f["gi"][0] != gi[0] or f["gi"][2] != gi[2])
        np.testing.assert_array_equal(f["g"], m["g"])
        np.testing.assert_allclose(f["o"], m["o"], atol=1e-12)
        np.testing.assert_array_equal(f["cnt"][:3], m["cnt"][:3])
        np.testing.assert_array_equal(f["gi"], m["gi"])
        np.testing.assert_array_equal(f["oi"], m["oi"])
    assert applied > 0.3 * 4000 and wrapped > 100 and fast_wrapped > 0.5 * wrapped


def test_fast_path_three_faces_in_order(orc):
    """Steps that cross three faces next to random null voxels of one octant:
    the fast path meets them in the order of the face distances and agrees
    with the face march (same voxel, counts, position up to rounding) on every
    step whose comparisons are transitive, including the steps whose outcome
    depends on that order."""
    rng = np.random.default_rng(5)
    dt = 2e-4
    n_applied = n_mixed = 0
    for trial in range(400):
        lab = np.ones((5, 5, 5), np.uint8)
        for S in range(1, 8):
            if rng.random() < 0.45:
                lab[2 + (S & 1), 2 + ((S >> 1) & 1), 2 + ((S >> 2) & 1)] = 0
        g = orc.Geometry(lab, 0.1, [0.0, 2.0], T2=[np.inf] * 2)
        o = rng.uniform(0.8, 0.99, 3)
        v = rng.uniform(0.45, 0.7, 3)
        v /= np.linalg.norm(v)
        f = orc.fast_step_f64(g, dt, [2, 2, 2], o, v, 1)
        m = orc.march_f64(g, dt, [2, 2, 2], o, v, 1)
        assert f is not None, (trial, lab[2:4, 2:4, 2:4])
        n_applied += 1
        np.testing.assert_array_equal(f["g"], m["g"])
        np.testing.assert_allclose(f["o"], m["o"], atol=1e-12)
        assert f["cnt"][0] == m["cnt"][0] and f["cnt"][1] == m["cnt"][1]
        assert f["cnt"][0] + f["cnt"][1] == 3  # three faces met
        n_mixed += 0 < f["cnt"][1] < 3
    assert n_applied == 400 and n_mixed > 100


def test_fast_path_exchange_faces_equal_march(orc):
    """With two non-null classes and a permeable membrane, a step meeting
    exchange faces takes the fast path when every draw reflects (u >= P) and
    then equals the march given the same draws (same voxel, class, counts,
    draws, position up to rounding); a crossing draw leaves the step to the
    march."""
    rng = np.random.default_rng(17)
    lab = rng.integers(1, 3, (7, 6, 8)).astype(np.uint8)  # two classes, no null voxel
    g = orc.Geometry(lab, 0.1, [0.0, 2.0, 1.5], T2=[np.inf] * 3, kappa=0.05)
    dt = 2e-4
    P = orc.derive(g, dt)["P"]
    assert 0.0 < P[1][2] < 1.0 and 0.0 < P[2][1] < 1.0
    n_draw = n_cross_declined = 0
    for _ in range(3000):
        while True:
            gv = rng.integers(1, [6, 5, 7])  # interior voxels: no lattice edge next door
            cls = int(lab[tuple(gv)])
            if cls != 0:
                break
        o = rng.random(3)
        u1, u2 = rng.random(2)
        sth = 2 * np.sqrt(u1 * (1 - u1))
        v = np.array([sth * np.cos(2 * np.pi * u2), sth * np.sin(2 * np.pi * u2), 1 - 2 * u1])
        for u in (0.999, 0.0):
            f = orc.fast_step_f64(g, dt, gv, o, v, cls, u_exch=(u,) * 4)
            m = orc.march_f64(g, dt, gv, o, v, cls, u_exch=(u,) * 4)
            if f is None:
                n_cross_declined += u == 0.0 and m["draws"] > 0
                continue
            np.testing.assert_array_equal(f["g"], m["g"])
            assert m["cls"] == cls
            np.testing.assert_allclose(f["o"], m["o"], atol=1e-12)
            assert list(f["cnt"]) == [m["cnt"][0], m["cnt"][1], m["cnt"][2], m["draws"]]
            assert u == 0.999 or f["cnt"][3] == 0  # a crossing draw never stays on the fast path
            n_draw += f["cnt"][3] > 0
    assert n_draw > 100 and n_cross_declined > 100
