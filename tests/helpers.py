"""Shared small geometries and run helpers for the parity tests."""
import numpy as np


def planes(nx=8, ny=8, nz=14, lo=0, hi=None, dtype=np.uint8):
    lab = np.ones((nx, ny, nz), dtype)
    hi = nz - 1 if hi is None else hi
    lab[:, :, lo] = 0
    lab[:, :, hi] = 0
    return lab


def random_blobs(shape, nclass, seed, p_null=0.15, dtype=np.uint8):
    rng = np.random.default_rng(seed)
    lab = rng.integers(1, nclass, size=shape).astype(dtype)
    # smooth a bit so compartments are contiguous
    for _ in range(2):
        lab = np.where(rng.random(shape) < 0.5, np.roll(lab, 1, axis=rng.integers(0, 3)), lab)
    lab[rng.random(shape) < p_null] = 0
    return np.ascontiguousarray(lab.astype(dtype))


def slabs_z(nlabel, nxy=6, depth=30, p_null=0.05):
    """Slabs of `depth` voxels along z, label k in the k-th, a few null voxels
    scattered: walkers deep in a slab do not reach another label for a while."""
    lab = np.zeros((nxy, nxy, depth * (nlabel - 1)), np.uint8)
    for k in range(1, nlabel):
        lab[:, :, depth * (k - 1):depth * k] = k
    lab[random_blobs(lab.shape, 2, seed=3, p_null=p_null) == 0] = 0
    return lab


def cylinder(n=24, nz=10, r=8.0, dtype=np.uint8):
    x = np.arange(n) + 0.5 - n / 2
    X, Y = np.meshgrid(x, x, indexing="ij")
    m = (X ** 2 + Y ** 2) <= r * r
    lab = np.zeros((n, n, nz), dtype)
    lab[m, :] = 1
    return lab


def pgse_table(dirs, bvals, Deltas, delta):
    rows = []
    for D_ in Deltas:
        for b in bvals:
            g = np.sqrt(b / (delta ** 2 * (D_ - delta / 3)))
            for d in dirs:
                d = np.asarray(d, float)
                d = d / np.linalg.norm(d)
                rows.append([D_, delta, g, d[0], d[1], d[2]])
    return np.array(rows)


def fib_dirs(n):
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)


def assert_same(gpu, ref, final_state=False):
    """Bit-exact equality of engine and oracle results (records, replicas,
    PGSE signal, counters and optionally the final walker states)."""
    for a, b in zip(gpu, ref):
        np.testing.assert_array_equal(a["counters"], b["counters"])
        if final_state:
            for k in ("fin_g", "fin_gi", "fin_c", "fin_cnt", "fin_o", "fin_oi"):
                np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        np.testing.assert_array_equal(a["rec"], b["rec"])
        np.testing.assert_array_equal(a["rec_rep"], b["rec_rep"])
        if b["sig"] is not None:
            np.testing.assert_array_equal(a["sig"], b["sig"])
            np.testing.assert_array_equal(a["sig_rep"], b["sig_rep"])


def planes_moments(a, D0, t, nterm=400, nodes=4000):
    """<dz^2>(t) and <dz^4>(t) for diffusion between reflecting walls a apart
    from a uniform start: the eigen-expansion
    P(z, t | z0) = (1/a) [1 + 2 sum_n cos(n pi z/a) cos(n pi z0/a) exp(-n^2 pi^2 D0 t / a^2)],
    with the double integrals of (z - z0)^k split by the binomial theorem into
    products of one-dimensional integrals (Gauss-Legendre)."""
    from math import comb
    x, w = np.polynomial.legendre.leggauss(nodes)
    z = 0.5 * a * (x + 1.0)
    w = 0.5 * a * w
    n = np.arange(1, nterm + 1)
    cosn = np.cos(np.pi * n[:, None] * z[None, :] / a)       # (nterm, nodes)
    I = np.stack([(cosn * (z ** j)[None, :]) @ w for j in range(5)])   # I[j][n] = int z^j cos(n pi z/a)
    M = np.array([np.sum(w * z ** j) for j in range(5)])              # int z^j
    t = np.asarray(t, float)
    decay = np.exp(-(n[None, :] ** 2) * np.pi ** 2 * D0 * t[:, None] / (a * a))
    out = []
    for k in (2, 4):
        c0 = sum(comb(k, j) * (-1) ** (k - j) * M[j] * M[k - j] for j in range(k + 1)) / a ** 2
        cn = sum(comb(k, j) * (-1) ** (k - j) * I[j] * I[k - j] for j in range(k + 1))
        out.append(c0 + 2.0 / a ** 2 * decay @ cn)
    return out[0], out[1]


def planes_x2_series(a, D0, t, nterm=400):
    """<dz^2>(t) between reflecting walls: a^2/6 - (16 a^2/pi^4) sum_{n odd} n^-4 exp(-n^2 pi^2 D0 t/a^2)."""
    n = np.arange(1, 2 * nterm, 2)[None, :]
    t = np.asarray(t)[:, None]
    return a * a / 6 - 16 * a * a / np.pi ** 4 * np.sum(n ** -4.0 * np.exp(-(n ** 2) * np.pi ** 2 * D0 * t / (a * a)), 1)


def gaussian_phase_signal(abe, grad, dt, TN, D0):
    """Free-diffusion PGSE signal exp(-Var(phi)/2) with the exact discrete
    variance Var(phi) = |g|^2 dt^2 2 D0 dt sum_k W_k^2, W_k = sum_{m >= k} w_m,
    w = +1 on lobe 1, -1 on lobe 2 (right-endpoint rule, main_PGSE.cu:340-346)."""
    out = np.zeros(len(abe))
    for j, (a_, b_, e_) in enumerate(abe):
        w = np.zeros(TN)
        w[:a_] = 1.0
        w[b_:e_] = -1.0
        W = np.cumsum(w[::-1])[::-1]
        var = grad[j, 2] ** 2 * dt ** 2 * 2 * D0 * dt * np.sum(W ** 2)
        out[j] = np.exp(-var / 2)
    return out
