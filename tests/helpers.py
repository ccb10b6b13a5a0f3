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
