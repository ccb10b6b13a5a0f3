"""Synthetic voxelised geometries for the benchmark workloads.

  planes      impermeable parallel planes (label-0 slabs), the parallel-plane
              setup of the reference's elastic-collision demos
  axon        beaded, undulating axon: radius r(z) = r0 (1 + sum of
              sinusoids), caliber CV measured per z-slice as
              part1_demo4_analysis.m:47-48 does (r = sqrt(area/pi)), centre
              line displaced by an undulation (analysis/visualizeaxon.m:42-67
              sweeps spheres along such a line)
  bundle      many axons with gamma-distributed radii, each in its own uint16
              lattice with an extra-axonal shell (labels 0 null, 1 intra,
              2 extra)
Lattices are x-major with z fastest, the order of fiber.txt.
"""
from __future__ import annotations

import numpy as np


def planes(nx=32, ny=32, nz=52, dtype=np.uint8):
    """Label 1 between label-0 slabs at z=0 and z=nz-1 (gap (nz-2) voxels)."""
    lab = np.ones((nx, ny, nz), dtype)
    lab[:, :, 0] = 0
    lab[:, :, nz - 1] = 0
    return lab


def _divisor_wavelengths(rng, n, lo, hi, length):
    """n wavelengths log-uniform in [lo, hi], snapped to length/k so the
    profile is periodic over `length`."""
    lam = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    k = np.maximum(1, np.round(length / lam))
    return length / k


def radius_profile(nz, res, r0, cv, rng, nsin=8, lam_lo=2.0, lam_hi=20.0):
    """r(z) = r0 * (1 + sum_k a_k sin(2 pi z / lam_k + phi_k)) with the
    amplitudes scaled so that std(r)/mean(r) = cv over the z samples."""
    L = nz * res
    z = (np.arange(nz) + 0.5) * res
    lam = _divisor_wavelengths(rng, nsin, lam_lo, lam_hi, L)
    phi = rng.uniform(0, 2 * np.pi, nsin)
    amp = rng.uniform(0.5, 1.0, nsin)
    s = np.sum(amp[:, None] * np.sin(2 * np.pi * z[None, :] / lam[:, None] + phi[:, None]), 0)
    if cv > 0:
        s = s / s.std() * cv
    else:
        s = np.zeros_like(s)
    return r0 * np.maximum(1 + s, 0.15)


def measured_cv(lab, res, label=1):
    """Caliber CV as part1_demo4_analysis.m:47-48: r = sqrt(area/pi) per slice."""
    area = np.sum(lab == label, axis=(0, 1)) * res * res
    r = np.sqrt(area / np.pi)
    return float(r.std(ddof=1) / r.mean())


def axon(n=64, nz=1000, res=0.1, r0=0.8, cv=0.2, undulation=0.5, lam_u=20.0, seed=1235, ecs=False,
         dtype=np.uint8, nsin=8):
    """Beaded, undulating axon in an n x n x nz lattice (label 1 inside).

    With ecs=True every other voxel of the lattice is extra-axonal (label 2).
    The caliber CV is matched on the voxelised shape by rescaling the
    sinusoid amplitudes."""
    rng = np.random.default_rng(seed)
    L = nz * res
    z = (np.arange(nz) + 0.5) * res
    lam_u = L / max(1, round(L / lam_u))
    psi = rng.uniform(0, 2 * np.pi)
    cx = undulation * np.cos(2 * np.pi * z / lam_u + psi)
    cy = undulation * np.sin(2 * np.pi * z / lam_u + psi)
    state = rng.bit_generator.state
    x = (np.arange(n) + 0.5 - n / 2) * res
    X, Y = np.meshgrid(x, x, indexing="ij")
    target = cv
    scale = 1.0
    lab = None
    for _ in range(8):
        rng.bit_generator.state = state
        r = radius_profile(nz, res, r0, target * scale, rng, nsin=nsin)
        d2 = (X[:, :, None] - cx[None, None, :]) ** 2 + (Y[:, :, None] - cy[None, None, :]) ** 2
        inside = d2 <= (r * r)[None, None, :]
        lab = inside.astype(dtype)
        if cv <= 0:
            break
        got = measured_cv(lab, res)
        if abs(got - target) < 0.005:
            break
        scale *= target / max(got, 1e-6)
    if ecs:
        lab = np.where(lab == 1, 1, 2).astype(dtype)
    return np.ascontiguousarray(lab)


def bundle(naxon=256, mean_r=0.6, cv_r=0.4, len_lo=20.0, len_hi=200.0, res=0.1, shell=1.0, caliber_cv=0.1,
           seed=1236, max_vox=5_000_000, dtype=np.uint16):
    """Axons with gamma radii (mean mean_r, CV cv_r) and uniform lengths, each
    in its own lattice: intra (1) inside r(z), extra-axonal (2) in a shell of
    width `shell` um around it, null (0) beyond.  Lattice sizes are uneven.
    Returns a list of uint16 arrays."""
    rng = np.random.default_rng(seed)
    k = 1.0 / cv_r ** 2
    radii = rng.gamma(k, mean_r / k, naxon)
    radii = np.clip(radii, 0.2, 3.0)
    lengths = rng.uniform(len_lo, len_hi, naxon)
    out = []
    for i in range(naxon):
        r0 = radii[i]
        nz = max(20, int(round(lengths[i] / res)))
        n = int(np.ceil(2 * (r0 * (1 + 3 * caliber_cv) + shell) / res)) + 2
        while n * n * nz > max_vox and nz > 20:
            nz = int(nz * 0.9)
        rr = radius_profile(nz, res, r0, caliber_cv, rng, nsin=4)
        x = (np.arange(n) + 0.5 - n / 2) * res
        X, Y = np.meshgrid(x, x, indexing="ij")
        d = np.sqrt(X ** 2 + Y ** 2)[:, :, None]
        lab = np.zeros((n, n, nz), dtype)
        lab[d <= (rr + shell)[None, None, :]] = 2
        lab[d <= rr[None, None, :]] = 1
        out.append(np.ascontiguousarray(lab))
    return out
