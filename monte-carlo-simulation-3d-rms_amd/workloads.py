"""The five benchmark workloads as job lists (one job = one reference setup
directory: a geometry, its walkers, its b-table or PGSE table).

  c1  impermeable parallel planes 5 um apart, 32x32x52 uint8 at 0.1 um (label-0
      slabs at z = 0 and 51), x and y periodic, 1e4 walkers, D0 = 2, 1e4 steps,
      records every 100 steps (the elastic-collision parallel-plane demo,
      part2_demo1_2_diffusivity.m:25-45)
  c2  one beaded, undulating axon, 64x64x1000 uint8, 1e6 walkers, 1e5 steps,
      PGSE 30 directions x b in {0.5,1,2,3} x Delta in {10,20,40} at delta = 5
  c3  256 axons with extra-axonal shells in uneven uint16 lattices, kappa in
      [0, 0.02], T2 80/60, D 2.0/1.5, 1e7 walkers, 5e4 steps, the PGSE table
  c4  the c2 axon (periodic lattice) with extra-axonal space, kappa = 0.01,
      1e6 walkers, 5e4 steps, 20 records, narrow-pulse 64 directions x b in
      {0.5,1,2,3}
  c5  4096 axons with extra-axonal space, 70x70x1000 uint16, caliber CV,
      undulation and kappa drawn per geometry, 4e7 walkers, 2e4 steps,
      20 records, narrow-pulse 16 directions x b in {1,2}

Each builder takes a walker count so tests can run the same geometries and
tables on fewer walkers.  The geometry class is the caller's (any class with
the constructor of engine.Geometry), so the same definitions feed every
consumer.  Seeds are fixed here.
"""
from __future__ import annotations

import math

import numpy as np

from . import geometry as G

DT = 2e-4
RES = 0.1


def fib_dirs(n):
    """n directions on a Fibonacci sphere."""
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)


def pgse_rows(dirs, bvals, Deltas, delta):
    """Rows (Delta, delta, |g|, gx, gy, gz) with b = g^2 delta^2 (Delta - delta/3)
    (readRMS_pgse.m:53), Delta outermost, then b, then direction."""
    rows = []
    for Delta in Deltas:
        for b in bvals:
            g = math.sqrt(b / (delta ** 2 * (Delta - delta / 3)))
            for d in dirs:
                rows.append([Delta, delta, g, *d])
    return np.array(rows)


def pgse_table_c2():
    """30 directions x b in {0.5,1,2,3} ms/um^2 x Delta in {10,20,40} ms, delta = 5 ms."""
    return pgse_rows(fib_dirs(30), (0.5, 1.0, 2.0, 3.0), (10.0, 20.0, 40.0), 5.0)


def narrow_table(ndir, bvals):
    """Narrow-pulse b-table: every direction at every b (b outermost)."""
    dirs = fib_dirs(ndir)
    return np.repeat(np.asarray(bvals, float), ndir), np.tile(dirs, (len(bvals), 1))


# ---------------------------------------------------------------- c1
C1 = dict(TN=10_000, timepoints=100, nwalk=10_000, seed=0x5EED0000, D0=2.0, gap_um=5.0)


def c1_labels():
    return G.planes(32, 32, 52)


def c1_jobs(Geometry, nwalk=C1["nwalk"]):
    geom = Geometry(c1_labels(), RES, [0.0, C1["D0"]], T2=[math.inf, math.inf], periodic="xy")
    return [dict(geom=geom, nwalk=int(nwalk), seed=C1["seed"])]


# ---------------------------------------------------------------- c2
C2 = dict(TN=100_000, timepoints=1000, nwalk=1_000_000, seed=0x5EED0001, D0=2.0)


def c2_labels():
    return G.axon(n=64, nz=1000, res=RES, r0=0.8, cv=0.2, undulation=0.5, seed=1235)


def c2_jobs(Geometry, nwalk=C2["nwalk"], periodic="xyz", labels=None):
    lab = c2_labels() if labels is None else labels
    geom = Geometry(lab, RES, [0.0, C2["D0"]], T2=[math.inf, math.inf], periodic=periodic)
    return [dict(geom=geom, nwalk=int(nwalk), seed=C2["seed"], grad=pgse_table_c2())]


# ---------------------------------------------------------------- c3
C3 = dict(TN=50_000, timepoints=1000, nwalk=10_000_000, naxon=256)


def c3_labels():
    return G.bundle(C3["naxon"], mean_r=0.6, cv_r=0.4, len_lo=20.0, len_hi=200.0, res=RES, seed=1236)


def c3_jobs(Geometry, total_walkers=C3["nwalk"], labels=None, which=None):
    """Walkers in proportion to each geometry's walkable voxels (at least 64);
    `which` selects geometries by index (all 256 by default) with the walker
    counts they have in the full bundle."""
    labs = c3_labels() if labels is None else labels
    kap = np.random.default_rng(2024).uniform(0.0, 0.02, len(labs))
    walkable = np.array([np.count_nonzero(l) for l in labs], float)
    nw = np.maximum(64, np.round(total_walkers * walkable / walkable.sum())).astype(np.int64)
    grad = pgse_table_c2()
    idx = range(len(labs)) if which is None else which
    return [dict(geom=Geometry(labs[i], RES, [0.0, 2.0, 1.5], T2=[math.inf, 80.0, 60.0], kappa=float(kap[i])),
                 nwalk=int(nw[i]), seed=0x5EED0002 | (int(i) << 32), grad=grad) for i in idx]


# ---------------------------------------------------------------- c4
C4 = dict(TN=50_000, timepoints=20, nwalk=1_000_000, seed=0x5EED0003, kappa=0.01, ndir=64,
          bvals=(0.5, 1.0, 2.0, 3.0))


def c4_labels():
    return G.axon(n=64, nz=1000, res=RES, r0=0.8, cv=0.2, undulation=0.5, seed=1235, ecs=True)


def c4_geometry(Geometry, D=(0.0, 2.0, 1.5), T2=(math.inf, 80.0, 60.0), kappa=C4["kappa"], labels=None,
                periodic="xyz"):
    """The c2 lattice (periodic edges) with label 2 outside the axon."""
    lab = c4_labels() if labels is None else labels
    return Geometry(lab, RES, list(D), T2=list(T2), kappa=kappa, periodic=periodic)


def c4_jobs(Geometry, nwalk=C4["nwalk"], labels=None, **phys):
    bval, bvec = narrow_table(C4["ndir"], C4["bvals"])
    return [dict(geom=c4_geometry(Geometry, labels=labels, **phys), nwalk=int(nwalk), seed=C4["seed"], bval=bval,
                 bvec=bvec)]


def c4_cotangent(A, rec, nclass, TD, nbvec):
    """Cotangent of the gradient objective sum_t [D_t(z) + K_t(z) + sum_j S_j(t)]
    with the narrow-pulse S = sigRe / sig0 (readRMS.m:44-48), for record
    totals `rec` (A = the analysis module)."""
    i_re = 1 + nclass + 21
    cot = A.akc_cotangent(rec, nclass, TD, [0.0, 0.0, 1.0])
    cot[:, i_re:i_re + nbvec] += (1.0 / rec[:, 0])[:, None]
    cot[:, 0] += -rec[:, i_re:i_re + nbvec].sum(1) / rec[:, 0] ** 2
    return cot


# ---------------------------------------------------------------- c5
C5 = dict(TN=20_000, timepoints=20, nwalk=40_000_000, ngeom=4096, ndir=16, bvals=(1.0, 2.0))


def c5_draws(ngeom=C5["ngeom"]):
    rng = np.random.default_rng(1238)
    return rng.uniform(0.0, 0.4, ngeom), rng.uniform(0.0, 1.0, ngeom), rng.uniform(0.0, 0.03, ngeom)


def c5_labels(k, cv, und):
    return G.axon(n=70, nz=1000, res=RES, r0=0.8, cv=float(cv), undulation=float(und), seed=1238 + k, ecs=True,
                  dtype=np.uint16)


def c5_jobs(Geometry, which, nwalk_per_geom, ngeom=C5["ngeom"]):
    """Geometries `which` (indices into the 4096 draws), each with its walkers."""
    cv, und, kap = c5_draws(ngeom)
    bval, bvec = narrow_table(C5["ndir"], C5["bvals"])
    out = []
    for k in which:
        lab = c5_labels(k, cv[k], und[k])
        geom = Geometry(lab, RES, [0.0, 2.0, 1.5], T2=[math.inf, 80.0, 60.0], kappa=float(kap[k]))
        out.append(dict(geom=geom, nwalk=int(nwalk_per_geom), seed=0x5EED0004 | (int(k) << 32), bval=bval,
                        bvec=bvec))
    return out
