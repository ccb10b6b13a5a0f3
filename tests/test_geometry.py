"""Synthetic geometry generators (the workloads of bench.py): the beaded,
undulating axon matches the requested caliber CV on the voxelised shape
(the estimator of part1_demo4_analysis.m:47-48), the bundle has uneven
uint16 lattices with an extra-axonal shell."""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def G(rms):
    return importlib.import_module(rms.__name__ + ".geometry")


def test_planes(G):
    lab = G.planes(8, 9, 12)
    assert lab.shape == (8, 9, 12) and lab.dtype == np.uint8
    assert np.all(lab[:, :, 0] == 0) and np.all(lab[:, :, -1] == 0) and np.all(lab[:, :, 1:-1] == 1)


@pytest.mark.parametrize("cv", [0.0, 0.1, 0.2, 0.3])
def test_axon_caliber_cv_matched(G, cv):
    lab = G.axon(n=48, nz=400, res=0.1, r0=0.8, cv=cv, undulation=0.3, seed=7)
    assert lab.shape == (48, 48, 400)
    got = G.measured_cv(lab, 0.1)
    assert got == pytest.approx(cv, abs=0.01 if cv > 0 else 0.02)
    # mean radius near r0 (voxel-area estimate)
    area = np.sum(lab == 1, axis=(0, 1)) * 0.01
    assert np.sqrt(area / np.pi).mean() == pytest.approx(0.8, rel=0.1)
    # z-periodic centreline and radius (wavelengths divide the length): first and last slices agree
    assert abs(int(np.sum(lab[:, :, 0])) - int(np.sum(lab[:, :, -1]))) < 0.2 * np.sum(lab[:, :, 0])


def test_axon_undulation_moves_centreline(G):
    lab = G.axon(n=64, nz=400, res=0.1, r0=0.6, cv=0.0, undulation=1.0, seed=3)
    x = np.arange(64) + 0.5 - 32
    cx = [np.sum(x[:, None] * (lab[:, :, k] == 1)) / np.sum(lab[:, :, k] == 1) for k in range(400)]
    amp = (np.max(cx) - np.min(cx)) / 2 * 0.1
    assert amp == pytest.approx(1.0, abs=0.15)


def test_axon_with_extra_axonal_space(G):
    lab = G.axon(n=32, nz=100, res=0.1, r0=0.8, cv=0.1, undulation=0.2, seed=5, ecs=True)
    assert set(np.unique(lab).tolist()) == {1, 2}


def test_bundle_uneven_uint16_lattices(G):
    labs = G.bundle(naxon=12, mean_r=0.6, cv_r=0.4, len_lo=20.0, len_hi=60.0, res=0.1, seed=11)
    assert len(labs) == 12
    sizes = [l.size for l in labs]
    assert len(set(sizes)) > 6  # uneven
    for l in labs:
        assert l.dtype == np.uint16
        u = set(np.unique(l).tolist())
        assert {1, 2} <= u <= {0, 1, 2}
        assert 200 <= l.shape[2] <= 600
