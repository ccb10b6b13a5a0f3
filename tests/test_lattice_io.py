"""Lattice inputs beyond fiber.txt: the reference's fiber.mat struct arrays
(part1_demo3_simulation.m:31-32, 75), the raw fiber.bin side channel the
CLIs accept, and the fixtures cut from the reference's own datasets."""
import importlib
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
FIXTURES = ["fiber_I_000.npz", "fiber_II_000.npz", "fiber_III_000.npz", "fiber_IV_000.npz"]


@pytest.fixture(scope="module")
def fileio(rms):
    return importlib.import_module(rms.__name__ + ".fileio")


def test_fiber_mat_struct_array_round_trip(fileio, tmp_path):
    from scipy.io import savemat
    rng = np.random.default_rng(0)
    shapes = [rng.integers(0, 3, (5, 6, 7)).astype(np.uint8), rng.integers(0, 2, (4, 4, 9)).astype(np.uint8)]
    rec = np.zeros((1, 2), dtype=[("shape", object), ("voxel_size", object)])
    for k, sh in enumerate(shapes):
        rec[0, k]["shape"] = sh
        rec[0, k]["voxel_size"] = np.array([[0.1 * (k + 1)]])
    path = str(tmp_path / "fiber.mat")
    savemat(path, {"fiber": rec})
    got = fileio.load_fiber_mat(path)
    assert len(got) == 2
    for (lab, vs), sh, k in zip(got, shapes, range(2)):
        np.testing.assert_array_equal(lab, sh)  # MATLAB (i,j,k) -> labels[i,j,k]
        assert vs == pytest.approx(0.1 * (k + 1))


def test_crop_keeps_content_and_margin(fileio):
    lab = np.zeros((20, 30, 8), np.uint8)
    lab[5:9, 10:14, :] = 1
    lab[6, 12, 3] = 2
    c = fileio.crop_to_content(lab, margin=2)
    assert c.shape == (8, 8, 8)
    assert np.array_equal(np.bincount(c.ravel(), minlength=3), np.bincount(lab.ravel(), minlength=3) - [lab.size - c.size, 0, 0])
    assert np.all(c[:2] == 0) and np.all(c[-2:] == 0) and np.all(c[:, :2] == 0) and np.all(c[:, -2:] == 0)


@pytest.mark.parametrize("name", FIXTURES)
def test_reference_fixtures(name):
    z = np.load(os.path.join(GOLD, name))
    lab = z["labels"]
    assert lab.dtype == np.uint8 and lab.ndim == 3 and lab.shape[2] == 180
    assert float(z["voxel_size"]) == pytest.approx(0.1)
    assert set(np.unique(lab).tolist()) <= {0, 1, 2}
    # cropped with a 2-voxel null margin in x and y
    assert np.all(lab[:2] == 0) and np.all(lab[-2:] == 0) and np.all(lab[:, :2] == 0) and np.all(lab[:, -2:] == 0)
    assert (lab == 1).sum() > 1000


def test_btable_fixture():
    z = np.load(os.path.join(GOLD, "btable_5shell_90ndir_b1000.npz"))
    bval, bvec = z["bval"], z["bvec"]
    assert bval.shape == (90,) and bvec.shape == (90, 3)
    np.testing.assert_allclose(np.unique(np.round(bval, 6)), [0.2, 0.4, 0.6, 0.8, 1.0])
    np.testing.assert_allclose(np.linalg.norm(bvec, axis=1), 1.0, atol=1e-6)


def test_fiber_bin_equals_fiber_txt(fileio, tmp_path):
    """The CLI parser reads fiber.bin (uint8 and uint16) to the same labels as
    fiber.txt."""
    from test_file_contract import _X
    lab = np.load(os.path.join(GOLD, "fiber_I_000.npz"))["labels"]
    d = tmp_path / "s"
    d.mkdir()
    fileio.input_rms(str(d), lab, _X())
    fileio.write_fiber_bin(str(d / "fiber.bin"), lab)
    assert os.path.getsize(d / "fiber.bin") == lab.size
    wide = lab.astype(np.uint16) * 300
    d2 = tmp_path / "w"
    d2.mkdir()
    fileio.write_fiber_bin(str(d2 / "fiber.bin"), wide)
    assert os.path.getsize(d2 / "fiber.bin") == 2 * lab.size
    np.testing.assert_array_equal(np.fromfile(d2 / "fiber.bin", "<u2").reshape(lab.shape), wide)


def test_cli_reader_fiber_bin(harness, fileio, tmp_path):
    from test_file_contract import _X
    lab = np.load(os.path.join(GOLD, "fiber_IV_000.npz"))["labels"]
    d = tmp_path / "s"
    d.mkdir()
    fileio.input_rms(str(d), lab, _X())
    fileio.write_fiber_bin(str(d / "fiber.bin"), lab)
    txt = subprocess.run([harness, "read", str(d), "0"], capture_output=True, text=True, check=True).stdout
    binr = subprocess.run([harness, "read", str(d), "0", "bin"], capture_output=True, text=True, check=True).stdout
    assert txt == binr
    with open(d / "fiber.bin", "ab") as f:
        f.write(b"\x00")
    r = subprocess.run([harness, "read", str(d), "0", "bin"], capture_output=True, text=True)
    assert r.returncode != 0 and "neither" in r.stderr
