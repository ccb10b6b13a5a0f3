"""The reference's text file contract (no GPU): the setupRMS-style writer,
the CLI's parser (rms_io.cpp through a small host harness), the output
formatting of main.cu:549-615 and the readRMS readers."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLI = os.path.join(ROOT, "monte-carlo-simulation-3d-rms_amd", "cli")


def _X(nbvec=3):
    rng = np.random.default_rng(0)
    bvec = rng.normal(size=(nbvec, 3))
    bvec /= np.linalg.norm(bvec, axis=1, keepdims=True)
    return dict(time_step=2e-4, step_num=500000, particle_num=1000, comp_num=3, voxel_size=0.1,
                diffusivity=[0.0, 2.0, 2.0 / 15], T2_relaxation=[80, 80, 20], bval=[0.2, 1.0, 2.5][:nbvec], bvec=bvec)


def test_setup_writer_format(rms, tmp_path):
    from rms_amd import fileio
    fiber = np.zeros((3, 2, 4), np.uint8)
    fiber[1, 1, 2] = 2
    fiber[2, 0, 3] = 1
    ok = fileio.input_rms(str(tmp_path), fiber, _X())
    assert ok
    txt = open(tmp_path / "fiber.txt").read()
    assert txt == " ".join(map(str, fiber.ravel())) + " "
    assert "\n" not in txt
    assert open(tmp_path / "simParamInput.txt").read() == \
        "2.000000e-04\n500000\n1000\n3\n3\n1.000000e-01\n3\n2\n4\n"
    assert open(tmp_path / "diffusivity.txt").read() == "0.00000000\n2.00000000\n0.13333333\n"
    assert len(open(tmp_path / "bvec.txt").read().split("\n")) == 3 * 3 + 1


def test_cli_parser_reads_writer_output(rms, harness, tmp_path):
    from rms_amd import fileio
    rng = np.random.default_rng(1)
    fiber = rng.integers(0, 3, size=(5, 4, 7)).astype(np.uint8)
    fileio.input_rms(str(tmp_path), fiber, _X())
    out = subprocess.run([harness, "read", str(tmp_path), "0"], capture_output=True, text=True, check=True).stdout
    l1, l2, l3 = out.strip().split("\n")
    f = l1.split()
    assert float(f[0]) == 2e-4 and f[1:5] == ["500000", "1000", "3", "3"] and float(f[5]) == 0.1
    assert f[6:] == ["5", "4", "7"]
    chk = 0
    for v in fiber.ravel():
        chk = (chk * 1000003 + int(v)) % (1 << 64)
    mx, c = l2.split()
    assert int(mx) == fiber.max() and int(c) == chk
    np.testing.assert_allclose([float(x) for x in l3.split()], [0.0, 2.0, 0.13333333])
    # PGSE variant: 8-field parameter file
    X = _X()
    X["grad"] = fileio.pgse_rows(20.0, 5.0, [1.0, 2.0], [[0, 0, 1], [1, 0, 0]])
    fileio.input_rms_pgse(str(tmp_path), fiber, X)
    out = subprocess.run([harness, "read", str(tmp_path), "1"], capture_output=True, text=True, check=True).stdout
    f = out.split("\n")[0].split()
    assert f[1:5] == ["500000", "1000", "0", "3"]
    g = np.loadtxt(tmp_path / "gradient_time.txt").reshape(-1, 6)
    np.testing.assert_array_equal(g, X["grad"])
    b = g[:, 2] ** 2 * g[:, 1] ** 2 * (g[:, 0] - g[:, 1] / 3)
    np.testing.assert_allclose(b, [1.0, 2.0], rtol=1e-14)


def test_cli_parser_rejects_malformed(harness, tmp_path):
    (tmp_path / "simParamInput.txt").write_text("2e-4\n1000\n")
    r = subprocess.run([harness, "read", str(tmp_path), "0"], capture_output=True, text=True)
    assert r.returncode == 1 and "expected 9 fields" in r.stderr


def test_output_formatting_matches_cpp_streams(harness, tmp_path):
    """precision(15) default float format == printf %.15g; default stream
    precision == %g (NPar_count, sim_para, diff_time; main.cu:555-615)."""
    subprocess.run([harness, "write", str(tmp_path)], check=True)
    a = [1.0, 0.1, 1.0 / 3.0, 123456789.123456789, 1e-300, -2.5e-7, 1e6, 4096.0, 0.0]
    want15 = "".join("\t".join("%.15g" % x for x in a[3 * i:3 * i + 3]) + "\n" for i in range(3))
    want6 = "".join("\t".join("%g" % x for x in a[3 * i:3 * i + 3]) + "\n" for i in range(3))
    assert open(tmp_path / "m15.txt").read() == want15
    assert open(tmp_path / "m6.txt").read() == want6
    assert open(tmp_path / "sim_para.txt").read() == "0.0002\n100000\n1000000\n0.1\n"


def test_readrms_normalisation(rms, tmp_path):
    """readRMS.m:115-120: dx2t = dx2 .* D_cnt ./ sig0 etc."""
    from rms_amd import analysis as A
    from rms_amd import fileio
    T = 4
    sig0 = np.array([10.0, 9.0, 8.0, 7.0])
    dx2 = np.arange(T * 6, dtype=float).reshape(T, 6) + 1
    dx4 = np.arange(T * 15, dtype=float).reshape(T, 15) + 1
    sigRe = np.ones((T, 2)) * -3
    for name, arr in (("sig0", sig0[:, None]), ("dx2", dx2), ("dx4", dx4), ("sigRe", sigRe),
                      ("NPar_count", np.ones((T, 2)) * 5)):
        np.savetxt(tmp_path / f"{name}.txt", arr, delimiter="\t", fmt="%.15g")
    (tmp_path / "sim_para.txt").write_text("0.0002\n4000\n10\n0.1\n")
    np.savetxt(tmp_path / "diff_time.txt", (np.arange(T) * 1000 + 1) * 2e-4, fmt="%g")
    np.savetxt(tmp_path / "diffusivity.txt", [0, 2])
    np.savetxt(tmp_path / "T2.txt", [80, 80])
    np.savetxt(tmp_path / "bval.txt", [1, 2])
    np.savetxt(tmp_path / "bvec.txt", [0, 0, 1, 1, 0, 0])
    r = fileio.ReadRMS(str(tmp_path))
    np.testing.assert_allclose(r.dx2t, dx2 * A.D_CNT / sig0[:, None])
    np.testing.assert_allclose(r.dx4t, dx4 * A.W_CNT / sig0[:, None])
    np.testing.assert_allclose(r.sig, 3 / sig0[:, None] * np.ones((T, 2)))
    K, D = r.akc_mom([0, 0, 1])
    np.testing.assert_allclose(D[:, 0], r.dx2t[:, 5] / 2 / r.TD)
    np.testing.assert_allclose(K[:, 0], r.dx4t[:, 14] / r.dx2t[:, 5] ** 2 - 3)


def test_cli_rejects_unknown_periodic_axis():
    """--periodic takes axes among x, y, z; anything else is a usage error
    (exit 2) before any input is read or a device is opened."""
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "monte-carlo-simulation-3d-rms_amd", "bin", "rms_main")
    if not os.path.exists(exe):
        pytest.skip("CLI not built")
    r = subprocess.run([exe, "--periodic", "xq"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--periodic AXES" in r.stderr
