"""The engine on the reference's own data: axons cut from the shipped
fiber.mat datasets with the demo's physics (part1_demo3_simulation.m:48-66:
3 compartments, D = [0 2 2/15] with mitochondria, T2 = [80 80 20]) and the
shipped 5-shell b-table, bit-exact against the oracle; and the drop-in
executables end to end (setup directory in, readRMS-parsed files out)."""
import importlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _fixture(name):
    z = np.load(os.path.join(GOLD, name))
    return z["labels"], float(z["voxel_size"])


def _btable():
    z = np.load(os.path.join(GOLD, "btable_5shell_90ndir_b1000.npz"))
    return z["bval"], z["bvec"]


@pytest.mark.parametrize("name,prec", [("fiber_I_000.npz", "f64"), ("fiber_I_000.npz", "f32"),
                                       ("fiber_IV_000.npz", "f32")])
def test_reference_axon_bitexact(rms, orc, engine, name, prec):
    print(f"case reference axon {name} {prec}", flush=True)
    lab, res = _fixture(name)
    bval, bvec = _btable()
    mito = name.startswith("fiber_I_")
    D = [0.0, 2.0, 2.0 / 15] if mito else [0.0, 2.0, 2.0]
    T2 = [80.0, 80.0, 20.0] if mito else [80.0, 80.0, 80.0]
    job = dict(nwalk=1000, seed=2024, bval=bval, bvec=bvec)
    gg = rms.Geometry(lab, res, D, T2=T2)
    og = orc.Geometry(lab, res, D, T2=T2)
    dt, TN, T = 2e-4, 1500, 30
    gpu = engine.run([dict(geom=gg, **job)], dt, TN, T, precision=prec, replicas=8, final_state=True)[0]
    ref = orc.run([dict(geom=og, **job)], dt, TN, T, prec=prec, replicas=8, final_state=True)[0]
    for k in ("rec", "rec_rep", "counters", "fin_g", "fin_o", "fin_c"):
        np.testing.assert_array_equal(gpu[k], ref[k], err_msg=k)
    if mito:
        assert gpu["counters"][4] > 0  # cytoplasm -> mitochondria draws happened
    # every walker is in a non-null compartment at every record
    np.testing.assert_array_equal(gpu["rec"][:, 1:4].sum(1), 1000.0)
    assert np.all(gpu["rec"][:, 1] == 0.0)


def _setup(tmp_path, fileio, lab, res, pgse=False, nwalk=640, TN=2000):
    bval, bvec = _btable()
    X = dict(time_step=2e-4, step_num=TN, particle_num=nwalk, comp_num=3, voxel_size=res,
             diffusivity=[0.0, 2.0, 2.0 / 15], T2_relaxation=[80.0, 80.0, 20.0], bval=bval, bvec=bvec)
    d = str(tmp_path / ("pgse" if pgse else "np"))
    if pgse:
        X["grad"] = fileio.pgse_rows(6.0, 2.0, bval[::9], bvec[::9])
        fileio.input_rms_pgse(d, lab, X)
    else:
        fileio.input_rms(d, lab, X)
    return d


def _engine_jobs(rms, d, lab, pgse, periodic=""):
    """The inputs exactly as the executables read them back from text."""
    D = np.loadtxt(os.path.join(d, "diffusivity.txt")).ravel()
    T2 = np.loadtxt(os.path.join(d, "T2.txt")).ravel()
    p = np.loadtxt(os.path.join(d, "simParamInput.txt")).ravel()
    res = p[4] if pgse else p[5]
    g = rms.Geometry(lab, res, list(D), T2=list(T2), periodic=periodic)
    job = dict(geom=g, nwalk=int(p[2]), seed=77)
    if pgse:
        job["grad"] = np.loadtxt(os.path.join(d, "gradient_time.txt")).ravel().reshape(-1, 6)
    else:
        job["bval"] = np.loadtxt(os.path.join(d, "bval.txt")).ravel()
        job["bvec"] = np.loadtxt(os.path.join(d, "bvec.txt")).ravel().reshape(-1, 3)
    return job, float(p[0]), int(p[1])


@pytest.mark.parametrize("lattice", ["txt", "bin"])
def test_cli_narrow_pulse_end_to_end(rms, engine, tmp_path, lattice):
    """rms_main in a setupRMS directory: stdout lines, output files parsed by
    the readRMS restatement, equal to the engine run on the same inputs (the
    CLI sums walkers in batch order, replicas = 1)."""
    print(f"case cli narrow pulse {lattice}", flush=True)
    fileio = importlib.import_module(rms.__name__ + ".fileio")
    lab, res = _fixture("fiber_I_000.npz")
    d = _setup(tmp_path, fileio, lab, res)
    if lattice == "bin":
        fileio.write_fiber_bin(os.path.join(d, "fiber.bin"), lab)
        os.remove(os.path.join(d, "fiber.txt"))
    out = fileio.run_setup(d, seed=77, timeout=120)
    lines = out.strip().splitlines()
    assert [l.split("=")[0] for l in lines[:3]] == ["diffusivity"] * 3
    assert [l.split("=")[0] for l in lines[3:12]] == ["permeation probability"] * 9
    assert lines[-1].startswith("Done! Elpased time ") and lines[-1].endswith(" s")
    r = fileio.ReadRMS(d)
    A = importlib.import_module(rms.__name__ + ".analysis")
    job, dt, TN = _engine_jobs(rms, d, lab, pgse=False)
    e = engine.run([job], dt, TN, 1000, precision="f64", replicas=1)[0]
    np.testing.assert_allclose(r.sig0, e["sig0"], rtol=1e-14)
    np.testing.assert_allclose(r.dx2t, e["dx2"] * A.D_CNT / e["sig0"][:, None], rtol=1e-13,
                               atol=1e-300)
    np.testing.assert_allclose(r.sig, np.abs(e["sigRe"]) / e["sig0"][:, None], rtol=1e-13, atol=1e-15)
    np.testing.assert_array_equal(r.NPar_count, e["NPar_count"])
    assert r.NPar_count.sum(1).tolist() == [640.0] * 1000
    np.testing.assert_allclose(r.TD, A.diff_times(dt, TN, 1000), rtol=1e-6)


@pytest.mark.parametrize("periodic", ["", "xyz"])
def test_cli_pgse_end_to_end(rms, engine, tmp_path, periodic):
    print(f"case cli pgse periodic={periodic}", flush=True)
    fileio = importlib.import_module(rms.__name__ + ".fileio")
    lab, res = _fixture("fiber_IV_000.npz")
    d = _setup(tmp_path, fileio, lab, res, pgse=True)
    fileio.run_setup(d, pgse=True, seed=77, timeout=120, periodic=periodic)
    r = fileio.ReadRMSPGSE(d)
    job, dt, TN = _engine_jobs(rms, d, lab, pgse=True, periodic=periodic)
    e = engine.run([job], dt, TN, 1000, precision="f64", replicas=1)[0]
    np.testing.assert_allclose(r.sig, np.abs(e["sig"]), rtol=1e-14)
    np.testing.assert_allclose(r.sig0, e["sig0"], rtol=1e-14)
    bval, _ = _btable()
    np.testing.assert_allclose(r.bval, bval[::9], rtol=1e-6)
