"""The C-ABI library loads and exports every entry point include/rms_amd.h
declares (no GPU calls)."""
import os
import re
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _declared():
    txt = open(os.path.join(ROOT, "include", "rms_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rms_[a-z_]+)\s*\(", txt)))


def test_header_symbols_exported(rms):
    lib = rms._capi.lib()
    decl = _declared()
    assert len(decl) >= 12
    for name in decl:
        assert hasattr(lib, name), name
    nm = subprocess.run(["nm", "-D", "--defined-only", rms._capi.LIB_PATH], capture_output=True, text=True,
                        check=True).stdout
    exported = set(re.findall(r" T (rms_\w+)", nm))
    assert set(decl) <= exported
    assert set(rms._capi.EXPORTS) == set(decl)


def test_pure_host_entry_points(rms):
    """rms_nval / rms_derive / rms_version need no device."""
    import ctypes as C
    import numpy as np
    lib = rms._capi.lib()
    assert lib.rms_version().startswith(b"rms_amd")
    assert lib.rms_nval(3, 90) == 1 + 3 + 21 + 90
    g = rms.Geometry(np.ones((2, 2, 2), np.uint8), 0.1, [0.0, 2.0, 0.5], T2=[80, 60, 20])
    d = g.derive(2e-4)
    np.testing.assert_array_equal(d["step"], np.sqrt(6 * 2e-4 * np.array([0, 2.0, 0.5])) / 0.1)
    assert d["P"][1, 2] == np.sqrt(0.25) and d["P"][2, 1] == 1.0
    assert C.sizeof(rms._capi.Geometry) == 4 * 4 + 8 + 8 + 4 + 4 + 8 + 8 * 4 * 2 + 8 * 16 + 4 + 4
    # rms_run_params: dt, TN, three int32 (+4 padding), chunk_steps, five int32 (+4 padding)
    assert C.sizeof(rms._capi.RunParams) == 8 + 8 + 3 * 4 + 4 + 8 + 5 * 4 + 4
    # rms_timing: three doubles, four int32, the int64 walker-step count
    assert C.sizeof(rms._capi.Timing) == 3 * 8 + 4 * 4 + 8
    # rms_adjoint_opts: estimator, fd_all_walkers, relative step
    assert C.sizeof(rms._capi.AdjointOpts) == 4 + 4 + 8


def test_product_path_has_no_cpu_fallback(rms):
    """The Python package never imports the oracle."""
    pkg = os.path.join(ROOT, "monte-carlo-simulation-3d-rms_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.lower() or f == "__init__.py", f


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU; --gpus N > 1 exits with a message (no GPU touched)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
