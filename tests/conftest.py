import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librms_amd.so)")


@pytest.fixture(scope="session")
def rms():
    import __graft_entry__ as ge
    return ge.load_package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def engine(rms):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    eng = rms.Engine(0)
    yield eng
    eng.close()


@pytest.fixture(scope="session")
def harness(tmp_path_factory):
    """Host build of the CLI's file-contract code (rms_io.cpp) with a small driver."""
    import subprocess
    d = tmp_path_factory.mktemp("h")
    exe = str(d / "io_harness")
    here = os.path.dirname(os.path.abspath(__file__))
    cli = os.path.join(ROOT, "monte-carlo-simulation-3d-rms_amd", "cli")
    subprocess.run(["g++", "-O1", "-std=c++17", os.path.join(here, "cpp", "io_harness.cpp"),
                    os.path.join(cli, "rms_io.cpp"), "-o", exe], check=True)
    return exe
