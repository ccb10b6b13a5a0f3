"""Build recipe for librms_amd.so and the drop-in CLIs (gfx950).

hipcc compiles the kernels and the C ABI into one shared library; the two
command-line programs that honour the reference file contract
(hpc_code/lib/main.cu and main_PGSE.cu) are plain C++ clients of that ABI.
Outputs stay in-tree: lib/librms_amd.so, bin/rms_main, bin/rms_main_pgse,
and bin/rms_math_check (exhaustive checks of the kernel's FP32 math shortcuts,
a GPU program the tests run).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
CLI = os.path.join(PKG, "cli")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIB_DIR, "librms_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RMS_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every a*b+c the engine evaluates is two roundings unless
# written as fma(), the order the host restatement reproduces bit for bit.
# -packed-fp32-ops off: no v_pk_* FP32 arithmetic is generated.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_lib(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-o", tmp,
           os.path.join(CSRC, "rms_kernels.hip"),
           os.path.join(CSRC, "rms_api.cpp"),
           os.path.join(CSRC, "rms_host.cpp")]
    out = _run(cmd)
    os.replace(tmp, LIB)
    if verbose and out:
        print(out)
    return LIB


def build_cli(verbose=False):
    os.makedirs(BIN_DIR, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    outs = []
    for name, src in (("rms_main", "rms_main.cpp"), ("rms_main_pgse", "rms_main_pgse.cpp")):
        exe = os.path.join(BIN_DIR, name)
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(CLI, src), os.path.join(CLI, "rms_cli.cpp"), os.path.join(CLI, "rms_io.cpp"),
               "-L", LIB_DIR, "-lrms_amd", "-Wl,-rpath,$ORIGIN/../lib", "-o", exe + ".tmp"]
        _run(cmd)
        os.replace(exe + ".tmp", exe)
        outs.append(exe)
    return outs


def build_math_check(verbose=False):
    os.makedirs(BIN_DIR, exist_ok=True)
    exe = os.path.join(BIN_DIR, "rms_math_check")
    flags = [f for f in HIP_FLAGS if f != "-fPIC"]
    out = _run([HIPCC, *flags, "-o", exe + ".tmp", os.path.join(CSRC, "rms_math_check.hip")])
    os.replace(exe + ".tmp", exe)
    if verbose and out:
        print(out)
    return exe


def build(verbose=False):
    lib = build_lib(verbose)
    clis = build_cli(verbose)
    chk = build_math_check(verbose)
    return [lib, *clis, chk]


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
