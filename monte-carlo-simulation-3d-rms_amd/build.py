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


# The tangent variants with the boundary terms of the D derivative hold three
# walkers' worth of state and spill heavily.  The greedy VGPR allocator of
# this toolchain (ROCm 7.2, clang 22) has placed a spill store of a value live
# in every lane at the start of a reconvergence block, before the exec-mask
# restore, so that lanes outside the branch never stored it and reloaded
# garbage (a sin/cos polynomial coefficient, in walk_kernel<float, 21, 3>).
# The basic allocator spills at definitions and reloads at uses only, which
# is correct under any exec mask; that translation unit is built with it, and
# the device assembly of every unit is scanned for the signature.
JUMP_FLAGS = ["-mllvm", "-vgpr-regalloc=basic"]
UNITS = (("rms_kernels.hip", []), ("rms_kernels_jump.hip", JUMP_FLAGS), ("rms_kernels_tangent.hip", []),
         ("rms_kernels_bricks.hip", []), ("rms_kernels_wrap.hip", []), ("rms_kernels_nocount.hip", []),
         ("rms_api.cpp", []), ("rms_host.cpp", []))


def spill_before_exec_restore(asm):
    """Blocks whose leading instructions store a VGPR spill before restoring
    exec (the spill then misses the lanes that rejoin there).  Returns
    (function, label) pairs."""
    hits = []
    func = None
    lines = asm.split("\n")
    for k, line in enumerate(lines):
        if line.startswith("_ZN") and line.split(":")[0].isidentifier():
            func = line.split(":")[0]
        if not (line.startswith(".LBB") or line.startswith("; %bb")):
            continue
        spills = 0
        for t in lines[k + 1:k + 200]:
            t = t.strip()
            if not t or t.startswith(";"):
                continue
            if "Folded Spill" in t:
                spills += 1
                continue
            if ("Folded Reload" in t or t.startswith(("s_waitcnt", "s_nop", "v_readlane", "v_writelane"))
                    or (t.startswith("s_") and "exec" not in t and not t.startswith(("s_cbranch", "s_branch")))):
                continue
            if t.startswith("s_or_b64 exec, exec") and spills:
                hits.append((func, line.split(":")[0]))
            break
    return hits


def build_lib(verbose=False):
    os.makedirs(LIB_DIR, exist_ok=True)
    tmpdir = os.path.join(LIB_DIR, ".obj")
    os.makedirs(tmpdir, exist_ok=True)
    procs, objs, asms = [], [], []
    for src, extra in UNITS:
        obj = os.path.join(tmpdir, src + ".o")
        cmd = [HIPCC, *HIP_FLAGS, *extra, "-c", "-o", obj, os.path.join(CSRC, src)]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
        if src.endswith(".hip"):
            asm = os.path.join(tmpdir, src + ".s")
            cmd = [HIPCC, *HIP_FLAGS, *extra, "--cuda-device-only", "-S", "-o", asm, os.path.join(CSRC, src)]
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
            asms.append(asm)
    out = []
    for cmd, pr in procs:
        txt = pr.communicate()[0]
        txt = "\n".join(l for l in txt.splitlines() if "packed-fp32-ops" not in l and l.strip())
        if pr.returncode != 0:
            raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + txt)
        if txt:
            out.append(txt)
    for asm in asms:
        hits = spill_before_exec_restore(open(asm).read())
        if hits:
            raise RuntimeError(f"{os.path.basename(asm)}: VGPR spill stored before an exec restore in "
                               + ", ".join(f"{f} {b}" for f, b in hits[:8]))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    shutil.rmtree(tmpdir, ignore_errors=True)
    if verbose and out:
        print("\n".join(out))
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
