"""The reference's on-disk contract, from Python.

Writers
  input_rms(target, fiber, X)        analysis/setupRMS.m:16-69 (narrow-pulse
                                     inputs: fiber.txt, bval.txt, bvec.txt,
                                     simParamInput.txt, diffusivity.txt, T2.txt)
  input_rms_pgse(target, fiber, X)   the inputs hpc_code/lib/main_PGSE.cu reads
                                     (8-field simParamInput.txt,
                                     gradient_Ngrad.txt, gradient_time.txt), for
                                     which the reference has no writer
Readers
  ReadRMS(root)                      analysis/readRMS.m:28-49 (+ akc_mom)
  ReadRMSPGSE(root)                  analysis/readRMS_pgse.m:29-58
Runner
  run_setup(target, pgse=False, ...) runs the drop-in program in a setup
                                     directory (part1_demo3_simulation.m:85-89)
"""
from __future__ import annotations

import os
import subprocess

import numpy as np

from . import analysis as A

PKG = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(PKG, "bin")


def _ml_u(v):
    return "%u" % int(v)


def _ml_e(v):
    """MATLAB %e: 6 digits, two-digit exponent at least (C printf agrees)."""
    return "%e" % float(v)


def _write(path, text):
    with open(path, "w") as f:
        f.write(text)


def write_fiber(path, fiber):
    """fiber.txt: '%u ' per voxel; i (x) outer, j (y), k (z) fastest, no
    newlines (setupRMS.m:17-23)."""
    fiber = np.asarray(fiber)
    if fiber.ndim != 3:
        raise ValueError("fiber must be 3-D")
    flat = np.ascontiguousarray(fiber).ravel()
    with open(path, "w") as f:
        chunk = 1 << 20
        for s in range(0, flat.size, chunk):
            part = flat[s:s + chunk]
            f.write(" ".join(map(str, part.tolist())) + " ")


def write_fiber_bin(path, fiber):
    """fiber.bin: the lattice as raw little-endian uint8 (labels < 256) or
    uint16, x-major / z fastest (the fiber.txt order).  The CLIs read it in
    place of fiber.txt when present; text is impractical above ~1e8 voxels."""
    fiber = np.asarray(fiber)
    if fiber.ndim != 3:
        raise ValueError("fiber must be 3-D")
    if fiber.size and (fiber.min() < 0 or fiber.max() > 65535):
        raise ValueError("labels must be in 0..65535")
    dt = "<u1" if (fiber.size == 0 or fiber.max() < 256) else "<u2"
    np.ascontiguousarray(fiber, dtype=dt).tofile(path)


def load_fiber_mat(path):
    """Read a fiber.mat of the reference (a MAT v5 struct array `fiber` with
    fields `shape` (uint8 label lattice) and `voxel_size`, as loaded by
    part1_demo3_simulation.m:31-32 and 57, 75).  Returns a list of
    (labels, voxel_size).  MATLAB's (i, j, k) index maps to labels[i, j, k],
    the order setupRMS.m:17-23 writes fiber.txt in."""
    try:
        from scipy.io import loadmat
    except ImportError as e:  # pragma: no cover - scipy is part of the image
        raise RuntimeError("reading .mat files needs scipy") from e
    m = loadmat(path, squeeze_me=False, struct_as_record=True)
    if "fiber" not in m:
        raise ValueError(f"{path}: no variable 'fiber'")
    f = m["fiber"].ravel()
    out = []
    for e in f:
        shape = np.ascontiguousarray(np.asarray(e["shape"]))
        if shape.ndim != 3:
            raise ValueError("fiber(k).shape must be a 3-D array")
        vs = float(np.asarray(e["voxel_size"]).ravel()[0])
        out.append((shape, vs))
    return out


def crop_to_content(labels, margin=2):
    """Crop x and y to the bounding box of the non-null voxels plus `margin`
    (full z).  The lattice edges are mirrors, so empty null margins beyond
    one voxel do not change the walk."""
    labels = np.asarray(labels)
    nz = np.nonzero(labels)
    if nz[0].size == 0:
        return labels
    lo = [max(0, int(a.min()) - margin) for a in nz[:2]]
    hi = [min(labels.shape[k], int(nz[k].max()) + margin + 1) for k in range(2)]
    return np.ascontiguousarray(labels[lo[0]:hi[0], lo[1]:hi[1], :])


def input_rms(target, fiber, X):
    """setupRMS.inputRMS: X has time_step, step_num, particle_num, comp_num,
    voxel_size, diffusivity, T2_relaxation, bval, bvec."""
    os.makedirs(target, exist_ok=True)
    write_fiber(os.path.join(target, "fiber.txt"), fiber)
    bval = np.atleast_1d(np.asarray(X["bval"], float))
    bvec = np.asarray(X["bvec"], float).reshape(-1, 3)
    _write(os.path.join(target, "bval.txt"), "".join("%.8f\n" % v for v in bval))
    _write(os.path.join(target, "bvec.txt"), "".join("%.8f\n" % v for row in bvec for v in row))
    nx, ny, nz = np.asarray(fiber).shape
    params = [_ml_e(X["time_step"]), _ml_u(X["step_num"]), _ml_u(X["particle_num"]), _ml_u(bvec.shape[0]),
              _ml_u(X["comp_num"]), _ml_e(X["voxel_size"]), _ml_u(nx), _ml_u(ny), _ml_u(nz)]
    _write(os.path.join(target, "simParamInput.txt"), "".join(p + "\n" for p in params))
    _write(os.path.join(target, "diffusivity.txt"), "".join("%.8f\n" % v for v in np.ravel(X["diffusivity"])))
    _write(os.path.join(target, "T2.txt"), "".join("%.8f\n" % v for v in np.ravel(X["T2_relaxation"])))
    step_max = np.sqrt(6 * np.max(X["diffusivity"]) * X["time_step"])
    return bool(step_max < X["voxel_size"])


def input_rms_pgse(target, fiber, X):
    """Inputs of main_PGSE.cu:440-506: the 8-field simParamInput.txt (no
    Nbvec), gradient_Ngrad.txt and gradient_time.txt (6 values per row:
    Delta, delta, |g|, gx, gy, gz).  X has the fields of input_rms except
    bval/bvec, plus `grad` (Ngrad x 6)."""
    os.makedirs(target, exist_ok=True)
    write_fiber(os.path.join(target, "fiber.txt"), fiber)
    grad = np.asarray(X["grad"], float).reshape(-1, 6)
    nx, ny, nz = np.asarray(fiber).shape
    params = [_ml_e(X["time_step"]), _ml_u(X["step_num"]), _ml_u(X["particle_num"]), _ml_u(X["comp_num"]),
              _ml_e(X["voxel_size"]), _ml_u(nx), _ml_u(ny), _ml_u(nz)]
    _write(os.path.join(target, "simParamInput.txt"), "".join(p + "\n" for p in params))
    _write(os.path.join(target, "diffusivity.txt"), "".join("%.8f\n" % v for v in np.ravel(X["diffusivity"])))
    _write(os.path.join(target, "T2.txt"), "".join("%.8f\n" % v for v in np.ravel(X["T2_relaxation"])))
    _write(os.path.join(target, "gradient_Ngrad.txt"), "%u\n" % grad.shape[0])
    _write(os.path.join(target, "gradient_time.txt"), "".join("%.17g\n" % v for row in grad for v in row))
    step_max = np.sqrt(6 * np.max(X["diffusivity"]) * X["time_step"])
    return bool(step_max < X["voxel_size"])


def pgse_rows(Delta, delta, bval, bvec):
    """Gradient rows for b-values and directions: |g| = sqrt(b/(delta^2
    (Delta - delta/3))), the inverse of readRMS_pgse.m:53."""
    rows = []
    for b, n in zip(np.ravel(bval), np.asarray(bvec, float).reshape(-1, 3)):
        n = n / np.linalg.norm(n) if np.linalg.norm(n) > 0 else n
        g = np.sqrt(b / (delta ** 2 * (Delta - delta / 3)))
        rows.append([Delta, delta, g, *n])
    return np.array(rows)


def _load(path):
    """MATLAB load() of a whitespace table."""
    a = np.loadtxt(path, ndmin=2)
    return a


class ReadRMS:
    """readRMS.readJob (readRMS.m:28-49)."""

    def __init__(self, root):
        param = np.loadtxt(os.path.join(root, "sim_para.txt")).ravel()
        self.dtime, self.Tstep, self.NPar, self.voxel_size = param[0], param[1], param[2], param[3]
        self.TD = np.loadtxt(os.path.join(root, "diff_time.txt")).ravel()
        self.D0 = np.loadtxt(os.path.join(root, "diffusivity.txt")).ravel()
        self.T2 = np.loadtxt(os.path.join(root, "T2.txt")).ravel()
        bval_p = os.path.join(root, "bval.txt")
        if os.path.exists(bval_p) and os.path.getsize(bval_p) > 0:
            self.bval = np.loadtxt(bval_p).ravel()
            self.bvec = np.loadtxt(os.path.join(root, "bvec.txt")).ravel().reshape(-1, 3)
        else:
            self.bval = np.zeros(0)
            self.bvec = np.zeros((0, 3))
        self.sig0 = np.loadtxt(os.path.join(root, "sig0.txt")).ravel()
        self.dx2t = _load(os.path.join(root, "dx2.txt")) * A.D_CNT / self.sig0[:, None]
        self.dx4t = _load(os.path.join(root, "dx4.txt")) * A.W_CNT / self.sig0[:, None]
        sp = os.path.join(root, "sigRe.txt")
        if os.path.exists(sp) and os.path.getsize(sp) > 0:
            self.sig = np.abs(_load(sp)) / self.sig0[:, None]
        else:
            self.sig = np.zeros((len(self.sig0), 0))
        self.NPar_count = _load(os.path.join(root, "NPar_count.txt"))

    def akc_mom(self, n):
        """[K, D] along unit directions n (readRMS.m:96-103)."""
        x2 = self.dx2t @ A.ndir2(n).T
        x4 = self.dx4t @ A.ndir4(n).T
        D = x2 / 2.0 / self.TD[:, None]
        K = x4 / x2 ** 2 - 3.0
        return K, D


class ReadRMSPGSE(ReadRMS):
    """readRMS_pgse.readJob (readRMS_pgse.m:29-58): b = g^2 delta^2
    (Delta - delta/3); the PGSE signal is |sig| without normalisation."""

    def __init__(self, root):
        param = np.loadtxt(os.path.join(root, "sim_para.txt")).ravel()
        self.dtime, self.Tstep, self.NPar, self.res = param[0], param[1], param[2], param[3]
        self.TD = np.loadtxt(os.path.join(root, "diff_time.txt")).ravel()
        self.D0 = np.loadtxt(os.path.join(root, "diffusivity.txt")).ravel()
        self.T2 = np.loadtxt(os.path.join(root, "T2.txt")).ravel()
        self.sig0 = np.loadtxt(os.path.join(root, "sig0.txt")).ravel()
        self.dx2t = _load(os.path.join(root, "dx2.txt")) * A.D_CNT / self.sig0[:, None]
        self.dx4t = _load(os.path.join(root, "dx4.txt")) * A.W_CNT / self.sig0[:, None]
        btab = np.loadtxt(os.path.join(root, "gradient_time.txt")).ravel().reshape(-1, 6)
        self.Del = btab[:, 0]
        self.del_ = btab[:, 1]
        g = btab[:, 2]
        self.bval = g ** 2 * self.del_ ** 2 * (self.Del - self.del_ / 3)
        self.bvec = btab[:, 3:6]
        self.sig = np.abs(np.loadtxt(os.path.join(root, "sig.txt")).ravel())
        self.NPar_count = _load(os.path.join(root, "NPar_count.txt"))


def run_setup(target, pgse=False, seed=None, precision=None, timepoints=None, device=None, timeout=None,
              periodic=None):
    """Run rms_main (or rms_main_pgse) with cwd = target, as
    part1_demo3_simulation.m:85-89 runs ./main_cuda.  periodic: axes whose
    lattice edges wrap around (e.g. "xy"; default: the reference's mirrors).
    Returns stdout."""
    exe = os.path.join(BIN, "rms_main_pgse" if pgse else "rms_main")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is missing: build the package first")
    cmd = [exe]
    if seed is not None:
        cmd += ["--seed", str(int(seed))]
    if precision is not None:
        cmd += ["--precision", precision]
    if timepoints is not None:
        cmd += ["--timepoints", str(int(timepoints))]
    if device is not None:
        cmd += ["--device", str(int(device))]
    if periodic:
        cmd += ["--periodic", str(periodic)]
    r = subprocess.run(cmd, cwd=target, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"{os.path.basename(exe)} failed ({r.returncode}): {r.stderr.strip()}")
    return r.stdout
