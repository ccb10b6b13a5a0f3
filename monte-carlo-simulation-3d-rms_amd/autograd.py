"""torch.autograd over the HIP engine.

`simulate` runs one job through `Plan.execute` and returns the record sums
(the sig0 / NPar_count / dx2 / dx4 / sigRe columns of hpc_code/lib/main.cu:
341-378, one row per record) and the PGSE signal (main_PGSE.cu:348-357) as
torch tensors that are differentiable with respect to the per-class
diffusivity D, the per-class T2 and a common membrane permeability kappa.
The backward pass is `Plan.adjoint` (rms_plan_adjoint): the same walk (same
counter-based random stream) with forward-mode tangents, contracted with the
incoming cotangents on the device.

Estimators used by the adjoint pass:
  D      chosen by d_estimator (rms_adjoint_opts): "auto" (the default) is
         the pathwise derivative of the displacement (faces fixed, step
         length d s_c / d D_c = s_c / (2 D_c), reflections folded, plus the
         score of the exchange draws) where the walk's end moves continuously
         with D, and central differences with common random numbers (the
         same seeds at D_c (1 +- h), walking only the walkers whose walk can
         change with D_c) where it jumps: exchange faces, lattice-edge
         mirrors, voxel edges whose face order decides the outcome.
         "boundary" adds the boundary terms of those jumps to the
         pathwise derivative instead (unbiased, slow, high variance; NaN with
         a PGSE table), "fd" differences every job, "pathwise" returns NaN
         where the walk jumps;
  T2     analytic: d/dT2_c exp(-sum t_c/T2_c) = (t_c / T2_c^2) exp(...);
  kappa  score function of the Bernoulli exchange draws,
         d log P / d kappa = 1 / (kappa (1 + pi0)), against a leave-one-out
         baseline (the mean of the other walkers of the 64-walker batch);
         NaN when a pair of non-null classes is impermeable only because of
         kappa (kappa = 0), where no draw happens.
Class 0 (the null compartment) has no gradient.  The inputs are host tensors
(float64); the simulation itself always runs on the engine's device.
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import Geometry


class _Simulate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, D, T2, kappa, spec):
        nc = D.numel()
        kap = None if kappa is None else float(kappa.detach().reshape(()).item())
        geom = Geometry(spec["labels"], spec["res"], D.detach().cpu().double().numpy().tolist(),
                        T2=T2.detach().cpu().double().numpy().tolist(), kappa=kap,
                        label_class=spec.get("label_class"))
        job = dict(geom=geom, nwalk=spec["nwalk"], seed=spec.get("seed", 1), bval=spec.get("bval"),
                   bvec=spec.get("bvec"), grad=spec.get("grad"))
        plan = spec["engine"].plan([job], spec["dt"], spec["TN"], spec.get("timepoints", 1000),
                                   precision=spec.get("precision", "f32"), replicas=spec.get("replicas", 32))
        try:
            out = plan.execute(final_state=False, want_replicas=False)[0]
        except Exception:
            plan.close()
            raise
        ctx.plan = plan
        ctx.nc = nc
        ctx.has_kappa = kappa is not None
        ctx.d_estimator = spec.get("d_estimator", "auto")
        ctx.fd_step = spec.get("fd_step", 0.0)
        ctx.ngrad = 0 if out["sig"] is None else len(out["sig"])
        rec = torch.from_numpy(np.ascontiguousarray(out["rec"]))
        sig = torch.from_numpy(out["sig"]) if ctx.ngrad else torch.zeros(0, dtype=torch.float64)
        return rec, sig

    @staticmethod
    def backward(ctx, g_rec, g_sig):
        plan = ctx.plan
        try:
            cr = None if g_rec is None else g_rec.detach().cpu().double().contiguous().numpy()
            cs = None if (g_sig is None or ctx.ngrad == 0) else g_sig.detach().cpu().double().contiguous().numpy()
            g = plan.adjoint([cr], [cs], estimator=ctx.d_estimator, fd_step=ctx.fd_step)[0]["grad"]
        finally:
            plan.close()
            ctx.plan = None
        nc = ctx.nc
        gD = torch.from_numpy(g[:nc].copy())
        gT = torch.from_numpy(g[4:4 + nc].copy())
        gk = torch.tensor(g[8], dtype=torch.float64) if ctx.has_kappa else None
        return gD, gT, gk, None


def simulate(engine, labels, res, D, T2, nwalk, dt, TN, timepoints=1000, kappa=None, seed=1, bval=None, bvec=None,
             grad=None, precision="f32", replicas=32, label_class=None, d_estimator="auto", fd_step=0.0):
    """Differentiable simulation of one geometry.

    D, T2: float64 tensors of length nclass (T2 = inf for no relaxation);
    kappa: None (the reference rule P = min(1, sqrt(Dj/Di))) or a scalar
    tensor applied to every pair of non-null classes.  Returns (rec, sig):
    rec [timepoints, nval] record sums, sig [ngrad] PGSE signal sums.
    d_estimator / fd_step: the D estimator of the backward pass (above) and
    the relative step of its central differences (0 = 0.05)."""
    D = torch.as_tensor(D, dtype=torch.float64)
    T2 = torch.as_tensor(T2, dtype=torch.float64)
    if kappa is not None:
        kappa = torch.as_tensor(kappa, dtype=torch.float64)
    spec = dict(engine=engine, labels=labels, res=res, nwalk=nwalk, dt=dt, TN=TN, timepoints=timepoints, seed=seed,
                bval=bval, bvec=bvec, grad=grad, precision=precision, replicas=replicas, label_class=label_class,
                d_estimator=d_estimator, fd_step=fd_step)
    return _Simulate.apply(D, T2, kappa, spec)


class _SimulateMany(torch.autograd.Function):
    @staticmethod
    def forward(ctx, D, T2, kappa, spec):
        jobs = []
        for j, js in enumerate(spec["jobs"]):
            kap = None if kappa is None else float(kappa[j].detach().item())
            geom = Geometry(js["labels"], js.get("res", spec["res"]), D[j].detach().cpu().double().numpy().tolist(),
                            T2=T2[j].detach().cpu().double().numpy().tolist(), kappa=kap,
                            label_class=js.get("label_class"), periodic=js.get("periodic", ""),
                            storage=js.get("storage", "auto"))
            jobs.append(dict(geom=geom, nwalk=js["nwalk"], seed=js.get("seed", 1 + j), bval=js.get("bval"),
                             bvec=js.get("bvec"), grad=js.get("grad")))
        plan = spec["engine"].plan(jobs, spec["dt"], spec["TN"], spec.get("timepoints", 1000),
                                   precision=spec.get("precision", "f32"), replicas=spec.get("replicas", 32))
        try:
            outs = plan.execute(final_state=False, want_replicas=False)
        except Exception:
            plan.close()
            raise
        ctx.plan = plan
        ctx.nc = D.shape[1]
        ctx.has_kappa = kappa is not None
        ctx.d_estimator = spec.get("d_estimator", "auto")
        ctx.fd_step = spec.get("fd_step", 0.0)
        ctx.ngrad = 0 if outs[0]["sig"] is None else len(outs[0]["sig"])
        rec = torch.from_numpy(np.stack([np.ascontiguousarray(o["rec"]) for o in outs]))
        sig = (torch.from_numpy(np.stack([o["sig"] for o in outs])) if ctx.ngrad
               else torch.zeros((len(outs), 0), dtype=torch.float64))
        return rec, sig

    @staticmethod
    def backward(ctx, g_rec, g_sig):
        plan = ctx.plan
        try:
            nj = len(plan.jobs)
            cr = [None] * nj if g_rec is None else [g_rec[j].detach().cpu().double().contiguous().numpy()
                                                    for j in range(nj)]
            cs = None if (g_sig is None or ctx.ngrad == 0) else [g_sig[j].detach().cpu().double().contiguous().numpy()
                                                                for j in range(nj)]
            g = np.stack([a["grad"] for a in plan.adjoint(cr, cs, estimator=ctx.d_estimator, fd_step=ctx.fd_step)])
        finally:
            plan.close()
            ctx.plan = None
        nc = ctx.nc
        gD = torch.from_numpy(g[:, :nc].copy())
        gT = torch.from_numpy(g[:, 4:4 + nc].copy())
        gk = torch.from_numpy(g[:, 8].copy()) if ctx.has_kappa else None
        return gD, gT, gk, None


def simulate_many(engine, jobs, D, T2, dt, TN, timepoints=1000, kappa=None, res=0.1, precision="f32", replicas=32,
                  d_estimator="auto", fd_step=0.0):
    """Differentiable simulation of many geometries in one plan (one
    persistent launch forward, one tangent pass backward).

    jobs: list of dicts with "labels" (and optionally "res", "nwalk",
    "seed", "bval", "bvec", "grad", "label_class", "periodic", "storage"),
    every geometry with the same class count and b-table size; D, T2:
    float64 tensors [njobs, nclass]; kappa: None (the reference rule) or a
    tensor [njobs], one permeability per geometry.  Returns (rec, sig): rec
    [njobs, timepoints, nval] record sums, sig [njobs, ngrad] PGSE signal
    sums.  The gradient of each geometry is the one the single-geometry
    `simulate` returns for it."""
    D = torch.as_tensor(D, dtype=torch.float64)
    T2 = torch.as_tensor(T2, dtype=torch.float64)
    if D.ndim != 2 or T2.shape != D.shape or D.shape[0] != len(jobs):
        raise ValueError("D and T2 must be [njobs, nclass]")
    if kappa is not None:
        kappa = torch.as_tensor(kappa, dtype=torch.float64).reshape(-1)
        if kappa.numel() != len(jobs):
            raise ValueError("kappa must have one value per job")
    spec = dict(engine=engine, jobs=jobs, dt=dt, TN=TN, timepoints=timepoints, res=res, precision=precision,
                replicas=replicas, d_estimator=d_estimator, fd_step=fd_step)
    return _SimulateMany.apply(D, T2, kappa, spec)
