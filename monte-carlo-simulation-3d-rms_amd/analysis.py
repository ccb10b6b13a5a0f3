"""Cumulant analysis of the record sums (the semantics of analysis/readRMS.m).

readRMS.m:13-18 define the multiplicities of the 6 second-order and 15
fourth-order moment components; readRMS.m:115-117 normalise by sig0;
akc_mom (readRMS.m:168-175) gives D(t) = x2/(2t) and K(t) = x4/x2^2 - 3 along
unit directions n.  part1_demo4_analysis.m:88-90 combines compartments by
volume fraction.  Standard errors come from the batch-means replicas the
engine returns (walker batches b = rho mod R).
"""
from __future__ import annotations

import numpy as np

D_CNT = np.array([1, 2, 2, 1, 2, 1], float)
W_CNT = np.array([1, 4, 4, 6, 12, 6, 4, 12, 12, 4, 1, 4, 6, 4, 1], float)
D_IND = np.array([[0, 0], [0, 1], [0, 2], [1, 1], [1, 2], [2, 2]])
W_IND = np.array([[0, 0, 0, 0], [0, 0, 0, 1], [0, 0, 0, 2], [0, 0, 1, 1], [0, 0, 1, 2],
                  [0, 0, 2, 2], [0, 1, 1, 1], [0, 1, 1, 2], [0, 1, 2, 2], [0, 2, 2, 2],
                  [1, 1, 1, 1], [1, 1, 1, 2], [1, 1, 2, 2], [1, 2, 2, 2], [2, 2, 2, 2]])


def ndir2(n):
    n = np.atleast_2d(np.asarray(n, float))
    return np.prod(n[:, D_IND], axis=2)


def ndir4(n):
    n = np.atleast_2d(np.asarray(n, float))
    return np.prod(n[:, W_IND], axis=2)


def diff_times(dt, TN, timepoints):
    """TD[i] = (i*(TN/timepoints)+1)*dt, main.cu:484-487."""
    Tstep = TN // timepoints
    return (np.arange(timepoints) * Tstep + 1).astype(np.float64) * dt


def moments(sig0, dx2, dx4):
    """dx2t, dx4t of readRMS.m:115-117 (multiplicity-weighted, /sig0)."""
    sig0 = np.asarray(sig0, float)[..., None]
    return np.asarray(dx2) * D_CNT / sig0, np.asarray(dx4) * W_CNT / sig0


def akc_mom(sig0, dx2, dx4, TD, n):
    """[K, D] along directions n (readRMS.m:168-175).  Returns arrays of shape
    (timepoints, ndir)."""
    dx2t, dx4t = moments(sig0, dx2, dx4)
    x2 = dx2t @ ndir2(n).T
    x4 = dx4t @ ndir4(n).T
    D = x2 / 2.0 / np.asarray(TD, float)[:, None]
    K = x4 / x2 ** 2 - 3.0
    return K, D


def akc_cotangent(rec, nclass, TD, n, wD=None, wK=None):
    """Cotangent of L = sum_t wD[t] D_t(n) + wK[t] K_t(n) with respect to the
    record sums rec [timepoints, nval] (chain rule through akc_mom and the
    sig0 normalisation), for the adjoint pass.  D_t = a2/(2 TD sig0),
    K_t = a4 sig0 / a2^2 - 3 with a2 = dx2 . (D_CNT n2), a4 = dx4 . (W_CNT n4)."""
    rec = np.asarray(rec, float)
    T = rec.shape[0]
    TD = np.asarray(TD, float)
    wD = np.ones(T) if wD is None else np.asarray(wD, float)
    wK = np.ones(T) if wK is None else np.asarray(wK, float)
    c2 = D_CNT * ndir2(n)[0]
    c4 = W_CNT * ndir4(n)[0]
    i2 = 1 + nclass
    i4 = i2 + 6
    sig0 = rec[:, 0]
    a2 = rec[:, i2:i2 + 6] @ c2
    a4 = rec[:, i4:i4 + 15] @ c4
    cot = np.zeros_like(rec)
    # D_t
    cot[:, i2:i2 + 6] += (wD / (2.0 * TD * sig0))[:, None] * c2
    cot[:, 0] += -wD * a2 / (2.0 * TD * sig0 ** 2)
    # K_t
    cot[:, i4:i4 + 15] += (wK * sig0 / a2 ** 2)[:, None] * c4
    cot[:, i2:i2 + 6] += (-2.0 * wK * a4 * sig0 / a2 ** 3)[:, None] * c2
    cot[:, 0] += wK * a4 / a2 ** 2
    return cot


def narrow_pulse_signal(sig0, sigRe):
    """sig = |sigRe| / sig0 (readRMS.m:119-120)."""
    return np.abs(np.asarray(sigRe)) / np.asarray(sig0, float)[:, None]


def pgse_bvalues(grad):
    """b = |g|^2 delta^2 (Delta - delta/3) per PGSE row (readRMS_pgse.m:48-53)."""
    g = np.asarray(grad, float).reshape(-1, 6)
    return g[:, 2] ** 2 * g[:, 1] ** 2 * (g[:, 0] - g[:, 1] / 3.0)


def split(rec, nclass, nbvec=0):
    rec = np.asarray(rec)
    return {
        "sig0": rec[..., 0],
        "NPar_count": rec[..., 1:1 + nclass],
        "dx2": rec[..., 1 + nclass:7 + nclass],
        "dx4": rec[..., 7 + nclass:22 + nclass],
        "sigRe": rec[..., 22 + nclass:22 + nclass + nbvec],
    }


def replica_stat(fn, rec_rep, *args):
    """Estimate fn(sums) and its standard error by delete-one jackknife over
    the R batch-means replicas (each replica sums the batches b = rho mod R).
    fn receives the summed record array (timepoints, nval)."""
    rep = np.asarray(rec_rep, float)
    R = rep.shape[0]
    tot = rep.sum(axis=0)
    est = np.asarray(fn(tot, *args))
    loo = np.stack([np.asarray(fn(tot - rep[r], *args)) for r in range(R)])
    se = np.sqrt((R - 1) / R * np.sum((loo - loo.mean(axis=0)) ** 2, axis=0))
    return est, se


def combine_compartments(f, D, K):
    """Volume-fraction combination of per-geometry D(t), K(t)
    (part1_demo4_analysis.m:88-90)."""
    f = np.asarray(f, float)
    D = np.asarray(D, float)
    K = np.asarray(K, float)
    Dm = np.sum(f * D, axis=-1)
    Km = 3 * (np.sum(f * D ** 2, axis=-1) - Dm ** 2) / Dm ** 2 + np.sum(f * D ** 2 * K, axis=-1) / Dm ** 2
    return Dm, Km


def fit_dinf(TD, D, sel=None):
    """D(t) = D_inf + c/sqrt(t) least squares (part1_demo4_analysis.m:84)."""
    TD = np.asarray(TD, float)
    D = np.asarray(D, float)
    if sel is not None:
        TD, D = TD[sel], D[sel]
    A = np.stack([np.ones_like(TD), 1 / np.sqrt(TD)], 1)
    x, *_ = np.linalg.lstsq(A, D, rcond=None)
    return x
