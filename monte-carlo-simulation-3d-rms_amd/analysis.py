"""Cumulant analysis of the record sums (the semantics of analysis/readRMS.m).

dki_shell / akc_shell / dki_lls restate readRMS.m:51-112 (signal-based
tensor fits); akc_lls needs the DKI class the reference does not ship.

readRMS.m:13-18 define the multiplicities of the 6 second-order and 15
fourth-order moment components; readRMS.m:44-45 normalise by sig0;
akc_mom (readRMS.m:96-103) gives D(t) = x2/(2t) and K(t) = x4/x2^2 - 3 along
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
    """dx2t, dx4t of readRMS.m:44-45 (multiplicity-weighted, /sig0)."""
    sig0 = np.asarray(sig0, float)[..., None]
    return np.asarray(dx2) * D_CNT / sig0, np.asarray(dx4) * W_CNT / sig0


def akc_mom(sig0, dx2, dx4, TD, n):
    """[K, D] along directions n (readRMS.m:96-103).  Returns arrays of shape
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


def dki_shell(sig, bvec, bval, TD):
    """Per-direction cumulant fit of the signal, then the moment tensors
    (readRMS.m:51-78): for each unique direction, least squares of log S on
    [-b, b^2/6, b^3/6, ...] (one column per distinct b-value) gives D and
    K = X2/D^2; <x^2> = 2 D t and <x^4> = (K+3) <x^2>^2 along the direction;
    the 6 + 15 tensor entries are the least-squares solutions over the
    directions.  sig [timepoints, nb], bvec [nb, 3], bval [nb]; returns
    [timepoints, 21] (dx2 entries first, then dx4), the layout akc_shell reads."""
    sig = np.atleast_2d(np.asarray(sig, float))
    bvec = np.asarray(bvec, float).reshape(-1, 3)
    bval = np.asarray(bval, float).ravel()
    TD = np.asarray(TD, float).ravel()
    bvecu, inv = np.unique(bvec, axis=0, return_inverse=True)
    inv = np.asarray(inv).ravel()
    n2 = ndir2(bvecu)
    n4 = ndir4(bvecu)
    nbval = np.unique(bval).size
    eps = np.finfo(float).eps
    out = np.zeros((sig.shape[0], 21))
    for i in range(sig.shape[0]):
        Di = np.zeros(bvecu.shape[0])
        Ki = np.zeros(bvecu.shape[0])
        for j in range(bvecu.shape[0]):
            sel = inv == j
            bj = bval[sel]
            A = np.column_stack([-bj] + [bj ** p / 6.0 for p in range(2, nbval + 1)])
            X = np.linalg.lstsq(A, np.log(sig[i, sel] + eps), rcond=None)[0]
            Di[j] = X[0]
            Ki[j] = X[1] / X[0] ** 2
        dx2g = Di * 2.0 * TD[i]
        dx4g = (Ki + 3.0) * dx2g ** 2
        out[i, :6] = np.linalg.lstsq(n2, dx2g, rcond=None)[0]
        out[i, 6:] = np.linalg.lstsq(n4, dx4g, rcond=None)[0]
    return out


def akc_shell(dt, TD, n):
    """[K, D] along directions n from dki_shell's tensors (readRMS.m:105-112)."""
    dt = np.atleast_2d(np.asarray(dt, float))
    x2 = dt[:, :6] @ ndir2(n).T
    x4 = dt[:, 6:21] @ ndir4(n).T
    D = x2 / 2.0 / np.asarray(TD, float)[:, None]
    K = x4 / x2 ** 2 - 3.0
    return K, D


def dki_lls(sig, bvec, bval):
    """Linear least-squares DKI fit of log S (readRMS.m:80-94): the b-matrix
    of the 6 diffusion entries (multiplicities D_CNT) and the 15 kurtosis
    entries (b^2/6, multiplicities W_CNT) on unit-normalised directions.
    sig [timepoints, nb]; returns [21, timepoints] like the reference."""
    sig = np.atleast_2d(np.asarray(sig, float))
    g = np.asarray(bvec, float).reshape(-1, 3).copy()
    b = np.asarray(bval, float).ravel()
    nrm = np.linalg.norm(g, axis=1)
    nrm[nrm == 0] = 1.0
    g /= nrm[:, None]
    B2 = -(b[:, None] * np.prod(g[:, D_IND], axis=2)) * D_CNT
    B4 = (b[:, None] ** 2 / 6.0 * np.prod(g[:, W_IND], axis=2)) * W_CNT
    Bm = np.hstack([B2, B4])
    eps = np.finfo(float).eps
    return np.linalg.lstsq(Bm, np.log(sig + eps).T, rcond=None)[0]


def narrow_pulse_signal(sig0, sigRe):
    """sig = |sigRe| / sig0 (readRMS.m:47-48)."""
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
