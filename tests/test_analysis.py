"""Analysis layer (readRMS.m / readRMS_pgse.m semantics, part1_demo4_analysis.m
combinations) against closed forms: signals built from known tensors are
fitted back, moment estimates of a Gaussian are exact."""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def A(rms):
    return importlib.import_module(rms.__name__ + ".analysis")


def _dirs(n, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_dki_shell_recovers_cumulants(A):
    """Signals from an exact second-order cumulant expansion whose moments are
    those of a two-compartment Gaussian mixture: <x^2>(n) is a quadratic form
    and <x^4>(n) a quartic form, so the shell fit returns them exactly."""
    T = 3
    TD = np.array([5.0, 10.0, 20.0])
    dirs = _dirs(30)
    bvals = np.array([0.5, 1.0, 2.0, 3.0])
    bvec = np.repeat(dirs, bvals.size, axis=0)
    bval = np.tile(bvals, dirs.shape[0])
    f = np.array([0.4, 0.6])
    C1 = np.diag([0.4, 0.5, 2.2])
    C2 = np.array([[1.2, 0.1, 0.0], [0.1, 0.9, 0.2], [0.0, 0.2, 1.0]])
    q1 = np.einsum("ij,jk,ik->i", dirs, C1, dirs)
    q2 = np.einsum("ij,jk,ik->i", dirs, C2, dirs)
    sig = np.zeros((T, bval.size))
    Dn_all, Kn_all = [], []
    for i in range(T):
        x2 = 2 * TD[i] * (f[0] * q1 + f[1] * q2)
        x4 = 3 * (2 * TD[i]) ** 2 * (f[0] * q1 ** 2 + f[1] * q2 ** 2)
        Dn = x2 / (2 * TD[i])
        Kn = x4 / x2 ** 2 - 3
        Dn_all.append(Dn)
        Kn_all.append(Kn)
        Db = np.repeat(Dn, bvals.size)
        Kb = np.repeat(Kn, bvals.size)
        sig[i] = np.exp(-bval * Db + bval ** 2 * Db ** 2 * Kb / 6.0)
    dt = A.dki_shell(sig, bvec, bval, TD)
    K, D = A.akc_shell(dt, TD, dirs)
    np.testing.assert_allclose(D, np.array(Dn_all), rtol=1e-9)
    np.testing.assert_allclose(K, np.array(Kn_all), atol=1e-8)


def test_dki_lls_gaussian_tensor(A):
    dirs = _dirs(40, seed=3)
    bvals = np.array([0.0, 1.0, 2.0])
    bvec = np.repeat(dirs, bvals.size, axis=0)
    bval = np.tile(bvals, dirs.shape[0])
    Dt = np.array([[1.5, 0.2, 0.1], [0.2, 1.0, 0.05], [0.1, 0.05, 0.7]])
    adc = np.einsum("ij,jk,ik->i", bvec, Dt, bvec)
    sig = np.exp(-bval * adc)[None, :]
    dt = A.dki_lls(sig, bvec, bval)
    d6 = dt[:6, 0]
    np.testing.assert_allclose(d6, [Dt[0, 0], Dt[0, 1], Dt[0, 2], Dt[1, 1], Dt[1, 2], Dt[2, 2]], atol=1e-9)
    np.testing.assert_allclose(dt[6:, 0], 0.0, atol=1e-9)


def test_akc_mom_gaussian_displacements(A, rms):
    """A displacement cloud with exactly known second and fourth moments."""
    rng = np.random.default_rng(1)
    T, N = 2, 200000
    sigma = np.array([0.5, 1.0, 2.0])
    x = rng.standard_normal((N, 3)) * sigma
    nclass = 1
    rec = np.zeros((T, rms.nval(nclass, 0)))
    for t in range(T):
        rec[t, 0] = N
        rec[t, 1] = N
        rec[t, 2:8] = [np.sum(x[:, a] * x[:, b]) for a, b in A.D_IND]
        rec[t, 8:23] = [np.sum(x[:, i] * x[:, j] * x[:, k] * x[:, l]) for i, j, k, l in A.W_IND]
    s = A.split(rec, nclass)
    TD = np.array([1.0, 2.0])
    K, D = A.akc_mom(s["sig0"], s["dx2"], s["dx4"], TD, np.eye(3))
    np.testing.assert_allclose(D[0], sigma ** 2 / 2.0, rtol=0.02)
    np.testing.assert_allclose(D[1], sigma ** 2 / 4.0, rtol=0.02)
    np.testing.assert_allclose(K, 0.0, atol=0.05)


def test_combine_compartments_and_dinf_fit(A):
    TD = np.linspace(1.0, 30.0, 40)
    Dinf, c = 1.3, 0.8
    D = Dinf + c / np.sqrt(TD)
    est = A.fit_dinf(TD, D)
    est = np.atleast_1d(np.asarray(est[0] if isinstance(est, tuple) else est, float))
    assert est[0] == pytest.approx(Dinf, rel=1e-9)
    f = np.array([0.3, 0.7])
    Dc = np.array([[2.0, 1.0]])
    Kc = np.array([[0.0, 0.0]])
    out = A.combine_compartments(f, Dc, Kc)
    Dm, Km = out
    assert float(np.ravel(Dm)[0]) == pytest.approx(0.3 * 2.0 + 0.7 * 1.0)
    # variance of the compartment diffusivities gives kurtosis 3 Var(D)/Dbar^2
    Dbar = 1.3
    assert float(np.ravel(Km)[0]) == pytest.approx(3 * (0.3 * 0.7 ** 2 + 0.7 * 0.3 ** 2) / Dbar ** 2, rel=1e-9)
