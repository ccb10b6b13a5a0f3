"""Pins of the oracle's primitives: Philox known answers (Random123 and the
rocRAND header as an independent implementation), the open-interval
uniforms, and the own sin/cos/exp against long-double libm."""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

# Random123 kat_vectors, philox4x32 with 10 rounds
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_random123_kat(orc, ctr, key, out):
    assert list(orc.philox(ctr, key)) == out


def test_philox_matches_rocrand_header(orc):
    data = json.load(open(os.path.join(HERE, "golden", "philox_rocrand.json")))
    assert len(data["vectors"]) >= 50
    for v in data["vectors"]:
        assert list(orc.philox(v["ctr"], v["key"])) == v["out"], v


def test_u01_open_interval_and_exact_grid(orc):
    assert orc.u01_f64(0, 0) == 2.0 ** -53
    assert orc.u01_f64(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert orc.u01_f32(0) == 2.0 ** -24
    assert orc.u01_f32(0xFFFFFFFF) == np.float32(1.0 - 2.0 ** -24)
    rng = np.random.default_rng(0)
    for a, b in rng.integers(0, 2 ** 32, size=(200, 2)):
        u = orc.u01_f64(int(a), int(b))
        k = ((int(a) >> 6) << 26) | (int(b) >> 6)
        assert u == (2 * k + 1) * 2.0 ** -53
        assert 0.0 < u < 1.0
    # the uniforms are unbiased: mean of many draws from the Philox stream
    us = [orc.u01_f64(*orc.philox([i, 0, 0, 0], [1, 2])[:2]) for i in range(4000)]
    assert abs(np.mean(us) - 0.5) < 4 * math.sqrt(1 / 12 / 4000)


def _ulp_err(got, want, dtype):
    want = np.asarray(want, dtype=np.longdouble)
    ulp = np.spacing(np.abs(want).astype(dtype)).astype(np.longdouble)
    ulp = np.maximum(ulp, np.longdouble(np.finfo(dtype).tiny))
    return float(np.max(np.abs(np.asarray(got, np.longdouble) - want) / ulp))


PI_LD = np.longdouble("3.14159265358979323846264338327950288")


def _turns_ref(t, dtype):
    """sin/cos(2*pi*t) in long double after the exact quarter-turn reduction
    (r = t - k/4 is exact), so the reference is not limited by pi's rounding."""
    t = np.asarray(t, dtype)
    k = np.floor(4 * t + dtype(0.5))
    r = (t - k * dtype(0.25)).astype(np.longdouble)
    s, c = np.sin(2 * PI_LD * r), np.cos(2 * PI_LD * r)
    q = k.astype(np.int64) & 3
    so = np.select([q == 0, q == 1, q == 2, q == 3], [s, c, -s, -c])
    co = np.select([q == 0, q == 1, q == 2, q == 3], [c, -s, -c, s])
    return so, co


def test_sincos_turns_f64_accuracy(orc):
    t = np.concatenate([np.linspace(1e-9, 1 - 1e-9, 3001), np.random.default_rng(1).random(2000)])
    s, c = zip(*[orc.sincos_turns(x, "f64") for x in t])
    rs, rc = _turns_ref(t, np.float64)
    assert _ulp_err(s, rs, np.float64) < 4
    assert _ulp_err(c, rc, np.float64) < 4
    # exact special values
    assert orc.sincos_turns(0.25, "f64") == (1.0, 0.0)
    assert orc.sincos_turns(0.5, "f64")[1] == -1.0


def test_sincos_turns_f32_accuracy(orc):
    """The FP32 direction's sin/cos (256-angle table + angle sum) on the
    draw's grid t = (2k+1) 2^-24 against long double: absolute error below
    2.5 ulp of 1, unit length within 3e-7."""
    rng = np.random.default_rng(3)
    m = np.concatenate([[1, (1 << 24) - 1, 1 << 23 | 1, (1 << 22) + 1, (3 << 22) + 1],
                        2 * rng.integers(0, 1 << 23, 6000) + 1, 2 * np.arange(0, 1 << 23, 1 << 13) + 1])
    t = (m * 2.0 ** -24).astype(np.float32)
    s, c = zip(*[orc.sincos_turns(float(x), "f32") for x in t])
    tl = m.astype(np.longdouble) * np.longdouble(2.0) ** -24
    rs, rc = np.sin(2 * PI_LD * tl), np.cos(2 * PI_LD * tl)
    assert np.max(np.abs(np.asarray(s, np.longdouble) - rs)) < 2.5 * 2.0 ** -24
    assert np.max(np.abs(np.asarray(c, np.longdouble) - rc)) < 2.5 * 2.0 ** -24
    n = np.asarray(s, np.float64) ** 2 + np.asarray(c, np.float64) ** 2
    assert np.max(np.abs(n - 1)) < 3e-7


def test_turn_table_matches_engine_header(orc):
    """The oracle's 256-angle table equals the engine's (rms_turn_table.h), and
    both are the float32 roundings of double sin/cos(2 pi k / 256)."""
    import ctypes as C
    import math
    import os
    import re
    ts, tc = (C.c_float * 256)(), (C.c_float * 256)()
    orc.lib().or_turn_table(ts, tc)
    hdr = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "monte-carlo-simulation-3d-rms_amd", "csrc", "rms_turn_table.h")
    pairs = re.findall(r"\{(-?0x[0-9a-fp.+-]+)f, (-?0x[0-9a-fp.+-]+)f\}", open(hdr).read())
    assert len(pairs) == 256
    for k, (hs, hc) in enumerate(pairs):
        a = 2 * math.pi * k / 256
        assert float.fromhex(hs) == ts[k] == float(np.float32(math.sin(a)))
        assert float.fromhex(hc) == tc[k] == float(np.float32(math.cos(a)))


def test_exp_accuracy(orc):
    x = np.concatenate([-np.logspace(-12, np.log10(700), 3000), [0.0, -1.0, -0.5]])
    got = [orc.exp(v) for v in x]
    assert _ulp_err(got, np.exp(x.astype(np.longdouble)), np.float64) < 3
    assert orc.exp(0.0) == 1.0
    assert orc.exp(-800.0) == 0.0


def test_sincos_any_accuracy(orc):
    x = np.concatenate([np.linspace(-50, 50, 4001), np.random.default_rng(2).uniform(-2e4, 2e4, 3000)])
    s, c = zip(*[orc.sincos(v) for v in x])
    xl = x.astype(np.longdouble)
    # absolute error relative to 1 (cos of large arguments): a few ulp of 1
    assert np.max(np.abs(np.asarray(c, np.longdouble) - np.cos(xl))) < 8 * 2.0 ** -53 * (1 + np.max(np.abs(x)) * 1e-12)
    assert np.max(np.abs(np.asarray(s, np.longdouble) - np.sin(xl))) < 8 * 2.0 ** -53 * (1 + np.max(np.abs(x)) * 1e-12)


def test_derive_matches_reference_formulas(orc):
    """main.cu:429-458: T2/dt, sqrt(6 dt D)/res, Pij rule and thresholds."""
    lab = np.ones((2, 2, 2), np.uint8)
    g = orc.Geometry(lab, 0.1, [0.0, 2.0, 0.5], T2=[80.0, 60.0, 20.0])
    d = orc.derive(g, 2e-4)
    np.testing.assert_array_equal(d["step"], np.sqrt(6.0 * 2e-4 * np.array([0.0, 2.0, 0.5])) / 0.1)
    np.testing.assert_array_equal(d["T2s"], np.array([80.0, 60.0, 20.0]) / 2e-4)
    P = d["P"]
    assert P[0, 0] == 0 and P[0, 1] == 0 and P[1, 0] == 0
    assert P[1, 1] == 1 and P[2, 2] == 1
    assert P[1, 2] == min(1.0, math.sqrt(0.5 / 2.0)) and P[2, 1] == 1.0
    np.testing.assert_array_equal(d["Pcat"], [[0, 0, 0], [0, 1, 2], [0, 1, 1]])
    # kappa extension: symmetric transmission keeps s_i P_ij = s_j P_ji
    gk = orc.Geometry(lab, 0.1, [0.0, 2.0, 0.5], kappa=0.02)
    dk = orc.derive(gk, 2e-4)
    s = dk["step"]
    assert abs(s[1] * dk["P"][1, 2] - s[2] * dk["P"][2, 1]) < 1e-15
    pi0 = 4 * 0.02 * 2e-4 / (0.1 * min(s[1], s[2]))
    assert abs(dk["P"][2, 1] - pi0 / (1 + pi0)) < 1e-15


def test_diff_times_and_pgse_lobes(orc):
    """TD = (i*(TN/1000)+1)*dt (main.cu:484-487); lobes use (i+1)dt < delta,
    Delta <= (i+1)dt < Delta+delta (main_PGSE.cu:341-345)."""
    TD = orc.diff_times(2e-4, 10000, 1000)
    np.testing.assert_array_equal(TD, (np.arange(1000) * 10 + 1) * 2e-4)
    grad = np.array([[10.0, 5.0, 0.1, 1, 0, 0], [2e-3, 1e-3, 1.0, 0, 1, 0], [50.0, 5.0, 1.0, 0, 0, 1]])
    abe = orc.pgse_lobes(2e-4, 100000, grad)
    for (a, b, e), gr in zip(abe, grad):
        i = np.arange(100000)
        t = (i + 1).astype(np.float64) * 2e-4
        l1 = t < gr[1]
        l2 = (~l1) & (t >= gr[0]) & (t < gr[0] + gr[1])
        assert a == l1.sum() and np.all(l1[:a])
        assert e - b == l2.sum()
        if l2.any():
            assert b == np.argmax(l2) and e == b + l2.sum()
