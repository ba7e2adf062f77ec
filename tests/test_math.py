"""ctsm_math.h (the log/exp/pow/log10 shared by the kernels and the oracle).

Pinned against 40-digit values from Python's decimal module (not against the
formulas of ctsm_math.h itself): every sampled result must be within 1 ulp of
the true value, and within 1 ulp of the host libm.  The GPU kernels compile
the same header, and tests/test_gpu_parity.py checks bitwise equality of the
physics that calls it.
"""
import ctypes as C
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle import oracle

getcontext().prec = 40


def _vec(name, *args):
    L = oracle.lib()
    fn = getattr(L, "oracle_cm_" + name)
    n = len(args[0])
    out = np.empty(n)
    ptrs = [np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double)) for a in args]
    fn(C.c_int(n), *ptrs, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def _ulp_err(got, exact):
    e = float(exact)
    return float(abs(Decimal(got) - exact) / Decimal(math.ulp(e)))


def _rng(k):
    return np.random.default_rng(0xC75D0100 + k)


def _samples_log():
    r = _rng(1)
    x = np.concatenate([10.0 ** r.uniform(-300, 300, 600), 1.0 + r.uniform(-1e-3, 1e-3, 200),
                        r.uniform(0.5, 2.0, 400), np.array([1.0, 2.0, 0.5, 10.0, math.e, 5e-324, 1.7e308])])
    return x


def test_log_within_one_ulp():
    x = _samples_log()
    got = _vec("log", x)
    worst = 0.0
    for xi, gi in zip(x, got):
        if xi == 1.0:
            assert gi == 0.0
            continue
        worst = max(worst, _ulp_err(gi, Decimal(xi).ln()))
    assert worst < 1.0, worst
    assert np.all(np.abs(got - np.log(x)) <= np.spacing(np.abs(np.log(x))) + 0.0)


def test_log10_within_one_ulp():
    x = _samples_log()
    x = x[x != 1.0]
    got = _vec("log10", x)
    worst = max(_ulp_err(g, Decimal(xi).log10()) for xi, g in zip(x, got))
    assert worst < 1.0, worst
    # exact powers of ten come out exact
    p = 10.0 ** np.arange(-20, 21)
    assert np.array_equal(_vec("log10", p), np.arange(-20, 21, dtype=float))


def test_exp_within_one_ulp():
    r = _rng(2)
    x = np.concatenate([r.uniform(-740, 709, 800), r.uniform(-1, 1, 400), r.uniform(-1e-8, 1e-8, 100),
                        np.array([0.0, 1.0, -1.0, 709.7, -745.0])])
    got = _vec("exp", x)
    worst = 0.0
    for xi, gi in zip(x, got):
        ex = Decimal(xi).exp()
        if ex < Decimal(2.2250738585072014e-308):    # subnormal results: absolute ulp of the subnormal range
            assert abs(Decimal(gi) - ex) <= Decimal(5e-324)
            continue
        worst = max(worst, _ulp_err(gi, ex))
    assert worst < 1.0, worst
    assert got[-5] == 1.0


def test_pow_within_one_ulp_on_physics_ranges():
    """Clapp-Hornberger forms: s**(-B), s**(2B+3), s**(2B+2), (1-s)**k, small and huge results."""
    r = _rng(3)
    n = 600
    s = np.concatenate([r.uniform(0.01, 1.0, n), 10.0 ** r.uniform(-6, 0, n), r.uniform(0.5, 2000.0, n)])
    y = np.concatenate([-r.uniform(2.5, 13.0, n), r.uniform(8.0, 30.0, n), r.uniform(-3.0, 3.0, n)])
    got = _vec("pow", s, y)
    worst = 0.0
    for si, yi, gi in zip(s, y, got):
        ex = (Decimal(yi) * Decimal(si).ln()).exp()
        if ex < Decimal(2.2250738585072014e-308):
            continue
        worst = max(worst, _ulp_err(gi, ex))
    assert worst < 1.0, worst
    ref = np.power(s, y)
    ok = np.isfinite(ref) & (ref > 1e-300)
    assert np.all(np.abs(got[ok] - ref[ok]) <= np.spacing(np.abs(ref[ok])))


@pytest.mark.parametrize("x,y,want", [
    (2.0, 10.0, 1024.0), (4.0, 0.5, 2.0), (9.0, 0.5, 3.0), (1.0, 123.4, 1.0), (7.3, 0.0, 1.0),
    (2.0, -3.0, 0.125), (10.0, 15.0, 1e15), (0.0, 2.0, 0.0), (0.0, -1.0, math.inf), (-2.0, 3.0, -8.0),
])
def test_pow_special_cases(x, y, want):
    assert _vec("pow", np.array([x]), np.array([y]))[0] == want


def test_atan_within_one_ulp():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    r = _rng(4)
    x = np.concatenate([r.uniform(-1, 1, 500), 10.0 ** r.uniform(-8, 8, 500), -(10.0 ** r.uniform(-3, 3, 200)),
                        1.0 + r.uniform(0, 1e-3, 100), (np.arange(0, 17) / 16.0), np.array([1e300, -1e-300, np.inf])])
    got = _vec("atan", x)
    worst = 0.0
    for xi, gi in zip(x, got):
        ex = mp.atan(mp.mpf(xi))
        e = float(ex)
        if e == 0.0:
            assert gi == 0.0
            continue
        worst = max(worst, float(abs(mp.mpf(gi) - ex) / mp.mpf(math.ulp(e))))
    assert worst < 1.0, worst
    assert np.all(np.abs(got - np.arctan(x)) <= np.spacing(np.abs(np.arctan(x))))
    assert _vec("atan", np.array([1.0]))[0] == math.atan(1.0)


def test_sin_within_one_ulp():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    r = _rng(6)
    deg = r.uniform(0.0, 90.0, 600)
    x = np.concatenate([(math.pi / 180.0) * deg, r.uniform(-10, 10, 400), 10.0 ** r.uniform(-9, 0, 200),
                        r.uniform(-1000, 1000, 100), np.array([math.pi / 2, math.pi, 1e-300, -0.5, 0.0])])
    got = _vec("sin", x)
    worst = 0.0
    for xi, gi in zip(x, got):
        ex = mp.sin(mp.mpf(xi))
        e = float(ex)
        if e == 0.0:
            assert gi == 0.0
            continue
        worst = max(worst, float(abs(mp.mpf(gi) - ex) / mp.mpf(math.ulp(e))))
    assert worst < 1.0, worst
    assert np.all(np.abs(got - np.sin(x)) <= 2 * np.spacing(np.abs(np.sin(x))))


def test_tan_within_two_and_a_half_ulp():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    r = _rng(7)
    deg = r.uniform(0.0, 89.0, 600)
    x = np.concatenate([(math.pi / 180.0) * deg, r.uniform(-10, 10, 400), 10.0 ** r.uniform(-9, 0, 200),
                        np.array([1e-300, -0.5, 0.0, 1.0])])
    got = _vec("tan", x)
    worst = 0.0
    for xi, gi in zip(x, got):
        ex = mp.tan(mp.mpf(xi))
        e = float(ex)
        if e == 0.0:
            assert gi == 0.0
            continue
        worst = max(worst, float(abs(mp.mpf(gi) - ex) / mp.mpf(math.ulp(e))))
    assert worst < 2.5, worst
    assert np.all(np.abs(got - np.tan(x)) <= 3 * np.spacing(np.abs(np.tan(x))))


def test_pow_and_log_hard_cases():
    """x near 1 with large exponents (log cancellation), results near the
    overflow / underflow thresholds, and table-boundary arguments."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 50
    r = _rng(5)
    x = np.concatenate([1.0 + r.uniform(-1e-6, 1e-6, 200), 1.0 + r.uniform(-1e-12, 1e-12, 100),
                        r.uniform(1.5, 3.0, 200),
                        np.array([(k + 0.5) / 128.0 * 2.0 ** e for k in range(91, 181, 7) for e in (-3, 0, 5)])])
    y = np.concatenate([r.uniform(-5e4, 5e4, 300), r.uniform(600.0, 1000.0, 200) * np.sign(r.uniform(-1, 1, 200)),
                        r.uniform(-30.0, 30.0, len(x) - 500)])
    got = _vec("pow", x, y)
    worst = 0.0
    for xi, yi, gi in zip(x, y, got):
        ex = mp.power(mp.mpf(xi), mp.mpf(yi))
        if ex > mp.mpf("1.7e308") or ex < mp.mpf("2.3e-308"):
            continue
        worst = max(worst, float(abs(mp.mpf(gi) - ex) / mp.mpf(math.ulp(float(ex)))))
    assert worst < 1.0, worst
    lx = _vec("log", x)
    for xi, gi in zip(x, lx):
        if xi == 1.0:
            continue
        ex = mp.log(mp.mpf(xi))
        assert float(abs(mp.mpf(gi) - ex) / mp.mpf(math.ulp(float(ex)))) < 1.0
