"""Pins of the history-buffer oracle (oracle/history.py) against closed forms.

hist_update_hbuf / hfields_normalize (histFileMod.F90:1362-2358) have no
golden vectors in the reference's tests; what fixes them:
  * 'A' over integer-valued series (exact in binary64) equals the mean;
    'S' equals the sum and is not normalised; 'I' holds the last value;
    'X' / 'M' the running maximum / minimum;
  * missing values (1e36) and inactive points are skipped: a point missing on
    some steps averages the others, one missing on every step stays 1e36, an
    inactive point is 1e36;
  * for 'X' / 'M' the first step fixes the count at 1, so a point missing on
    the first step starts from 1e36: the maximum stays 1e36, the minimum is
    that of the later steps (the reference's order of tests);
  * hfields_zero clears buffer and count.
"""
import numpy as np
import pytest

from oracle import history as H

SPVAL = H.SPVAL


def _run(flag, series, active=None):
    n2, n1 = series[0].shape
    hbuf = np.zeros((n2, n1))
    nacs = np.zeros((n2, n1), dtype=np.int32)
    for x in series:
        H.hist_update(hbuf, nacs, x, active, flag)
    H.hist_normalize(hbuf, nacs, flag)
    return hbuf, nacs


def _series(steps=7, shape=(3, 50), seed=1):
    rng = np.random.default_rng(seed)
    return [rng.integers(-500, 500, shape).astype(np.float64) for _ in range(steps)]


@pytest.mark.parametrize("flag", ["A", "S", "I", "X", "M"])
def test_closed_forms(flag):
    s = _series()
    h, n = _run(flag, s)
    a = np.stack(s)
    expect = {"A": a.sum(0) / len(s), "S": a.sum(0), "I": a[-1], "X": a.max(0), "M": a.min(0)}[flag]
    assert np.array_equal(h, expect)
    assert np.all(n == (len(s) if flag in "AS" else 1))


def test_missing_and_inactive_points():
    s = _series(steps=6, shape=(2, 40), seed=2)
    miss = np.zeros((6, 2, 40), dtype=bool)
    miss[1:4, :, 5] = True           # missing on some steps
    miss[:, :, 6] = True             # missing on every step
    for t in range(6):
        s[t][miss[t]] = SPVAL
    active = np.ones(40, dtype=bool)
    active[7] = False
    h, n = _run("A", s, active)
    a = np.stack(s)
    ok = ~miss & active[None, None, :]
    cnt = ok.sum(0)
    mean = np.where(cnt > 0, np.where(ok, a, 0.0).sum(0) / np.maximum(cnt, 1), SPVAL)
    assert np.array_equal(h, mean)
    assert np.array_equal(n, cnt)
    assert np.all(h[:, 6] == SPVAL) and np.all(h[:, 7] == SPVAL)
    # instantaneous: missing / inactive points are spval on that step
    hi, _ = _run("I", s, active)
    assert np.all(hi[:, 7] == SPVAL) and np.all(hi[:, 6] == SPVAL)


def test_extremes_keep_the_first_steps_missing_value():
    s = _series(steps=4, shape=(1, 10), seed=3)
    s[0][0, 2] = SPVAL
    others = [k for k in range(10) if k != 2]
    a = np.stack(s)[:, 0, :]
    h, _ = _run("X", s)
    assert h[0, 2] == SPVAL
    assert np.array_equal(h[0, others], a[:, others].max(0))
    h, _ = _run("M", s)
    assert h[0, 2] == a[1:, 2].min()
    assert np.array_equal(h[0, others], a[:, others].min(0))


def test_normalise_uses_single_precision_count_and_zero_clears():
    hbuf = np.array([[10.0, 7.0, 3.0]])
    nacs = np.array([[4, 0, 3]], dtype=np.int32)
    H.hist_normalize(hbuf, nacs, "A")
    assert np.array_equal(hbuf, [[2.5, 7.0, 1.0]])
    h2 = hbuf.copy()
    H.hist_normalize(h2, nacs, "S")
    assert np.array_equal(h2, hbuf)
    H.hist_zero(hbuf, nacs)
    assert np.all(hbuf == 0) and np.all(nacs == 0)
