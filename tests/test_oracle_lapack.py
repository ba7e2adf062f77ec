"""Pins of the oracle's LAPACK restatements (dgtsv, dgbsv).

The reference calls the machine LAPACK (SoilWaterMovementMod.F90:1287,
BandDiagonalMod.F90:197) whose version CTSM does not pin; scipy's bundled
LAPACK is the independent implementation checked against here, plus dense
solves and residuals.  Pivot sequences must match exactly because the pivot
choice changes the rounding the kernels have to reproduce.
"""
import numpy as np
import pytest
from scipy.linalg import lapack

from oracle import oracle as O


def _tri(n, rng, pivot_bias=0.0):
    dl = rng.uniform(-1, 1, n - 1)
    du = rng.uniform(-1, 1, n - 1)
    d = rng.uniform(-1, 1, n) + np.sign(rng.uniform(-1, 1, n)) * (1.0 - pivot_bias)
    b = rng.uniform(-1, 1, n)
    return dl, d, du, b


@pytest.mark.parametrize("n", [2, 3, 5, 20, 38])
@pytest.mark.parametrize("bias", [0.0, 0.9])
def test_dgtsv_matches_scipy(n, bias):
    rng = np.random.default_rng(100 + n)
    for _ in range(50):
        dl, d, du, b = _tri(n, rng, bias)
        x, info = O.dgtsv(dl, d, du, b)
        _, _, _, xs, infos = lapack.dgtsv(dl.copy(), d.copy(), du.copy(), b.copy())
        assert info == infos == 0
        assert np.allclose(x, xs, rtol=1e-12, atol=1e-12 * np.max(np.abs(xs)))
        A = np.diag(d) + np.diag(dl, -1) + np.diag(du, 1)
        assert np.allclose(A @ x, b, rtol=1e-10, atol=1e-10)


def test_dgtsv_pivoting_path_exact():
    """A system whose sub-diagonal dominates forces the row-interchange branch
    at every step; the 2x2 case reduces to Cramer's rule."""
    dl = np.array([4.0])
    d = np.array([1.0, 2.0])
    du = np.array([3.0])
    b = np.array([5.0, 6.0])
    x, info = O.dgtsv(dl, d, du, b)
    # [[1,3],[4,2]] x = [5,6] -> x = [0.8, 1.4]
    assert info == 0 and np.allclose(x, [0.8, 1.4], rtol=0, atol=1e-15)
    n = 10
    dl = np.full(n - 1, 10.0)
    d = np.full(n, 1.0)
    du = np.full(n - 1, 1.0)
    b = np.arange(1.0, n + 1)
    x, _ = O.dgtsv(dl, d, du, b)
    _, _, _, xs, _ = lapack.dgtsv(dl.copy(), d.copy(), du.copy(), b.copy())
    assert np.allclose(x, xs, rtol=1e-13)


def test_dgtsv_singular_info():
    x, info = O.dgtsv([0.0, 1.0], [0.0, 1.0, 1.0], [1.0, 1.0], [1.0, 1.0, 1.0])
    assert info == 1
    x, info = O.dgtsv([1.0], [1.0, 1.0], [1.0], [1.0, 2.0])   # [[1,1],[1,1]]
    assert info == 2


def _band(n, kl, ku, rng, dominant):
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(max(0, i - kl), min(n, i + ku + 1)):
            A[i, j] = rng.uniform(-1, 1)
        if dominant:
            A[i, i] += 5.0 * np.sign(A[i, i] or 1.0)
    ab = np.zeros((2 * kl + ku + 1, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            ab[kl + ku + i - j, j] = A[i, j]
    return A, ab


@pytest.mark.parametrize("n", [1, 2, 3, 4, 26, 38])
@pytest.mark.parametrize("dominant", [True, False])
def test_dgbsv_matches_scipy(n, dominant):
    rng = np.random.default_rng(7 * n + dominant)
    for _ in range(40):
        A, ab = _band(n, 2, 2, rng, dominant)
        b = rng.uniform(-1, 1, n)
        x, ipiv, info = O.dgbsv(2, 2, ab, b)
        lub, piv, xs, infos = lapack.dgbsv(2, 2, ab.copy(), b.copy())
        assert info == infos == 0
        assert np.array_equal(ipiv, piv + 1), "pivot sequence differs from LAPACK"
        assert np.allclose(x, xs, rtol=1e-11, atol=1e-11 * np.max(np.abs(xs)))
        assert np.allclose(A @ x, b, atol=1e-9)
        if n <= 4:
            assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)


def test_dgbsv_zero_pivot_info():
    A = np.array([[0.0, 1.0, 0.0], [0.0, 1.0, 1.0], [0.0, 0.0, 1.0]])
    ab = np.zeros((7, 3))
    for j in range(3):
        for i in range(max(0, j - 2), min(3, j + 3)):
            ab[4 + i - j, j] = A[i, j]
    x, ipiv, info = O.dgbsv(2, 2, ab, np.ones(3))
    assert info == 1
