"""Pins of the Rys roots/weights (FP64 Chebyshev table and the quad-precision
Golub-Welsch reference) against what the mathematics fixes.

Convention (foofiles/rys.foo:89-134): roots u = t^2 in [0,1), weights with
sum_i w_i u_i^k = F_k(X), k = 0..2n-1.
"""
import numpy as np
import pytest
from scipy import special

XS = [0.0, 1e-8, 1e-3, 0.2, 0.99, 1.0, 2.5, 7.3, 13.0, 19.9, 26.4, 33.3, 39.9, 40.0, 47.9,
      53.9, 61.9, 67.9, 73.9, 77.9, 78.0, 95.0, 250.0, 1.0e4]


def boys_scipy(m, X):
    """F_m(X) = gamma(m+1/2) P(m+1/2, X) / (2 X^(m+1/2)) -- independent of the oracle."""
    if X == 0:
        return 1.0 / (2 * m + 1)
    a = m + 0.5
    return special.gamma(a) * special.gammainc(a, X) / (2.0 * X ** a)


@pytest.mark.parametrize("n", range(1, 8))
def test_table_vs_quad_reference(oracle_mod, n):
    worst = 0.0
    for X in XS:
        if X > 120:
            continue  # quad Golub-Welsch moments lose accuracy far into the Hermite regime
        u, w = oracle_mod.rys_table(n, X)
        uq, wq = oracle_mod.rys_quad(n, X)
        worst = max(worst, np.max(np.abs(u - uq) / uq), np.max(np.abs(w - wq) / wq))
    assert worst < 2e-15


@pytest.mark.parametrize("n", range(1, 8))
def test_quadrature_reproduces_boys_moments(oracle_mod, n):
    for X in [0.0, 0.5, 3.0, 11.0, 30.0, 60.0]:
        u, w = oracle_mod.rys_table(n, X)
        for k in range(2 * n):
            ref = boys_scipy(k, X)
            assert abs(np.sum(w * u ** k) - ref) <= 5e-14 * ref, (n, X, k)


def test_one_root_closed_form(oracle_mod):
    for X in [0.0, 0.3, 4.0, 25.0, 90.0]:
        u, w = oracle_mod.rys_table(1, X)
        F0, F1 = boys_scipy(0, X), boys_scipy(1, X)
        assert abs(w[0] - F0) < 1e-15 * F0
        assert abs(u[0] - F1 / F0) < 1e-14 * (F1 / F0)


@pytest.mark.parametrize("n", range(1, 8))
def test_zero_argument_is_gauss_legendre(oracle_mod, n):
    x, wl = np.polynomial.legendre.leggauss(2 * n)
    pos = x > 0
    u, w = oracle_mod.rys_table(n, 0.0)
    assert np.allclose(np.sort(u), np.sort(x[pos] ** 2), rtol=5e-14, atol=0)
    assert np.allclose(w[np.argsort(u)], wl[pos][np.argsort(x[pos] ** 2)], rtol=5e-14, atol=0)


@pytest.mark.parametrize("n", range(1, 8))
def test_large_argument_is_gauss_hermite(oracle_mod, n):
    s, wh = np.polynomial.hermite.hermgauss(2 * n)
    pos = s > 0
    X = 500.0
    u, w = oracle_mod.rys_table(n, X)
    order = np.argsort(s[pos])
    assert np.allclose(u, (s[pos] ** 2 / X)[order], rtol=1e-13, atol=0)
    assert np.allclose(w, (wh[pos] / np.sqrt(X))[order], rtol=1e-13, atol=0)


def test_quad_boys_matches_scipy(oracle_mod):
    for m in range(0, 14):
        for X in [0.0, 0.01, 1.0, 10.0, 35.0]:
            ref = boys_scipy(m, X)
            assert abs(oracle_mod.boys_quad(m, X) - ref) < 1e-13 * ref
