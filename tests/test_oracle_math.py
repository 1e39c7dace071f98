"""The oracle's numerical building blocks against independent references.

The reference ships no golden vectors (no tests, rhs_x/ and precomputations/
are run-time outputs), so each oracle function is pinned here against
published algorithms in numpy/scipy or against closed forms.
"""
import numpy as np
import pytest
import scipy.special as sp


def test_gauss_legendre_10_matches_golub_welsch(oracle):
    import mpmath as mp
    mp.mp.dps = 40
    x, w = oracle.gauss_legendre(10)
    xr, wr = np.polynomial.legendre.leggauss(10)  # Golub-Welsch
    o = np.argsort(x)
    np.testing.assert_allclose(x[o], xr, rtol=0, atol=5e-16)
    np.testing.assert_allclose(w[o], wr, rtol=0, atol=5e-16)
    # 40-digit nodes/weights: the oracle's doubles are correctly rounded (<= 1/2 ulp + tiny)
    for xi, wi in zip(x, w):
        z = mp.findroot(lambda t: mp.legendre(10, t), mp.mpf(xi))
        dp = mp.diff(lambda t: mp.legendre(10, t), z)
        wz = 2 / ((1 - z ** 2) * dp ** 2)
        assert abs(mp.mpf(xi) - z) <= 0.6 * np.spacing(abs(xi))
        assert abs(mp.mpf(wi) - wz) <= 0.6 * np.spacing(wi)
    # ordering of the reference table: (-x, +x) pairs, |x| ascending
    assert np.all(np.diff(np.abs(x)[::2]) > 0) and np.all(x[::2] < 0) and np.all(x[1::2] > 0)
    # exact for degree <= 19, not for 20
    for d in range(20):
        assert abs(np.sum(w * x ** d) - (2.0 / (d + 1) if d % 2 == 0 else 0.0)) < 1e-14
    assert abs(np.sum(w * x ** 20) - 2.0 / 21) > 1e-6


@pytest.mark.parametrize("x", [0.05, 0.41, 0.95, 1.9, 3.7, 10.0, 27.3, 60.0, 113.0, 170.0])
def test_bessel_arrays_vs_scipy(oracle, x):
    n = np.arange(61)
    for exact in (False, True):
        J, Y = oracle.bessel(x, 60, exact)
        Jr, Yr = sp.jv(n, x), sp.yv(n, x)
        # J relative to itself where it decays monotonically (n > x), relative
        # to |H_n| = |J_n + i Y_n| in the oscillatory range where J has zeros
        scaleJ = np.where(n > x, np.abs(Jr), np.hypot(Jr, Yr))
        ok = scaleJ > 1e-290
        assert np.max(np.abs(J - Jr)[ok] / scaleJ[ok]) < 1e-12, (x, exact)
        scaleY = np.where(n > x, np.abs(Yr), np.hypot(Jr, Yr))
        assert np.max(np.abs(Y - Yr) / scaleY) < 1e-12, (x, exact)


def test_cpqr_factorisation_and_rank(oracle):
    rng = np.random.default_rng(3)
    m, n, r = 40, 30, 12
    A = (rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))) @ (
        rng.standard_normal((r, n)) + 1j * rng.standard_normal((r, n)))
    A += 1e-10 * (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))
    rank, Q, rd, perm = oracle.cpqr(A, 1e-8, m)
    assert rank == r
    np.testing.assert_allclose(Q.conj().T @ Q, np.eye(m), atol=1e-13)
    R = Q.conj().T @ A[:, perm]
    assert np.max(np.abs(np.tril(R, -1))) < 1e-12 * np.abs(R).max()
    np.testing.assert_allclose(np.abs(np.diag(R)), rd, rtol=1e-10, atol=1e-13 * rd[0])
    assert np.all(np.diff(rd[:r]) <= 1e-12 * rd[0])  # pivoting orders |R_ii|
    # the first `rank` columns span the range of A
    Qr = Q[:, :rank]
    assert np.linalg.norm(A - Qr @ (Qr.conj().T @ A)) < 1e-8 * np.linalg.norm(A)


def test_dense_solvers(oracle):
    rng = np.random.default_rng(4)
    n = 37
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    X = np.linalg.solve(A, B)
    np.testing.assert_allclose(oracle.qr_solve(A, B), X, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(oracle.lu_solve(A, B), X, rtol=1e-11, atol=1e-11)
