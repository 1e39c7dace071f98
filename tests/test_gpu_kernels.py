"""Batched dense gfx950 kernels (through the C ABI self-test hook) against
numpy in float64 / complex128."""
import numpy as np
import pytest

import aifmm_ls_amd as A

pytestmark = pytest.mark.gpu


def rnd(rng, *s):
    return rng.standard_normal(s) + 1j * rng.standard_normal(s)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (36, 36, 36), (33, 70, 17), (5, 300, 64), (129, 3, 250)])
def test_gemm(m, n, k):
    rng = np.random.default_rng(m * n + k)
    a, b = rnd(rng, m, k), rnd(rng, k, n)
    ref = a @ b
    np.testing.assert_allclose(A.selftest_gemm(a, b), ref, rtol=1e-13, atol=1e-12 * np.abs(ref).max())
    c = A.selftest_gemm(a, b, c64=True)
    assert np.abs(c - ref).max() < 1e-5 * np.abs(ref).max() * np.sqrt(k)


@pytest.mark.parametrize("m,n", [(1, 1), (36, 1), (36, 290), (77, 65), (200, 3)])
def test_lu_solve(m, n):
    rng = np.random.default_rng(m + 7 * n)
    a, b = rnd(rng, m, m), rnd(rng, m, n)
    ref = np.linalg.solve(a, b)
    x = A.selftest_lu_solve(a, b)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-11
    x = A.selftest_lu_solve(a, b, c64=True)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-3


@pytest.mark.parametrize("m,n", [(1, 1), (36, 1), (36, 290), (77, 65), (200, 3), (130, 17)])
def test_qr_solve(m, n):
    rng = np.random.default_rng(3 * m + n)
    a, b = rnd(rng, m, m), rnd(rng, m, n)
    ref = np.linalg.solve(a, b)
    x, nz = A.selftest_qr_solve(a, b)
    assert nz == m
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-11
    x, nz = A.selftest_qr_solve(a, b, c64=True)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-3


def test_qr_solve_rank_deficient_matches_oracle():
    """Singular A: the solve keeps the nonzero pivots only (unknowns past them
    are 0), as ColPivHouseholderQR::solve; compared with the oracle's CPQR."""
    from oracle import oracle as O
    rng = np.random.default_rng(11)
    m, r = 24, 9
    # exactly singular: m - r zero columns in shuffled positions
    a = np.concatenate([rnd(rng, m, r), np.zeros((m, m - r))], axis=1)[:, rng.permutation(m)]
    b = rnd(rng, m, 3)
    x, nz = A.selftest_qr_solve(a, b)
    xo = O.qr_solve(a, b)
    assert nz == r
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-10


@pytest.mark.parametrize("m,n,r", [(36, 20, 8), (40, 130, 12), (100, 60, 31)])
def test_cpqr_rank_and_basis(m, n, r):
    rng = np.random.default_rng(m + n + r)
    a = rnd(rng, m, r) @ rnd(rng, r, n) + 1e-11 * rnd(rng, m, n)
    rank, Q = A.selftest_cpqr(a, 1e-8, r)
    assert rank == r
    np.testing.assert_allclose(Q.conj().T @ Q, np.eye(r), atol=1e-12)
    assert np.linalg.norm(a - Q @ (Q.conj().T @ a)) < 1e-9 * np.linalg.norm(a)
