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


# 64/65, 100/101 and 128/129 rows: the LDS kernels' one- and two-row-per-lane
# solves and their limits (complex128 up to 100 rows, complex64 up to 128 in
# 160 KiB of LDS), then the global-memory kernels
@pytest.mark.parametrize("m,n", [(1, 1), (36, 1), (36, 290), (64, 5), (65, 5), (77, 65), (100, 20), (101, 3),
                                 (128, 17), (129, 2), (200, 3)])
def test_lu_solve(m, n):
    rng = np.random.default_rng(m + 7 * n)
    a, b = rnd(rng, m, m), rnd(rng, m, n)
    ref = np.linalg.solve(a, b)
    x = A.selftest_lu_solve(a, b)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-11
    x = A.selftest_lu_solve(a, b, c64=True)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-3


def lu_estimate(a):
    """Growth factor max|U|/max|A| times max|u_ii|/min|u_ii| of the LU with
    partial pivoting by modulus (the pivot rule of the kernels), in FP64."""
    u = np.array(a, np.complex128)
    n = u.shape[0]
    for k in range(n):
        p = k + int(np.argmax(np.abs(u[k:, k])))
        u[[k, p]] = u[[p, k]]
        if u[k, k] != 0:
            u[k + 1:, k] /= u[k, k]
            u[k + 1:, k + 1:] -= np.outer(u[k + 1:, k], u[k, k + 1:])
    U = np.triu(u)
    d = np.abs(np.diag(U))
    return np.abs(U).max() / np.abs(a).max() * d.max() / d.min() if d.min() > 0 else np.inf


def test_growth_gate():
    """The complex64 pivot-block gate: a block keeps its complex64 LU when its
    growth factor times condition estimate is at most the threshold, else it
    is factored and solved in FP64; blocks above 128 rows and singular ones
    always go to FP64."""
    rng = np.random.default_rng(5)
    a, b = rnd(rng, 40, 40), rnd(rng, 40, 3)
    a32 = a.astype(np.complex64).astype(np.complex128)  # what the kernels see
    ref = np.linalg.solve(a32, b.astype(np.complex64).astype(np.complex128))
    est_ref = lu_estimate(a32)
    x, est, fp64 = A.selftest_gated_lu_solve(a, b, 1e8)
    assert not fp64
    assert abs(est - est_ref) < 1e-3 * est_ref
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-3
    x, est, fp64 = A.selftest_gated_lu_solve(a, b, 0.5 * est_ref)
    assert fp64
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-6  # FP64 solve, result rounded to complex64
    # badly conditioned: diagonal scaling over 1e6
    s = np.logspace(0, 6, 40)
    x, est, fp64 = A.selftest_gated_lu_solve(a * s[None, :], b, 1e4)
    assert fp64 and est > 1e4
    # 101-128 rows: the complex64 factor fits the LDS, the FP64 one does not
    # (formed by the FP64 kernels gated on the block's flag)
    a2, b2 = rnd(rng, 110, 110), rnd(rng, 110, 4)
    ref2 = np.linalg.solve(a2.astype(np.complex64).astype(np.complex128), b2.astype(np.complex64).astype(np.complex128))
    x, est, fp64 = A.selftest_gated_lu_solve(a2, b2, 1e30)
    assert not fp64 and np.linalg.norm(x - ref2) / np.linalg.norm(ref2) < 1e-3
    x, est, fp64 = A.selftest_gated_lu_solve(a2, b2, 0.0)
    assert fp64 and np.linalg.norm(x - ref2) / np.linalg.norm(ref2) < 1e-6
    # above 128 rows and exactly singular blocks: FP64
    x, est, fp64 = A.selftest_gated_lu_solve(rnd(rng, 129, 129), rnd(rng, 129, 1), 1e30)
    assert fp64 and np.isinf(est)
    z = a.copy()
    z[:, 7] = 0
    x, est, fp64 = A.selftest_gated_lu_solve(z, b, 1e30)
    assert fp64


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


def q_times_r(F, tau, k):
    """Q R from the first k reflectors and the upper triangle of F's first k
    rows.  The kernels apply H_j = I - tau_j v v^H (v_j = 1 at row j, F below
    it) to A P, so A P = H_0^H ... H_{k-1}^H R."""
    m, n = F.shape
    Rm = np.triu(F)[:k]
    Q = np.eye(m, dtype=complex)[:, :k]
    for j in range(k - 1, -1, -1):  # Q = H_0 ... H_{k-1} applied to I[:, :k]
        v = np.zeros(m, complex)
        v[j] = 1
        v[j + 1:] = F[j + 1:, j]
        Q = Q - np.outer(np.conj(tau[j]) * v, v.conj() @ Q)
    return Q @ Rm


def cpqr_cases():
    rng = np.random.default_rng(21)
    out = []
    for m, n in [(1, 5), (17, 3), (36, 1), (36, 36), (36, 120), (36, 198), (36, 256), (24, 24)]:
        out.append((f"full{m}x{n}", rnd(rng, m, n), False))
    # low rank + noise with the early stop, as the recompression stacks are
    a = rnd(rng, 36, 12) @ rnd(rng, 12, 150) + 1e-9 * rnd(rng, 36, 150)
    out.append(("lowrank36x150", a, True))
    # duplicated and zero columns (ties in the pivot choice), badly scaled rows
    b = rnd(rng, 36, 40)
    b[:, 7] = b[:, 3]
    b[:, 20] = 0
    b[:, 30] = b[:, 3]
    b[30:] *= 1e-25
    out.append(("ties36x40", b, True))
    c = rnd(rng, 36, 8) @ rnd(rng, 8, 60)
    c[:, ::7] = 0
    out.append(("zeros36x60", c, True))
    return out


@pytest.mark.parametrize("c64", [False, True], ids=["c128", "c64"])
def test_cpqr_register_kernel_equals_global(c64):
    """k_cpqr_reg (one thread per column) against k_cpqr on the same input:
    the same rank, nonzero-pivot count and pivot order, factors equal to
    rounding (the two kernels sum dot products in different orders), and
    A P = Q R from the register kernel's own outputs."""
    tol = 2e-5 if c64 else 1e-12
    # complex64 rounding leaves ~1e-7 ||A|| in the trailing columns, so its
    # rank test needs a threshold above that to be a decision about A
    thr = 1e-4 if c64 else 1e-7
    for name, a, early in cpqr_cases():
        Fg, tg, pg, rg, zg = A.selftest_cpqr_raw(a, thr, reg=False, early=early, c64=c64)
        Fr, tr, pr, rr, zr = A.selftest_cpqr_raw(a, thr, reg=True, early=early, c64=c64)
        m, n = a.shape
        k = rr if early else min(m, n)
        print(f"  {name}: rank {rr}, nonzero {zr}", flush=True)
        assert rr == rg and zr == zg, name
        assert np.array_equal(pr[:k], pg[:k]), name
        assert sorted(pr.tolist()) == list(range(n)), name
        sc = np.linalg.norm(a)
        assert np.linalg.norm(Fr[:, :k] - Fg[:, :k]) <= tol * sc + 1e-300, name
        # tau_j moves with the rounding of R_jj's direction: |d tau_j| ~ eps ||A|| / |R_jj|
        eps = 6e-8 if c64 else 1.1e-16
        rd = np.abs(np.diag(Fg)[:k])
        assert np.all(np.abs(tr[:k] - tg[:k]) <= 100 * eps * sc / np.maximum(rd, 1e-300)), name
        if early:
            assert np.all(tr[k:] == 0), name
        QR = q_times_r(Fr, tr, k)
        assert np.linalg.norm(a[:, pr[:k]] - QR[:, :k]) <= 10 * tol * sc, name
        if not early and m <= n:
            assert np.linalg.norm(a[:, pr] - QR) <= 10 * tol * sc, name
    # the register kernel refuses what it cannot hold
    with pytest.raises(A.AifmmError):
        A.selftest_cpqr_raw(rnd(np.random.default_rng(1), 37, 4), 1e-7, reg=True, c64=c64)
    with pytest.raises(A.AifmmError):
        A.selftest_cpqr_raw(rnd(np.random.default_rng(1), 4, 257), 1e-7, reg=True, c64=c64)
