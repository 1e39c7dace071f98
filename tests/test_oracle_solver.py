"""Tree lists, ACA/NCA and the hierarchical solve of the oracle, pinned by
brute force (coordinates, dense factorisations) and by invariants."""
import numpy as np
import pytest


def bxy(level, b):
    ix = iy = 0
    for t in range(level - 1, -1, -1):
        c = (b >> (2 * t)) & 3
        ix = 2 * ix + (c in (1, 2))
        iy = 2 * iy + (c in (2, 3))
    return ix, iy


@pytest.fixture(scope="module")
def p3(oracle):
    return oracle.Problem(n_levels=3, kappa=10.0, tol_pow=8, qchoice=0)


def test_child_order_and_points(p3):
    xy, _ = p3.points()
    # leaf k owns points 36k..36k+35, centred on the box of (ix, iy)
    for k in (0, 1, 2, 3, 17, 63):
        ix, iy = bxy(3, k)
        c = xy[36 * k:36 * k + 36].mean(0)
        np.testing.assert_allclose(c, [-1 + (2 * ix + 1) / 8, -1 + (2 * iy + 1) / 8], atol=1e-15)
    # children 0..3 of the root sit at (-,-), (+,-), (+,+), (-,+)
    assert [bxy(1, c) for c in range(4)] == [(0, 0), (1, 0), (1, 1), (0, 1)]


def test_interaction_lists_partition_far_field(p3):
    """Every ordered pair of distinct leaves is either adjacent (near field)
    or separated at exactly one level j >= 2, where their ancestors are in
    each other's interaction list; lists hold distinct, non-adjacent boxes."""
    L = 3
    lists = {j: [p3.lists(j, k) for k in range(4 ** j)] for j in range(2, L + 1)}
    for j in range(2, L + 1):
        for k in range(4 ** j):
            nb, inner, outer = lists[j][k]
            x, y = bxy(j, k)
            il = [b for b in np.concatenate([inner, outer]) if b >= 0]
            assert len(set(il)) == len(il)
            for b in il:
                bx, by = bxy(j, b)
                assert max(abs(bx - x), abs(by - y)) >= 2
                # parents are colleagues
                assert max(abs(bx // 2 - x // 2), abs(by // 2 - y // 2)) <= 1
            for s, b in enumerate(nb):
                off = [(-1, -1), (0, -1), (1, -1), (1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0)][s]
                inside = 0 <= x + off[0] < 2 ** j and 0 <= y + off[1] < 2 ** j
                assert (b >= 0) == inside
                if b >= 0:
                    assert bxy(j, b) == (x + off[0], y + off[1])
    for a in range(4 ** L):
        for b in range(4 ** L):
            if a == b:
                continue
            ax, ay = bxy(L, a)
            bx, by = bxy(L, b)
            if max(abs(ax - bx), abs(ay - by)) <= 1:
                continue
            hits = 0
            for j in range(2, L + 1):
                aa, bb = a >> (2 * (L - j)), b >> (2 * (L - j))
                _, inner, outer = lists[j][aa]
                hits += int(bb in set(inner) | set(outer))
            assert hits == 1, (a, b)


def test_aca_returns_raw_pivot_rows_and_columns(p3):
    xy, q = p3.points()
    leafq = np.abs(q).reshape(-1, 36).max(1)
    k = int(np.argmax(leafq))
    _, inner, outer = p3.lists(3, k)
    il = [b for b in np.concatenate([inner, outer]) if b >= 0]
    rows = np.arange(36 * k, 36 * k + 36)
    cols = np.concatenate([np.arange(36 * b, 36 * b + 36) for b in il])
    I, J = np.meshgrid(rows, cols, indexing="ij")
    K = p3.entries(I.ravel(), J.ravel()).reshape(I.shape)
    for tp in (6, 8, 10):
        rb, cb, Ac, Ar = p3.aca(rows, cols, tp)
        assert len(set(rb)) == len(rb) and len(set(cb)) == len(cb)
        np.testing.assert_array_equal(Ac, K[:, cb])
        np.testing.assert_array_equal(Ar, K[rb, :])
        approx = Ac @ np.linalg.solve(K[np.ix_(rb, cb)], Ar)
        err = np.linalg.norm(K - approx) / np.linalg.norm(K)
        # first pivot row = first row whose max exceeds 1e-12 (ACA.hpp:397-413)
        first = next(i for i in range(36) if np.abs(K[i]).max() > 1e-12)
        a = np.abs(K[first])
        # first column within a relative 1e-12 of the maximum (ties to the lowest index)
        assert rb[0] == first and cb[0] == int(np.nonzero(a >= a.max() * (1 - 1e-12))[0][0])
        if abs(K[rb[0], cb[0]]) <= 10.0 ** -tp:
            # a weak first pivot ends the loop at rank 1 (ACA.hpp:435)
            assert len(rb) == 1
        else:
            # the norm criterion stops near sqrt(tol)-relative updates
            assert err < 100 * np.sqrt(10.0 ** -tp), (tp, err)
    # rows reordered strongest-first: the iteration runs and converges
    order = np.argsort(-np.abs(K).max(1))
    rb, cb, Ac, Ar = p3.aca(rows[order], cols, 6)
    Ko = K[order]
    err = np.linalg.norm(Ko - Ac @ np.linalg.solve(Ko[np.ix_(rb, cb)], Ar)) / np.linalg.norm(Ko)
    assert len(rb) > 1 and err < 0.1
    # transposed problem (rows > cols) takes the column-first branch
    K2 = p3.entries(J.T.ravel(), I.T.ravel()).reshape(J.T.shape)  # A(IL points, box points)
    rb2, cb2, Ac2, Ar2 = p3.aca(cols, rows, 8)
    np.testing.assert_array_equal(Ar2, K2[rb2, :])
    np.testing.assert_array_equal(Ac2, K2[:, cb2])


@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_solve_exact_at_two_levels(oracle, schedule):
    """At L = 2 every far block is an M2L block of the base system, so the
    hierarchical solve reproduces the dense FP64 solution."""
    p = oracle.Problem(n_levels=2, kappa=10.0, tol_pow=8, qchoice=0)
    A = p.dense()
    b = p.rhs(0.0)
    x, _ = p.solve(b, schedule)
    xd = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-12


@pytest.fixture(scope="module")
def l3_solutions(oracle):
    p = oracle.Problem(n_levels=3, kappa=10.0, tol_pow=8, qchoice=0)
    A = p.dense()
    b = p.rhs(0.0)
    xd = np.linalg.solve(A, b)
    xs = {s: p.solve(b, s)[0] for s in (0, 1, 2)}
    return A, b, xd, xs


def test_three_level_solve_against_dense(l3_solutions):
    A, b, xd, xs = l3_solutions
    for s, x in xs.items():
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 5e-5, s
        assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 5e-5, s


def test_schedules_agree(l3_solutions):
    _, _, xd, xs = l3_solutions
    n = np.linalg.norm(xs[0])
    # efficient (reference) vs eager vs colour schedule: 10x the 1e-8 tolerance
    assert np.linalg.norm(xs[0] - xs[1]) / n < 2e-7
    assert np.linalg.norm(xs[0] - xs[2]) / n < 2e-7
