"""The adaptive discretisation on the GPU (aifmm_ls_create_adaptive): the
quadtree with 2:1 level restriction (LS_Assembly.hpp:609-742, 892-1072) and
getMatrixEntry between leaves of different levels (:283-541), against the
oracle's restatement.

Checks: leaves (level, box number, centre) in the oracle's order and the
points bitwise; entries of random pairs and of every near pair of sampled
leaves (all four neighbour families) against the oracle; the dense product
against the oracle's O(N^2) one; a tree that comes out uniform gives the
uniform problem's entries and product; the solver entry points refuse the
adaptive discretisation."""
import numpy as np
import pytest

import aifmm_ls_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [  # kappa, qchoice, bumps, tree adaptivity
    (15.0, 1, None, 2),
    (10.0, 6, [[1.0, 0.03, 0.37, 0.21]], 3),
    (40.0, 6, [[1.2, 0.08, 0.5, 0.1], [0.5, 0.2, -0.3, -0.5], [0.9, 0.05, 0.0, 0.6]], 3),
]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module", params=range(len(CASES)), ids=["k15-q1-t2", "k10-bump-t3", "k40-bumps-t3"])
def pair(request):
    kappa, qc, bumps, ta = CASES[request.param]
    bm = np.array(bumps) if bumps else None
    print(f"\n[adaptive] kappa {kappa}, contrast {qc}, tree adaptivity {ta}", flush=True)
    g = A.LSProblem(0, kappa, 8, qchoice=qc, bumps=bm, tree_adaptivity=ta)
    o = O.AdaptiveProblem(kappa, ta, qchoice=qc, bumps=bm)
    yield g, o
    g.close()


def test_leaves_and_points(pair):
    g, o = pair
    lv, bx, c = g.leaves()
    lo, bo, co = o.leaves()
    assert np.array_equal(lv, lo) and np.array_equal(bx, bo) and np.array_equal(c, co)
    xy, q = g.points()
    xo, qo = o.points()
    assert g.N == o.N
    assert np.array_equal(xy, xo) and np.array_equal(q, qo)
    print(f"  {len(lv)} leaves at levels {np.bincount(lv).tolist()}, N = {g.N}", flush=True)


def test_entries(pair):
    g, o = pair
    lv, _, c = o.leaves()
    n = len(lv)
    rng = np.random.default_rng(11)
    I = rng.integers(0, g.N, 3000)
    J = rng.integers(0, g.N, 3000)
    # every near pair of 40 sampled target leaves, a few entries each
    li, lj = [], []
    for a in rng.choice(n, min(40, n), replace=False):
        t, _ = o.neighbor(np.full(n, a), np.arange(n))
        for b in np.where(t >= 0)[0]:
            li.append(a)
            lj.append(b)
    li, lj = np.array(li), np.array(lj)
    fams = set(o.neighbor(li, lj)[0].tolist())
    I = np.concatenate([I, 36 * li + rng.integers(0, 36, li.size), 36 * li + 5])
    J = np.concatenate([J, 36 * lj + rng.integers(0, 36, lj.size), 36 * lj + 5])
    v = g.entries(I, J)
    vo = o.entries(I, J)
    scale = np.maximum(np.abs(vo), 1e-14)
    err = np.max(np.abs(v - vo) / scale)
    print(f"  {I.size} entries ({li.size} near pairs, neighbour types {sorted(fams)}): max rel {err:.1e}", flush=True)
    assert err < 1e-10


def test_dense_product(pair):
    g, o = pair
    rng = np.random.default_rng(2)
    x = rng.standard_normal(g.N) + 1j * rng.standard_normal(g.N)
    y = g.dense_matvec(x)
    yo = o.dense_matvec(x)
    e = rel(y, yo)
    print(f"  dense product vs the oracle: {e:.1e}", flush=True)
    assert e < 1e-12


def test_uniform_tree_case():
    """kappa 30, contrast 3: every leaf at level 3, in box-number order, so the
    adaptive problem is the 3-level uniform one."""
    g = A.LSProblem(0, 30.0, 8, qchoice=3, tree_adaptivity=2)
    u = A.LSProblem(3, 30.0, 8, qchoice=3)
    lv, bx, _ = g.leaves()
    assert np.all(lv == 3) and np.array_equal(bx, np.arange(64))
    assert np.array_equal(g.points()[0], u.points()[0])
    rng = np.random.default_rng(4)
    I = rng.integers(0, u.N, 2000)
    J = rng.integers(0, u.N, 2000)
    assert np.array_equal(g.entries(I, J), u.entries(I, J))
    x = rng.standard_normal(u.N) + 1j * rng.standard_normal(u.N)
    assert rel(g.dense_matvec(x), u.dense_matvec(x)) < 1e-14
    g.close()
    u.close()


def test_solver_refuses_adaptive():
    g = A.LSProblem(0, 15.0, 8, qchoice=1, tree_adaptivity=2)
    with pytest.raises(A.AifmmError):
        g.assemble()
    with pytest.raises(A.AifmmError):
        g.factor(A.PREC_C64)
    with pytest.raises(A.AifmmError):
        g.tables()
    # the fields work: scattered field of a vector through the volume potential
    b = g.rhs(0.0)
    assert b.shape == (g.N,) and np.all(np.isfinite(b))
    g.close()
