"""Dense product y = A x with the discretised operator itself
(aifmm_ls_dense_matvec: the generator's rhs = A x over all leaf pairs,
FMM2DTree_gen_rhs_x.hpp:2808-2903).

Against the oracle's dense matrix (every entry as getMatrixEntry, zeroed below
1e-16) times x at 2 and 3 levels; at N = 147,456 against full rows of entries
(aifmm_ls_entries) and by linearity.  Bar: far entries are summed per source
leaf through its moments and not zeroed one by one, so a row may differ from
the entry-by-entry sum by rounding plus at most 1e-16 |x_j| per far entry; the
tests hold 1e-12 relative in the 2-norm and, per row, 1e-13 of the row's
absolute sum plus 1e-16 sum |x_j|."""
import numpy as np
import pytest

import aifmm_ls_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def c2_bumps(seed=2024, n=8):
    rng = np.random.default_rng(seed)
    return np.c_[rng.uniform(0.2, 1.5, n), rng.uniform(0.05, 0.2, n), rng.uniform(-0.7, 0.7, (n, 2))]


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("L,kap,tp,qc,bumps", [(2, 10.0, 8, 0, None), (3, 20.0, 8, 6, [[1.0, 0.15, 0.1, 0.2]]),
                                               (3, 40.0, 7, 6, c2_bumps())])
def test_dense_product_against_oracle_matrix(L, kap, tp, qc, bumps):
    print(f"\n[case] dense product L={L} kappa={kap}", flush=True)
    g = A.LSProblem(L, kap, tp, qc, bumps=bumps)
    o = O.Problem(L, kap, tp, qc, bumps=bumps)
    Ad = o.dense()
    rng = np.random.default_rng(L)
    X = rand_c(rng, g.N, 2)
    Y = g.dense_matvec(X)
    Yr = Ad @ X
    d = max(rel(Y[:, c], Yr[:, c]) for c in range(2))
    # the solution-like vector the generator multiplies: the plane-wave solve
    xs = np.linalg.solve(Ad, o.rhs(0.0))
    ds = rel(g.dense_matvec(xs), Ad @ xs)
    print(f"  random x {d:.2e}, solution x {ds:.2e}", flush=True)
    assert d < 1e-12 and ds < 1e-12
    g.close()


def test_dense_product_rows_and_linearity_at_c2_size():
    """N = 147,456, kappa 40, the 8-bump contrast: six full rows against
    entry-by-entry sums, and y(x1 + 2i x2) = y(x1) + 2i y(x2)."""
    print("\n[case] dense product N = 147,456", flush=True)
    g = A.LSProblem(6, 40.0, 7, 6, bumps=c2_bumps())
    rng = np.random.default_rng(5)
    x1, x2 = rand_c(rng, g.N), rand_c(rng, g.N)
    Y = g.dense_matvec(np.stack([x1, x2, x1 + 2j * x2], axis=1))
    assert rel(Y[:, 2], Y[:, 0] + 2j * Y[:, 1]) < 1e-13
    cols = np.arange(g.N, dtype=np.int64)
    worst = 0.0
    for i in rng.integers(0, g.N, 6):
        e = g.entries(np.full(g.N, i, np.int64), cols)
        s = np.sum(e * x1)
        bound = 1e-13 * np.sum(np.abs(e * x1)) + 1e-16 * np.sum(np.abs(x1))
        worst = max(worst, abs(Y[i, 0] - s) / bound)
        assert abs(Y[i, 0] - s) <= bound, (i, Y[i, 0], s)
    print(f"  rows: worst difference {worst:.2f} of the bound", flush=True)
    g.close()
