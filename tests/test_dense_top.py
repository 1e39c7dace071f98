"""The top of the one-column sweep as one dense matrix (Factor::dense_top_*):
formed at factor time by running the sweep's top levels on identity columns,
it must give the staged sweep's result to working-precision rounding.  Sweeps
built for several columns keep the staged form, so a one-column and a
two-column direct solve of the same right-hand side compare the two forms in
one process."""
import numpy as np
import pytest

import aifmm_ls_amd as A
from aifmm_ls_amd.workloads import gaussian_bumps

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("levels,expect", [(3, 0), (4, 2), (5, 3)])
def test_dense_top_level(levels, expect):
    g = A.LSProblem(levels, 20.0, 7, qchoice=6, bumps=gaussian_bumps(7, 4))
    g.assemble()
    g.factor(A.PREC_C64)
    st = g.stats()
    print(f"levels {levels}: dense top level {st['dense_top_level']}, n {st['dense_top_n']}", flush=True)
    assert st["dense_top_level"] == expect
    assert (st["dense_top_n"] > 0) == (expect > 0)
    g.close()


# direct sweeps (no refinement): dense one-column against staged two-column;
# bounds: complex128 rounding (observed far below), complex64 rounding
# amplified by the top levels' conditioning
@pytest.mark.parametrize("prec,bound", [(A.PREC_C128, 1e-8), (A.PREC_C64, 1e-3)])
def test_dense_top_matches_staged_sweep(prec, bound):
    g = A.LSProblem(5, 20.0, 7, qchoice=6, bumps=gaussian_bumps(7, 4))
    g.assemble()
    g.factor(prec)
    assert g.stats()["dense_top_level"] == 3
    b = g.rhs(0.3)
    x1, _ = g.solve(b, tol=1e-7, max_refine=0)
    x2, _ = g.solve(np.stack([b, b], axis=1), tol=1e-7, max_refine=0)
    d = [float(np.linalg.norm(x1 - x2[:, c]) / np.linalg.norm(x2[:, c])) for c in range(2)]
    print(f"precision {prec}: dense vs staged sweep {d[0]:.2e} {d[1]:.2e}", flush=True)
    assert d[0] < bound and d[1] < bound
    assert np.array_equal(x2[:, 0], x2[:, 1])
    g.close()


def test_dense_top_refined_solution_against_fp64():
    g = A.LSProblem(5, 20.0, 7, qchoice=6, bumps=gaussian_bumps(7, 4))
    g.assemble()
    g.factor(A.PREC_C128)
    b = g.rhs(0.3)
    xr, _ = g.solve(b, tol=1e-13, max_refine=60)
    g.factor(A.PREC_C64)
    x, st = g.solve(b, tol=1e-7, max_refine=40)
    err = float(np.linalg.norm(x - xr) / np.linalg.norm(xr))
    print(f"refined complex64 with dense top: {int(st[1])} steps, rel.err vs FP64 {err:.2e}", flush=True)
    assert err < 1e-6
    g.close()


def test_one_column_program_kept_after_wider_solves():
    """A one-column solve after a wider one runs the one-column program again
    (its 16-row tiles and the dense top), not the wider staged one: the result
    is bitwise the first one-column solve's."""
    g = A.LSProblem(5, 20.0, 7, qchoice=6, bumps=gaussian_bumps(7, 4))
    g.assemble()
    g.factor(A.PREC_C64)
    b = g.rhs(0.3)
    x1, _ = g.solve(b, tol=1e-7, max_refine=0)
    xw, _ = g.solve(np.stack([b] * 80, axis=1), tol=1e-7, max_refine=0)  # above the dense top's 64 columns
    d = float(np.linalg.norm(x1 - xw[:, 0]) / np.linalg.norm(x1))
    x1b, _ = g.solve(b, tol=1e-7, max_refine=0)
    x4, _ = g.solve(np.stack([b] * 4, axis=1), tol=1e-7, max_refine=0)
    x1c, _ = g.solve(b, tol=1e-7, max_refine=0)
    print(f"one column after 80: dense vs staged {d:.2e}", flush=True)
    assert np.array_equal(x1, x1b) and np.array_equal(x1, x1c)
    assert d < 1e-3
    # the 80-column program again (its column grouping may differ with the width)
    assert all(np.linalg.norm(x4[:, c] - xw[:, 0]) <= 1e-10 * np.linalg.norm(xw[:, 0]) for c in range(4))
    g.close()
