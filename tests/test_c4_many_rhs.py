"""Many right-hand sides at the full size: one factorisation at 7 levels
(N = 589,824), kappa 60, tolerance 1e-6, a smooth 4-bump contrast, complex64
factors, then 256 plane waves at equally spaced angles in [0, 2 pi) in one
FP64-refined solve.

No oracle runs at this size; the checks are ones that hold at any size:
every column reaches the refinement tolerance against the FP64 operator
(checked again from the returned solution with an independent matvec), the
columns equal the same right-hand sides solved alone to the tolerance times
the conditioning, and the solve is linear in b."""
import numpy as np
import pytest

import aifmm_ls_amd as A

pytestmark = pytest.mark.gpu

TOL = 1e-9


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_c4_256_plane_waves():
    rng = np.random.default_rng(2024)
    bm = np.c_[rng.uniform(0.2, 1.0, 4), rng.uniform(0.15, 0.25, 4), rng.uniform(-0.5, 0.5, (4, 2))]
    print("\n[c4] 7 levels, N = 589,824, 256 right-hand sides", flush=True)
    g = A.LSProblem(7, 60.0, 6, qchoice=6, bumps=bm)
    g.assemble()
    g.factor(A.PREC_C64)
    n = 256
    B = np.stack([g.rhs(t) for t in 2 * np.pi * np.arange(n) / n], axis=1)
    X, st = g.solve(B, tol=TOL, max_refine=40)
    print(f"  solve {st[0]:.2f} s, {st[1]:.0f} steps, max relres {st[2]:.1e}", flush=True)
    assert X.shape == (g.N, n)
    assert st[2] <= TOL
    cols = [0, 37, 128, 255]
    R = g.matvec(X[:, cols]) - B[:, cols]
    assert np.linalg.norm(R, axis=0).max() / np.linalg.norm(B[:, cols], axis=0).min() <= 2 * TOL
    for c in (0, 128):
        xa, sa = g.solve(B[:, c], tol=TOL, max_refine=40)
        assert sa[2] <= TOL
        assert rel(X[:, c], xa) < 1e3 * TOL
    # linearity: b = B0 + 2i B1 gives x = X0 + 2i X1
    xl, sl = g.solve(B[:, 0] + 2j * B[:, 1], tol=TOL, max_refine=40)
    assert rel(xl, X[:, 0] + 2j * X[:, 1]) < 1e3 * TOL
    g.close()
