"""Box programs in every launch shape against numpy.

One stage of random programs, y_p = S_p + sum_t alpha_t A_{p,t} X_{p,t} and
out_p = B_p y_p (aifmm_ls_selftest_prog), run by k_prog in the default shape
and in every workgroup size (256, 512, 1024 threads) and count of
right-hand-side columns per pass (1, 2, 4, 8), in complex128 and complex64.
The cases cover one-column and many-column runs, boxes below and above the
tiling height (16 rows for a program built for one column, 64 for several)
and above one row block (256 rows), terms shorter and longer than the column
slices, and more terms than a program slot holds.

The products are exact linear algebra on random data, with no amplification:
complex128 must match to 1e-12, complex64 to 1e-4 relative."""
import numpy as np
import pytest

import aifmm_ls_amd as A

pytestmark = pytest.mark.gpu

SHAPES = [0, 256, 512, 1024] + [nth + 4096 * ch for nth in (256, 512) for ch in (1, 2, 4, 8)] + [1024 + 4096 * 2]
CASES = [  # nprog, m, nterms, cols, nrhs
    (40, 12, 3, 36, 1),
    (40, 36, 25, 36, 1),
    (9, 300, 4, 70, 1),
    (64, 36, 6, 5, 1),
    (30, 36, 7, 36, 9),
    (7, 150, 20, 24, 16),
]


def reference(Ab, X, S, Bm):
    alpha = np.array([1.0 if t % 2 == 0 else -0.5 + 0.25j for t in range(Ab.shape[1])])
    Y = S + np.einsum("t,ptmc,ptcr->pmr", alpha, Ab, X)
    return Y, np.einsum("pij,pjr->pir", Bm, Y)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("case", CASES, ids=[f"p{c[0]}_m{c[1]}_t{c[2]}_c{c[3]}_r{c[4]}" for c in CASES])
def test_box_programs_match_numpy_in_every_shape(case):
    nprog, m, nterms, cols, nrhs = case
    rng = np.random.default_rng(sum(case))
    cx = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    Ab, X, S, Bm = cx(nprog, nterms, m, cols), cx(nprog, nterms, cols, nrhs), cx(nprog, m, nrhs), cx(nprog, m, m)
    Yr, Or = reference(Ab, X, S, Bm)
    worst = {0: 0.0, 1: 0.0}
    for c64 in (False, True):
        for shape in SHAPES:
            if shape // 4096 > 1 and nrhs == 1:
                continue  # several columns per pass need several right-hand sides
            Y, O = A.selftest_program(Ab, X, S, Bm, shape=shape, c64=c64)
            ey, eo = rel(Y, Yr), rel(O, Or)
            worst[c64] = max(worst[c64], ey, eo)
            assert ey < (1e-4 if c64 else 1e-12) and eo < (1e-4 if c64 else 1e-12), (shape, c64, ey, eo)
    print(f"\n  {case}: worst complex128 {worst[0]:.1e}, complex64 {worst[1]:.1e}", flush=True)
