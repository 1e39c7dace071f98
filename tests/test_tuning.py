"""Launch tuning: the stages of the sweep and of the FP64 operator run in the
launch shape the tuner picks (256, 512 or 1024 threads per workgroup; 1, 2,
4 or 8 right-hand-side columns per pass), or in one fixed shape for every
stage (the test modes of aifmm_ls_set_tuning).

Every size gives the same solution up to summation order.
- Complex128 direct solves agree with the default sizes to FP64 rounding
  amplified by the direct method (about 1e-11 at 5 levels, 1e-10 at 6).
- Complex64 direct solves are compared with the complex128 direct solve of
  the same problem.  Every size uses the same factors, so only the sweep's
  summation order differs.  Each size must be as accurate as the default
  sizes, within a factor of five.  The direct method amplifies complex64
  rounding: at 5 levels its own error is 3e-2 to 1.4e-1 depending on the
  summation order, and two orders differ by up to 1.6e-1; at 6 levels by
  about 3e-3.  The kernels themselves are checked against exact products,
  with no amplification, in test_gpu_programs.py.
- Refined solves agree to the refinement tolerance.
- Repeated solves with one tuning are bitwise equal.

A faulty kernel variant would show up as a much larger error in the direct
solves; flexible GMRES would hide it behind extra steps.
"""
import numpy as np
import pytest

import aifmm_ls_amd as A
from aifmm_ls_amd.workloads import gaussian_bumps

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def prob():
    print("\n[tuning] 5 levels, N = 36,864, kappa 40, 8 bumps", flush=True)
    g = A.LSProblem(5, 40.0, 7, qchoice=6, bumps=gaussian_bumps(2024, 8))
    g.assemble()
    B = np.stack([g.rhs(t) for t in np.linspace(0, 2 * np.pi, 8, endpoint=False)], axis=1)
    yield g, B
    A.set_tuning(1)
    g.close()


def solves(g, prec, mode, B):
    A.set_tuning(mode)
    g.factor(prec)  # a new factorisation tunes afresh
    x1, _ = g.solve(B[:, 0], max_refine=0)
    x1b, _ = g.solve(B[:, 0], max_refine=0)
    assert np.array_equal(x1, x1b)
    xm, _ = g.solve(B, max_refine=0)
    xr, sr = g.solve(B, tol=TOL, max_refine=40)
    assert sr[2] <= TOL
    return x1, xm, xr, g.stats()


@pytest.mark.parametrize("prec", [A.PREC_C128, A.PREC_C64], ids=["c128", "c64"])
def test_every_workgroup_size_gives_the_same_solution(prob, prec):
    g, B = prob
    x10, xm0, xr0, _ = solves(g, prec, 0, B)
    if prec == A.PREC_C64:
        # the complex128 direct solve at the default sizes, for the accuracy test
        y1, ym, _, _ = solves(g, A.PREC_C128, 0, B)
        e10, em0 = rel(x10, y1), rel(xm0, ym)
        print(f"  default sizes: complex64 direct vs complex128 direct: 1 RHS {e10:.1e}, 8 RHS {em0:.1e}", flush=True)
        A.set_tuning(0)
        g.factor(prec)
    # workgroup sizes, then columns per pass (8 at 256 threads, 1 at 1024
    # threads) on the 8-RHS solves, then the timed tuner
    for mode in (256, 512, 1024, 256 + 8 * 4096, 1024 + 1 * 4096, 1):
        x1, xm, xr, st = solves(g, prec, mode, B)
        d1, dm, dr = rel(x1, x10), rel(xm, xm0), rel(xr, xr0)
        msg = (f"  mode {mode:4d}: direct vs default sizes 1 RHS {d1:.1e}, 8 RHS {dm:.1e}; refined {dr:.1e}; "
               f"stages resized {st['sweep_tuned_stages']:.0f}, time ratio {st['sweep_tuned_time_ratio']:.3f}")
        if prec == A.PREC_C128:
            print(msg, flush=True)
            assert d1 < 1e-8 and dm < 1e-8
        else:
            e1, em = rel(x1, y1), rel(xm, ym)
            print(msg + f"; vs complex128 1 RHS {e1:.1e}, 8 RHS {em:.1e}", flush=True)
            assert e1 <= 5 * e10 + 1e-6 and em <= 5 * em0 + 1e-6
        assert dr < 1e3 * TOL
        assert 0 < st["sweep_tuned_time_ratio"] <= 1.0
        if mode not in (1, 1024):
            assert st["sweep_tuned_stages"] >= 1
