"""The growth gate of the complex64 factorisation's pivot blocks (the self
blocks P2P[k][k] and L2M_f, which the reference factors in FP64 at
FMM2DTree_find_error.hpp:1819, 2092): each block keeps its complex64 LU when
its growth factor times its condition estimate is at most the threshold, and
is factored in FP64 otherwise.

Checks on a 4-level problem with the 8-bump contrast at kappa 40:
- threshold 0 sends every block through the gated FP64 path, which must give
  the ungated FP64 path's direct solve bitwise;
- an infinite threshold keeps every block that fits the LDS in complex64; its
  direct solve is no worse than the FP64-pivot one by more than a factor of
  three, measured against the FP64 method's refined solution;
- the default threshold (1e4) splits the blocks, and its refined solution
  matches the FP64 method's to the refinement tolerance."""
import numpy as np
import pytest

import aifmm_ls_amd as A
from aifmm_ls_amd.workloads import gaussian_bumps

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_growth_gate_paths():
    print("\n[gate] 4 levels, kappa 40, 8 bumps", flush=True)
    g = A.LSProblem(4, 40.0, 7, qchoice=6, bumps=gaussian_bumps(2024, 8))
    b = g.rhs(0.3)
    g.factor(A.PREC_C128)
    xr, st = g.solve(b, tol=1e-13, max_refine=60)
    assert st[2] < 1e-12
    A.set_tuning(0)  # one launch shape, so sweeps of equal factors sum alike
    try:
        g.set_growth_gate(-1)
        g.factor(A.PREC_C64)
        s_off = g.stats()
        x_off, _ = g.solve(b)
        assert s_off["pivot_blocks_c64"] == 0 and s_off["pivot_blocks_fp64"] == 0
        g.set_growth_gate(0)
        g.factor(A.PREC_C64)
        s0 = g.stats()
        x0, _ = g.solve(b)
        total = s0["pivot_blocks_c64"] + s0["pivot_blocks_fp64"]
        assert s0["pivot_blocks_c64"] == 0 and total > 0
        assert np.array_equal(x0, x_off)
        g.set_growth_gate(float("inf"))
        g.factor(A.PREC_C64)
        si = g.stats()
        xi, _ = g.solve(b)
        assert si["pivot_blocks_c64"] + si["pivot_blocks_fp64"] == total
        assert si["pivot_blocks_c64"] > 0.9 * total
        e_off, e_inf = rel(x_off, xr), rel(xi, xr)
        g.set_growth_gate(1e4)
        g.factor(A.PREC_C64)
        sd = g.stats()
        xd, std = g.solve(b, tol=1e-10, max_refine=40)
        print(f"  pivot blocks {total:.0f}: complex64 at threshold inf {si['pivot_blocks_c64']:.0f}, at 1e4 "
              f"{sd['pivot_blocks_c64']:.0f}; direct error vs FP64 method: FP64 pivots {e_off:.2e}, complex64 "
              f"pivots {e_inf:.2e}; refined at 1e4: {std[1]:.0f} steps, {rel(xd, xr):.2e}", flush=True)
        assert e_inf < 3 * e_off + 1e-6
        assert sd["pivot_blocks_c64"] + sd["pivot_blocks_fp64"] == total
        assert std[2] <= 1e-10
        assert rel(xd, xr) < 1e-8
    finally:
        A.set_tuning(1)
        g.set_growth_gate(1e4)
        g.close()
