"""The zero-pivot decisions that separate the GPU factorisation's default rules
from Eigen's (the reference's ColPivHouseholderQR, FMM2DTree_find_error.hpp
:2092 for L2M_f and :2477 for the base system), pinned on the oracle alone.

A pivot counts as zero below eps x the largest column norm.  Eigen uses the
FP64 epsilon; the GPU's complex128 factorisation uses 1e-12 for the L2M_f
pivot blocks.  The GPU parity tests compare with the oracle under the same
rules; these tests show which decision each rule changes, by how much, and
that under Eigen's rule the reference's result at those inputs is fixed by
rounding rather than by the input."""
import numpy as np


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def c2_bumps(seed=2024, n=8):
    rng = np.random.default_rng(seed)
    return np.c_[rng.uniform(0.2, 1.5, n), rng.uniform(0.05, 0.2, n), rng.uniform(-0.7, 0.7, (n, 2))]


def test_l2mf_rule_at_the_reference_run(oracle):
    """`6 10 6 1 0 8 5` (5 levels, kappa 10, Qchoice 0, TOL_POW 8): the 1e-12
    rule zeroes one pivot of one block, the 2 x 2 L2M_f of leaf 522, and that
    one decision is the whole difference to Eigen's rule (2.4e-4).  Under
    Eigen's rule a 1e-15 change of the far-field entries changes the
    recompression decisions that follow; under the 1e-12 rule it does not."""
    O = oracle
    eig = O.Problem(5, 10.0, 8, 0)
    b = eig.rhs(0.0)
    xe, _ = eig.solve(b, O.SCHED_COLORED)
    te = eig.recompress_trace()
    assert eig.zero_pivots().size == 0
    rule = O.Problem(5, 10.0, 8, 0, piv_eps=1e-12)
    xr, _ = rule.solve(b, O.SCHED_COLORED)
    tr = rule.recompress_trace()
    np.testing.assert_array_equal(rule.zero_pivots(), [[5, 522, 1, 2]])
    d = rel(xr, xe)
    print(f"\n  Eigen's rule vs 1e-12 at leaf 522: {d:.3e}")
    assert 1e-5 < d < 1e-3
    pe = O.Problem(5, 10.0, 8, 0, perturb=1e-15, perturb_seed=1)
    pe.solve(b, O.SCHED_COLORED)
    pr = O.Problem(5, 10.0, 8, 0, piv_eps=1e-12, perturb=1e-15, perturb_seed=1)
    xp, _ = pr.solve(b, O.SCHED_COLORED)
    assert not np.array_equal(pe.recompress_trace(), te)
    np.testing.assert_array_equal(pr.recompress_trace(), tr)
    assert rel(xp, xr) < 1e-8


def test_singular_base_system_under_eigens_rule(oracle):
    """3 levels, kappa 40, tol 1e-7, the 8-bump contrast of the N = 147,456
    workload: the level-2 base system is singular to rounding (smallest over
    largest |R_ii| below 1e-12), so under Eigen's rule the solution moves by
    about 2e-6 under a 1e-15 change of the far-field entries, and by about as
    much when only L2P / P2M change in the last bit.  Dropping the base
    pivots below 1e-12 makes the solution continuous again."""
    O = oracle
    kw = dict(bumps=c2_bumps(), piv_eps=1e-12)

    def sens(base_eps):
        o = O.Problem(3, 40.0, 7, 6, base_piv_eps=base_eps, **kw)
        b = o.rhs(0.0)
        x0, st = o.solve(b, O.SCHED_COLORED)
        d = 0.0
        for seed in (1, 2):
            p = O.Problem(3, 40.0, 7, 6, base_piv_eps=base_eps, perturb=1e-15, perturb_seed=seed, **kw)
            xp, _ = p.solve(b, O.SCHED_COLORED)
            d = max(d, rel(xp, x0))
        return d, st[7], st[6]

    s_eig, ratio, n = sens(0.0)
    s_rule, _, _ = sens(1e-12)
    print(f"\n  base n = {n:.0f}, |R_ii| ratio {ratio:.1e}; sensitivity Eigen {s_eig:.2e}, 1e-12 rule {s_rule:.2e}")
    assert ratio < 1e-12
    assert s_eig > 1e-7
    assert s_rule < 1e-7
