"""The uneven batch: 32 independent problems (levels 3-5, kappa ~ U[5,40],
tolerance 1e-6, 2-6 Gaussian bumps, one plane wave at a random angle each),
solved together with one level-synchronous launch sequence: stage i of every
problem's sweep (and of its FP64 matvec) shares one launch, and the plain FP64
refinement runs in lockstep.

Checks: every problem converges to the refinement tolerance against its own
FP64 operator; each solution equals the problem solved alone (LSProblem,
flexible GMRES) to the refinement tolerance times its conditioning; the
merged direct solve equals the per-problem direct solve (the same programs,
grouped into fewer launches) to FP64 rounding with complex128 factors; and for the level-3 members, the complex128 direct
solve of the batch agrees with the oracle's colour schedule within 10x the
compression tolerance, or 20x the oracle's own last-bit sensitivity."""
import numpy as np
import pytest

import aifmm_ls_amd as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def draw(seed=303, n=32):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nb = int(rng.integers(2, 7))
        bumps = np.column_stack([rng.uniform(0.2, 1.5, nb), rng.uniform(0.05, 0.2, nb),
                                 rng.uniform(-0.7, 0.7, nb), rng.uniform(-0.7, 0.7, nb)])
        out.append(dict(levels=int(rng.integers(3, 6)), kappa=float(rng.uniform(5, 40)), bumps=bumps,
                        theta=float(rng.uniform(0, 2 * np.pi))))
    return out


@pytest.fixture(scope="module")
def batch():
    cfg = draw()
    print(f"\n[c3] levels {[c['levels'] for c in cfg]}", flush=True)
    gs = []
    for c in cfg:
        g = A.LSProblem(c["levels"], c["kappa"], 6, qchoice=6, bumps=c["bumps"])
        g.assemble()
        gs.append(g)
    bs = [g.rhs(c["theta"]) for g, c in zip(gs, cfg)]
    yield cfg, gs, bs
    for g in gs:
        g.close()


def test_batch_refined_equals_alone(batch):
    cfg, gs, bs = batch
    for g in gs:
        g.factor(A.PREC_C64)
    xs, rr, st = A.solve_batch(gs, bs, tol=TOL, max_refine=40)
    print(f"  batch of {len(gs)}: {st[0] * 1e3:.1f} ms, {st[1]:.0f} lockstep steps, max relres {st[2]:.1e}",
          flush=True)
    assert np.all(rr <= TOL), rr
    worst = 0.0
    for g, b, x in zip(gs, bs, xs):
        r = np.linalg.norm(g.matvec(x) - b) / np.linalg.norm(b)
        assert r <= 2 * TOL
        xa, sa = g.solve(b, tol=TOL, max_refine=40)
        assert sa[2] <= TOL
        worst = max(worst, rel(x, xa))
    print(f"  worst batch vs alone {worst:.2e}", flush=True)
    assert worst < 1e3 * TOL


def test_batch_direct_equals_single_direct(batch):
    cfg, gs, bs = batch
    # complex128 factors: the merged launches must reproduce each problem's
    # own sweep up to FP64 rounding
    sub = list(range(8))
    for i in sub:
        gs[i].factor(A.PREC_C128)
    xs, _, _ = A.solve_batch([gs[i] for i in sub], [bs[i] for i in sub], max_refine=0)
    w64 = max(rel(x, gs[i].solve(bs[i], max_refine=0)[0]) for i, x in zip(sub, xs))
    # complex64 factors: a merged stage of more than 1024 programs runs
    # 256-thread workgroups, which sum in another order; float rounding is
    # then amplified like any perturbation of the direct method (whose own
    # error is about 1e-2 here)
    for g in gs:
        g.factor(A.PREC_C64)
    xs, _, _ = A.solve_batch(gs, bs, max_refine=0)
    w32 = max(rel(x, g.solve(b, max_refine=0)[0]) for g, b, x in zip(gs, bs, xs))
    print(f"  direct, merged vs single: complex128 {w64:.1e}, complex64 {w32:.1e}", flush=True)
    assert w64 < 1e-9
    assert w32 < 1e-2


def test_batch_fp64_direct_vs_oracle(batch):
    cfg, gs, bs = batch
    small = [i for i, c in enumerate(cfg) if c["levels"] == 3][:4]
    assert small
    for i in small:
        gs[i].factor(A.PREC_C128)
    xs, _, _ = A.solve_batch([gs[i] for i in small], [bs[i] for i in small], max_refine=0)
    for i, x in zip(small, xs):
        c = cfg[i]
        o = O.Problem(3, c["kappa"], 6, qchoice=6, bumps=c["bumps"])
        bo = o.rhs(c["theta"])
        assert rel(bs[i], bo) < 1e-14
        xo, _ = o.solve(bs[i], O.SCHED_COLORED)
        sens = 0.0
        for seed in (1, 2):
            xp, _ = O.Problem(3, c["kappa"], 6, qchoice=6, bumps=c["bumps"], perturb=1e-15,
                              perturb_seed=seed).solve(bs[i], O.SCHED_COLORED)
            sens = max(sens, rel(xp, xo))
        d = rel(x, xo)
        print(f"  problem {i}: GPU vs oracle colour {d:.1e} (sensitivity {sens:.1e})", flush=True)
        assert d < max(10 * 1e-6, 20 * sens)


@pytest.mark.parametrize("prec", [A.PREC_C128, A.PREC_C64], ids=["c128", "c64"])
def test_merged_factorisations_equal_alone(batch, prec):
    """aifmm_ls_factor_batch: the eliminations of several problems of levels
    3, 4 and 5 as one merged launch sequence.  Every descriptor of a merged
    launch is computed as it is alone, so each problem's recompression
    ranks, its zeroed pivots and its direct solve (launch tuning off, so
    both sweeps sum in one order) must equal those of its own
    factorisation bitwise."""
    cfg, gs, bs = batch
    sub = [i for lv in (3, 4, 5) for i in [k for k, c in enumerate(cfg) if c["levels"] == lv][:3]]
    A.set_tuning(0)
    try:
        alone = []
        for i in sub:
            gs[i].factor(prec)
            x, _ = gs[i].solve(bs[i], max_refine=0)
            alone.append((x, gs[i].factor_trace(0), gs[i].factor_trace(1)))
        info = A.factor_batch([gs[i] for i in sub], prec)
        print(f"  merged factorisation of {len(sub)} problems: {info}", flush=True)
        assert info["rounds"] > 0 and info["launches"] > 0
        for i, (xa, ta, za) in zip(sub, alone):
            x, _ = gs[i].solve(bs[i], max_refine=0)
            assert np.array_equal(gs[i].factor_trace(0), ta)
            assert np.array_equal(gs[i].factor_trace(1), za)
            assert np.array_equal(x, xa), (i, rel(x, xa))
        # refined from the merged factors, as from any factorisation
        xr, sr = gs[sub[-1]].solve(bs[sub[-1]], tol=TOL, max_refine=40)
        assert sr[2] <= TOL
    finally:
        A.set_tuning(1)


def test_merged_factorisation_refusals(batch):
    cfg, gs, bs = batch
    with pytest.raises(A.AifmmError):
        A.factor_batch([gs[0], gs[0]])
    with pytest.raises(A.AifmmError):
        A.factor_batch(gs[:2], precision=5)


def test_batch_rejects_mixed_precision(batch):
    cfg, gs, bs = batch
    gs[0].factor(A.PREC_C64)
    gs[1].factor(A.PREC_C128)
    with pytest.raises(A.AifmmError):
        A.solve_batch(gs[:2], bs[:2])
    with pytest.raises(A.AifmmError):
        A.solve_batch([gs[0], gs[0]], [bs[0], bs[0]])
