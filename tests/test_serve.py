"""Serving loop (aifmm_ls_server_*): each result is the same alone and in a mix
of concurrent requests, queued solves of one problem are merged, queued
one-column solves of different problems run as one batch, problems are
admitted by device memory with least-recently-used eviction, and a solve for
an evicted problem factorises it again.

"Alone" is the same request through LSProblem (one handle, one solve at a
time).  Both are FP64-refined to the same relative residual, so they agree to
the refinement tolerance times the problem's conditioning, not bitwise: a
merged solve runs its right-hand sides through the multi-RHS kernels."""
import numpy as np
import pytest

import aifmm_ls_amd as A

pytestmark = pytest.mark.gpu

TOL = 1e-10


def bumps(seed, n):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(0.2, 1.5, n), rng.uniform(0.05, 0.2, n), rng.uniform(-0.7, 0.7, n),
                            rng.uniform(-0.7, 0.7, n)])


PROBLEMS = [dict(n_levels=3, kappa=12.0, tol_pow=8, qchoice=6, bumps=bumps(1, 3)),
            dict(n_levels=4, kappa=25.0, tol_pow=7, qchoice=6, bumps=bumps(2, 4)),
            dict(n_levels=4, kappa=8.0, tol_pow=8, qchoice=0)]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def alone(cfg):
    g = A.LSProblem(**cfg)
    g.assemble()
    g.factor(A.PREC_C64)
    return g


def test_mix_equals_alone_and_requests_overlap():
    print("\n[serve] mix vs alone", flush=True)
    ref = [alone(c) for c in PROBLEMS]
    s = A.Server(64 << 30, slots=3)
    pids = [s.factor(precision=A.PREC_C64, **c) for c in PROBLEMS]
    for _, t in pids:
        s.wait(t)
    rng = np.random.default_rng(7)
    reqs = []
    for i in range(12):
        p = i % 3
        k = int(rng.choice([1, 2, 3, 5]))
        B = np.stack([ref[p].rhs(th) for th in rng.uniform(0, 2 * np.pi, k)], axis=1)
        reqs.append((p, B, s.solve(pids[p][0], B, tol=TOL, max_refine=40)))
    worst = 0.0
    for p, B, t in reqs:
        x, info = s.wait(t)
        assert info["relres"] <= TOL
        xa, st = ref[p].solve(B, tol=TOL, max_refine=40)
        assert st[2] <= TOL
        worst = max(worst, rel(x, xa))
        # the residual against the FP64 operator, from the result itself
        r = np.linalg.norm(ref[p].matvec(x) - B) / np.linalg.norm(B)
        assert r <= 2 * TOL
    st = s.stats()
    print(f"  worst mix vs alone {worst:.2e}; batches {st['batches']:.0f}, merged {st['merged_requests']:.0f}, "
          f"peak running {st['peak_running']:.0f}", flush=True)
    assert worst < 1e3 * TOL
    assert st["peak_running"] >= 2
    assert st["completed"] == 15
    s.close()
    for g in ref:
        g.close()


def test_same_problem_solves_are_merged():
    cfg = PROBLEMS[1]
    ref = alone(cfg)
    s = A.Server(64 << 30, slots=1)
    pid, t = s.factor(precision=A.PREC_C64, **cfg)
    # queue behind the factorisation, so all are waiting when the slot frees
    Bs = [ref.rhs(0.3 * i)[:, None] for i in range(6)]
    ts = [s.solve(pid, B, tol=TOL, max_refine=40) for B in Bs]
    s.wait(t)
    rhs_counts = []
    for B, tt in zip(Bs, ts):
        x, info = s.wait(tt)
        rhs_counts.append(info["batch_rhs"])
        xa, _ = ref.solve(B, tol=TOL, max_refine=40)
        assert rel(x, xa) < 1e3 * TOL
    assert max(rhs_counts) == 6, rhs_counts
    assert s.stats()["merged_requests"] >= 5
    s.close()
    ref.close()


def test_one_column_solves_of_different_problems_batch():
    """One slot, busy with a 64-column solve while one-column solves of the
    three problems queue (the one of problem 0 cannot join the 64-column
    solve: 64 columns is the merge limit): when the slot frees, they run as
    one batch solve (merged launch sequences, lockstep refinement), and each
    result equals the same solve alone."""
    print("\n[serve] cross-problem batch", flush=True)
    ref = [alone(c) for c in PROBLEMS]
    s = A.Server(64 << 30, slots=1)
    pids = [s.factor(precision=A.PREC_C64, **c)[0] for c in PROBLEMS]
    B0 = np.stack([ref[0].rhs(0.1 * i) for i in range(64)], axis=1)
    t0 = s.solve(pids[0], B0, tol=TOL, max_refine=40)
    bs = [ref[p].rhs(1.0 + p) for p in (1, 2, 0)]
    ts = [s.solve(pids[p], b, tol=TOL, max_refine=40) for p, b in zip((1, 2, 0), bs)]
    s.wait(t0)
    worst, sizes = 0.0, []
    for p, b, t in zip((1, 2, 0), bs, ts):
        x, info = s.wait(t)
        sizes.append(info["batch_problems"])
        assert info["relres"] <= TOL
        xa, _ = ref[p].solve(b, tol=TOL, max_refine=40)
        worst = max(worst, rel(x, xa))
    st = s.stats()
    print(f"  batch sizes {sizes}, cross-problem batches {st['cross_problem_batches']:.0f}, "
          f"worst batch vs alone {worst:.2e}", flush=True)
    assert max(sizes) == 3 and st["cross_problem_batches"] >= 1
    assert worst < 1e3 * TOL
    s.close()
    for g in ref:
        g.close()


def test_admission_by_memory_evicts_lru_and_refactors():
    cfg = PROBLEMS[2]
    g = alone(cfg)
    b = g.rhs(0.0)
    xa, _ = g.solve(b, tol=TOL, max_refine=40)
    one = g.device_bytes()
    g.close()
    print(f"\n[serve] one L=4 problem holds {one / 2**20:.1f} MiB", flush=True)
    # room for two problems of this shape, not three
    s = A.Server(int(2.5 * one), slots=2)
    p1, t1 = s.factor(precision=A.PREC_C64, **cfg)
    s.wait(t1)
    x1, _ = s.wait(s.solve(p1, b, tol=TOL, max_refine=40))
    assert rel(x1, xa) < 1e3 * TOL
    p2, t2 = s.factor(precision=A.PREC_C64, **cfg)
    s.wait(t2)
    st = s.stats()
    assert st["resident_problems"] == 2 and st["evictions"] == 0, st
    p3, t3 = s.factor(precision=A.PREC_C64, **cfg)
    s.wait(t3)
    st = s.stats()
    assert st["evictions"] == 1 and st["resident_problems"] == 2, st
    assert st["resident_bytes"] <= st["budget"]
    # p1 was the least recently used: it is gone, and a solve rebuilds it
    x1b, info = s.wait(s.solve(p1, b, tol=TOL, max_refine=40))
    assert info["refactored"]
    assert rel(x1b, xa) < 1e3 * TOL
    st = s.stats()
    assert st["refactorisations"] == 1 and st["resident_bytes"] <= st["budget"], st
    # a released problem refuses new work
    s.release(p2)
    with pytest.raises(A.AifmmError):
        s.solve(p2, b)
    s.close()


def test_requests_that_cannot_fit_fail_cleanly():
    s = A.Server(1 << 20, slots=1)
    pid, t = s.factor(precision=A.PREC_C64, **PROBLEMS[0])
    # the first problem of a shape is admitted on its estimate, capped at the
    # budget; its measured footprint then exceeds a 1 MiB budget
    s.wait(t)
    with pytest.raises(A.AifmmError):
        s.wait(s.solve(pid, np.ones(64 * 36, np.complex128)))
    with pytest.raises(A.AifmmError):
        s.solve(pid, np.ones(10, np.complex128))
    s.close()
