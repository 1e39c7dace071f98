"""The serving loop's scheduling core (aifmm_ls_amd/csrc/serve_core.h) on the
host, with stub problem handles in place of GPU factorisations
(tests/serve_stub/stub_server.cpp, built here with g++).

Seeded traces from several submitting threads mix factorisations, solves of
1-9 right-hand sides and releases against a device-memory budget of a few
problems, so evictions, re-factorisations, merged and cross-problem batches
and destroys of evicted handles all happen while other slots work.  Stub
destroys sleep, as device frees synchronise.  Checked: every request ends;
results are right; no handle is used by two slots at once, used after its
destroy or destroyed while in use; the live stub bytes never exceed the
budget; and submission never waits for a destroy (destroys run outside the
scheduler's lock)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def stub(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("stub") / "libstub_server.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-shared", "-fPIC", "-Wall",
                    os.path.join(HERE, "serve_stub", "stub_server.cpp"), "-o", out], check=True)
    L = ctypes.CDLL(out)
    L.stub_trace.argtypes = [ctypes.c_int] * 7 + [ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    return L


def run(L, seed, nreq=60, slots=4, nsub=4, build_ms=4, solve_ms=2, destroy_ms=25, budget_problems=3):
    out = np.zeros(15)
    L.stub_trace(seed, nreq, slots, nsub, build_ms, solve_ms, destroy_ms, budget_problems,
                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    keys = ["completed", "failed", "max_submit_ms", "peak_bytes", "budget", "evictions", "violations",
            "cross_batches", "wrong", "builds", "destroys", "live_bytes_end", "batch_calls", "refused", "peak_builds"]
    return dict(zip(keys, out.tolist()))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_trace_with_evictions_and_releases(stub, seed):
    r = run(stub, seed)
    print(f"\n  seed {seed}: {r}")
    assert r["violations"] == 0
    assert r["wrong"] == 0
    assert r["failed"] == 0
    assert r["evictions"] > 0
    assert r["peak_bytes"] <= r["budget"]
    assert r["builds"] == r["destroys"] and r["live_bytes_end"] == 0
    # a destroy sleeps 25 ms: a submission that waited for one would show it
    assert r["max_submit_ms"] < 15


def test_one_slot_and_tight_budget(stub):
    """One slot, room for one problem: every switch of problem evicts."""
    r = run(stub, 7, nreq=40, slots=1, nsub=3, budget_problems=1)
    print(f"\n  {r}")
    assert r["violations"] == 0 and r["wrong"] == 0 and r["failed"] == 0
    assert r["peak_bytes"] <= r["budget"]


def test_cross_problem_batches_form(stub):
    """Slow solves and many slots' worth of one-column requests queue up, so
    one-column solves of different resident problems run as one batch."""
    r = run(stub, 11, nreq=80, slots=2, nsub=6, solve_ms=8, destroy_ms=5, budget_problems=500)
    print(f"\n  {r}")
    assert r["violations"] == 0 and r["wrong"] == 0 and r["failed"] == 0
    assert r["cross_batches"] > 0 and r["batch_calls"] == r["cross_batches"]


def test_factorisations_leave_slots_to_solves(stub):
    """Four slots, slow builds: at most two slots factorise at once (the
    others keep serving resident problems in arrival order), and every
    request still ends."""
    r = run(stub, 5, nreq=80, slots=4, nsub=4, build_ms=20, solve_ms=2, destroy_ms=5, budget_problems=6)
    print(f"\n  {r}")
    assert r["violations"] == 0 and r["wrong"] == 0 and r["failed"] == 0
    assert 1 <= r["peak_builds"] <= 2
