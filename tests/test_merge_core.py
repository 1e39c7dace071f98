"""The lockstep merge of batched launches (aifmm_ls_amd/csrc/merge.h, used by
aifmm_ls_factor_batch) on the host, with stub launches in place of kernels
(tests/merge_stub/merge_check.cpp, built here with g++ together with
merge.cpp).

Seeded runs of 1-40 member threads submit sequences of batches of three
kinds (some empty); some members stop early and one launch fails.  Checked:
every run ends; every submitted batch is launched exactly once and in its
member's order, one per round; each kind is launched at most once per round;
exactly the members whose batches were in the failed launch see its error,
and the others finish their sequences."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("merge") / "libmerge_check.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-shared", "-fPIC", "-Wall",
                    os.path.join(HERE, "merge_stub", "merge_check.cpp"),
                    os.path.join(ROOT, "aifmm_ls_amd", "csrc", "merge.cpp"), "-o", out], check=True)
    L = ctypes.CDLL(out)
    L.merge_check.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    return L


@pytest.mark.parametrize("members", [1, 2, 5, 17, 40])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_lockstep_merge(lib, members, seed):
    out = np.zeros(7)
    lib.merge_check(seed, members, 60, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    violations, rounds, launches, launched, submitted, failing_round, threw = out
    assert violations == 0, out
    # the group does not count the launch that failed
    assert launches == launched - (1 if failing_round >= 0 else 0)
    # merging: never more rounds than the longest member sequence allows
    assert rounds <= 60
    if members > 3:
        assert failing_round >= 0 and threw >= 1
    else:
        assert threw == 0
    if members >= 17:
        assert launched < submitted  # batches of several members shared launches
