"""The source-compatible userkernel header (include/aifmm/kernel.hpp, the
reference's kernel.hpp:54-86, 130-136 surface): getMatrixEntry, getRow,
getCol and getMatrix from a C++ program built against it (tests/
shim_check.cpp), against the oracle's entries, and the points and RHS it
exposes."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "aifmm_ls_amd", "_lib")


@pytest.mark.gpu
def test_userkernel_accessors(tmp_path):
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    exe = str(tmp_path / "shim_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "shim_check.cpp"), "-o", exe, "-L", LIBDIR, "-laifmm_ls",
                    f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath-link,/opt/rocm/lib"], check=True)
    print("\n[shim] userkernel at 2 levels", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0].split()[1:] == ["576", "2"]
    ij = np.array([[int(t) for t in l.split()[:2]] for l in lines[1:]], np.int64)
    v = np.array([complex(float(l.split()[2]), float(l.split()[3])) for l in lines[1:]])
    o = O.Problem(2, 10.0, 9, 0)
    ref = o.entries(ij[:, 0], ij[:, 1])
    big = np.abs(ref) > 1e-14
    assert ij.shape[0] == 7 + 6 + 3 + 6
    assert np.all(np.abs(v - ref)[big] <= 1e-11 * np.abs(ref)[big])
    assert np.all(np.abs(v - ref)[~big] <= 1e-13)
