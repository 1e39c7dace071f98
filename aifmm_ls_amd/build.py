"""Build libaifmm_ls.so in-tree for gfx950 with hipcc (no JIT, no torch).

    python -m aifmm_ls_amd.build        # or aifmm_ls_amd.build.build()
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
OBJ = os.path.join(OUT, "obj")
LIB = os.path.join(OUT, "libaifmm_ls.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AIFMM_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
SOURCES = ["k_entry.hip", "k_dense.hip", "k_tables.hip", "problem.cpp", "factor.cpp", "h2.cpp", "prof.cpp", "prog.cpp", "k_prog.hip", "k_vec.hip", "krylov.cpp", "capi.cpp", "serve.cpp", "noalloc.hip", "precomp.cpp", "merge.cpp", "adaptive.cpp"]


def _compile(src):
    obj = os.path.join(OBJ, src + ".o")
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC, *FLAGS, *lang, "-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 2)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    build_drivers()
    if verbose:
        print("built", LIB)
    return LIB


DRIVERS = ["testFMM2D_find_error", "testFMM2D_gen_rhs_x"]
ROOT = os.path.dirname(HERE)


def build_drivers():
    """Host C++ drivers over the source-compatible headers in include/aifmm,
    linked against the in-tree library (rpath $ORIGIN)."""
    for name in DRIVERS:
        out = os.path.join(OUT, name)
        cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
               os.path.join(HERE, "drivers", name + ".cpp"), "-o", out + ".tmp", "-L", OUT, "-laifmm_ls",
               "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver {name} failed to build:\n" + r.stderr[-4000:])
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build()
    sys.exit(0)
