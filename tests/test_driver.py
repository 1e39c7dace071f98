"""The find_error driver built on the source-compatible FMM2DTree / userkernel
headers (include/aifmm), and the rhs_x/*.bin vector format it shares with the
reference's generator (cereal PortableBinary of a std::vector<double>:
endianness byte, uint64 count, doubles)."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "aifmm_ls_amd", "_lib", "testFMM2D_find_error")
GEN = os.path.join(ROOT, "aifmm_ls_amd", "_lib", "testFMM2D_gen_rhs_x")


def write_vec(path, v):
    v = np.ascontiguousarray(v, np.float64)
    with open(path, "wb") as f:
        f.write(struct.pack("<B", 1))
        f.write(struct.pack("<Q", v.size))
        f.write(v.astype("<f8").tobytes())


def read_vec(path):
    with open(path, "rb") as f:
        little = struct.unpack("<B", f.read(1))[0]
        e = "<" if little else ">"
        n = struct.unpack(e + "Q", f.read(8))[0]
        return np.frombuffer(f.read(8 * n), dtype=e + "f8").astype(np.float64)


def test_vector_file_layout(tmp_path):
    v = np.array([1.5, -2.0, 3.25e-300, np.pi])
    p = tmp_path / "v.bin"
    write_vec(p, v)
    raw = p.read_bytes()
    assert raw[0] == 1 and struct.unpack("<Q", raw[1:9])[0] == 4 and len(raw) == 9 + 32
    np.testing.assert_array_equal(read_vec(p), v)
    # big-endian writer, as a PortableBinary archive from such a host would be
    q = tmp_path / "b.bin"
    with open(q, "wb") as f:
        f.write(struct.pack("B", 0) + struct.pack(">Q", 4) + v.astype(">f8").tobytes())
    np.testing.assert_array_equal(read_vec(q), v)


def test_driver_built_and_usage():
    for exe in (DRIVER, GEN):
        if not os.path.exists(exe):
            pytest.skip("drivers not built")
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_find_error_protocol(tmp_path):
    """The reference's two-run protocol at L=3 (kappa 10, Qchoice 0, TOL_POW 8):
    x_true = the dense FP64 solution of the oracle's matrix, b = A x_true written
    to rhs_x/ in the archive layout; the driver reads b, solves, and reports the
    forward error against x_true read back from rhs_x/."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    o = O.Problem(3, 10.0, 8, 0)
    A = o.dense()
    x_true = np.linalg.solve(A, o.rhs(0.0))
    b = A @ x_true
    N = o.N
    d = tmp_path / "rhs_x"
    d.mkdir()
    tag = f"0_10_{N}.bin"
    write_vec(d / f"rhs_real_{tag}", b.real)
    write_vec(d / f"rhs_imag_{tag}", b.imag)
    write_vec(d / f"x_real_{tag}", x_true.real)
    write_vec(d / f"x_imag_{tag}", x_true.imag)
    print(f"\n[driver] L=3 N={N}", flush=True)
    r = subprocess.run([DRIVER, "6", "10", "6", "1", "0", "8", "3"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "rhs from rhs_x/rhs_real_" in r.stdout
    err = float(r.stdout.strip().splitlines()[-1].split(":")[1])
    xe, _ = o.solve(b, O.SCHED_EFFICIENT)
    ref_err = np.linalg.norm(xe - x_true) / np.linalg.norm(x_true)
    print(f"  driver forward error {err:.3e}, reference schedule {ref_err:.3e}", flush=True)
    assert err < 1.5 * ref_err + 1e-9
    # with complex64 factors and refinement
    r = subprocess.run([DRIVER, "6", "10", "6", "1", "0", "8", "3", "1", "20", "1e-13"], capture_output=True,
                       text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    err2 = float(r.stdout.strip().splitlines()[-1].split(":")[1])
    assert err2 < 1.5 * ref_err + 1e-9


@pytest.mark.gpu
def test_generator_then_find_error(tmp_path):
    """The reference's two-run protocol with this build's own binaries only
    (Readme.md: run the generator, then find_error with the same arguments):
    testFMM2D_gen_rhs_x solves for the incident field, writes x and the dense
    rhs = A x to rhs_x/ and prints |A x - b| / |b|; testFMM2D_find_error then
    solves A y = rhs and reports |y - x| / |x|.  Checked: the files against
    the oracle's dense matrix (rhs = A x to 1e-12), the printed residual
    against numpy's, and the forward error against the reference schedule's
    own on the same right-hand side."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    args = ["6", "10", "6", "1", "0", "8", "3"]
    print("\n[driver] generator L=3", flush=True)
    r = subprocess.run([GEN, *args], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "gen rhs x" in r.stdout
    gen_err = float(r.stdout.strip().splitlines()[-1].split(":")[1])
    o = O.Problem(3, 10.0, 8, 0)
    N = o.N
    tag = f"0_10_{N}.bin"
    d = tmp_path / "rhs_x"
    x = read_vec(d / f"x_real_{tag}") + 1j * read_vec(d / f"x_imag_{tag}")
    rhs = read_vec(d / f"rhs_real_{tag}") + 1j * read_vec(d / f"rhs_imag_{tag}")
    Ad = o.dense()
    b0 = o.rhs(0.0)
    assert np.linalg.norm(rhs - Ad @ x) / np.linalg.norm(Ad @ x) < 1e-12
    res = np.linalg.norm(Ad @ x - b0) / np.linalg.norm(b0)
    assert abs(gen_err - res) <= 1e-5 * res + 1e-15  # printed to 6 significant digits, as the reference does
    print(f"[driver] find_error on the generator's files (residual of x: {res:.3e})", flush=True)
    r = subprocess.run([DRIVER, *args], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "rhs from rhs_x/rhs_real_" in r.stdout
    fwd = float(r.stdout.strip().splitlines()[-1].split(":")[1])
    xe, _ = o.solve(rhs, O.SCHED_EFFICIENT)
    ref = np.linalg.norm(xe - x) / np.linalg.norm(x)
    print(f"  forward error {fwd:.3e}, reference schedule {ref:.3e}", flush=True)
    assert fwd < 1.5 * ref + 1e-9


@pytest.mark.gpu
def test_driver_writes_then_reads_precomputations(tmp_path):
    """The reference's ctor writes precomputations/ under the working
    directory and its Readme reuses them by commenting the writes out
    (LS_Assembly.hpp:262-275, LS/Readme.md:1-5); the drivers do the same under
    AIFMM_PRECOMPUTATIONS=write / read.  Both runs print the same error."""
    args = ["6", "10", "6", "1", "0", "8", "3"]
    print("\n[driver] generator, then find_error writing and reading precomputations/", flush=True)
    r = subprocess.run([GEN, *args], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stderr
    out = []
    for mode in ("write", "read"):
        env = dict(os.environ, AIFMM_PRECOMPUTATIONS=mode)
        r = subprocess.run([DRIVER, *args], capture_output=True, text=True, cwd=tmp_path, timeout=120, env=env)
        assert r.returncode == 0, (mode, r.stderr)
        out.append(r.stdout.strip().splitlines()[-1])
        md = tmp_path / "precomputations" / "M" / "M_6_10_6"
        nd = tmp_path / "precomputations" / "Neighbor" / "Neighbor_6_10_6"
        assert len(os.listdir(md)) == 12 and len(os.listdir(nd)) == 2 * 53 + 1
    print(f"  {out[0]} | {out[1]}", flush=True)
    assert out[0] == out[1]
    # a deeper tree than the files hold: the read fails as the reference's does
    env = dict(os.environ, AIFMM_PRECOMPUTATIONS="read")
    r = subprocess.run([DRIVER, "6", "10", "6", "1", "0", "8", "4"], capture_output=True, text=True, cwd=tmp_path,
                       timeout=120, env=env)
    assert r.returncode != 0 and "could not open" in (r.stderr + r.stdout)
