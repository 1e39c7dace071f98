"""Precomputation tables and their text cache.

The reference computes the far table M at levels 0..10 and 53 neighbour
operators per level (9 colleague, 12 fine, 20 separated fine, 12 coarse) at
levels 2..L, writes them as Eigen text under precomputations/, and reads them
back (LS_Assembly.hpp:262-275, 1166-1228, 1290-1459).  Checked here:
  * every family's device tables against the oracle's (whose families are
    pinned by direct integrals and mirror symmetry in test_oracle_tables.py);
  * the files: names, count, Eigen's layout, values = the tables at 6
    significant digits, read back by ReadFromTextFile's rule;
  * a problem built from the cache equals a freshly computed (rounded) one:
    tables bitwise, solution to rounding;
  * refusals: missing file, ragged file, a cache keyed for another kappa with
    the same int(kappa); a cache without the key file (the reference's) is read.
"""
import os
import re
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FAMILY_FILE = ["colleagueNeighborInteraction", "fineNeighborInteraction", "separatedFineNeighborInteraction",
               "coarseNeighborInteraction"]
FAMILY_SIZE = [9, 12, 20, 12]


def read_eigen(path):
    """ReadFromTextFile (LS_Assembly.hpp:1186-1215): rows = newlines, columns =
    words / rows, complex words "(re,im)"."""
    text = open(path).read()
    rows = text.count("\n")
    words = text.split()
    assert rows > 0 and len(words) % rows == 0, path
    cols = len(words) // rows

    def val(w):
        m = re.fullmatch(r"\(([^,()]+),([^,()]+)\)", w)
        return complex(float(m.group(1)), float(m.group(2))) if m else complex(float(w), 0.0)

    return np.array([val(w) for w in words]).reshape(rows, cols)


def round6(z):
    f = np.vectorize(lambda v: float("%.6g" % v))
    return f(z.real) + 1j * f(z.imag)


def test_reader_rule():
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "m")
        with open(p, "w") as f:
            f.write("   (1,-2) (0.5,3e-05)\n(-1.25,0)      (7,8)\n")
        np.testing.assert_array_equal(read_eigen(p), np.array([[1 - 2j, 0.5 + 3e-5j], [-1.25, 7 + 8j]]))


@pytest.fixture(scope="module")
def pair():
    import aifmm_ls_amd as A
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    print("\n[precomp] L=3 kappa=10", flush=True)
    g = A.LSProblem(3, 10.0, 8, 0, quantize=False)
    o = O.Problem(3, 10.0, 8, 0, quantize=False)
    yield g, o
    g.close()


@pytest.mark.gpu
def test_family_tables_match_oracle(pair):
    g, o = pair
    worst = 0.0
    for level in (2, 3):
        for f in range(4):
            for n in range(FAMILY_SIZE[f]):
                a, b = g.table(f, level, n), o.table(f, level, n)
                worst = max(worst, np.linalg.norm(a - b) / np.linalg.norm(b))
    for level in (0, 1, 3, 7, 10):
        a, b = g.table(-1, level), o.table(-1, level)
        worst = max(worst, np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"  worst table distance device vs oracle {worst:.2e}", flush=True)
    assert worst < 1e-12


@pytest.mark.gpu
def test_leaf_tables_are_the_level_L_families(pair):
    g, _ = pair
    M, T, _ = g.tables()
    np.testing.assert_array_equal(g.table(-1, 3), M)
    for n in range(9):
        np.testing.assert_array_equal(g.table(0, 3, n), T[n])


@pytest.fixture(scope="module")
def written(pair, tmp_path_factory):
    g, _ = pair
    d = tmp_path_factory.mktemp("cache")
    t = g.write_precomputations(str(d))
    print(f"  wrote the precomputation in {t * 1e3:.1f} ms of device quadrature", flush=True)
    return d


@pytest.mark.gpu
def test_written_files(pair, written):
    g, _ = pair
    md = written / "precomputations" / "M" / "M_6_10_6"
    nd = written / "precomputations" / "Neighbor" / "Neighbor_6_10_6"
    assert sorted(os.listdir(md)) == sorted([f"M_6_{l}" for l in range(11)] + ["aifmm_key"])
    want = [f"{FAMILY_FILE[f]}_{j}_{n}" for j in (2, 3) for f in range(4) for n in range(FAMILY_SIZE[f])]
    assert sorted(os.listdir(nd)) == sorted(want + ["aifmm_key"])
    for l in (0, 3, 10):
        np.testing.assert_array_equal(read_eigen(md / f"M_6_{l}"), round6(g.table(-1, l)))
    for j in (2, 3):
        for f in range(4):
            for n in (0, FAMILY_SIZE[f] - 1):
                p = nd / f"{FAMILY_FILE[f]}_{j}_{n}"
                np.testing.assert_array_equal(read_eigen(p), round6(g.table(f, j, n)))
                # Eigen's layout: coefficients right-aligned to one width, one
                # space apart, one line per row, a newline after the last row
                text = open(p).read()
                lines = text.split("\n")
                assert lines[-1] == "" and len(lines) == 37
                assert len({len(x) for x in lines[:-1]}) == 1
                w = max(len(t) for t in text.split())
                assert len(lines[0]) == 36 * w + 35


@pytest.mark.gpu
def test_cached_problem_equals_fresh(written):
    import aifmm_ls_amd as A
    fresh = A.LSProblem(3, 10.0, 8, 0, quantize=True)
    cached = A.LSProblem(3, 10.0, 8, 0, precomputations=str(written))
    Mf, Tf, _ = fresh.tables()
    Mc, Tc, _ = cached.tables()
    np.testing.assert_array_equal(Mc, Mf)
    np.testing.assert_array_equal(Tc, Tf)
    b = fresh.rhs(0.0)
    for p in (fresh, cached):
        p.assemble()
        p.factor(A.PREC_C128)
    xf, _ = fresh.solve(b)
    xc, _ = cached.solve(b)
    # same tables, same operator: equal up to the order of the solve's own
    # reductions (launch tuning may pick another split per problem)
    assert np.linalg.norm(xc - xf) <= 1e-13 * np.linalg.norm(xf)
    # a 2-level problem reads the same cache (levels 0..2 of M, level 2 of the families)
    two = A.LSProblem(2, 10.0, 8, 0, precomputations=str(written))
    ref2 = A.LSProblem(2, 10.0, 8, 0, quantize=True)
    np.testing.assert_array_equal(two.tables()[1], ref2.tables()[1])
    for p in (fresh, cached, two, ref2):
        p.close()


@pytest.mark.gpu
def test_cache_refusals(written, tmp_path):
    import shutil
    import aifmm_ls_amd as A
    d = tmp_path / "c"
    shutil.copytree(written, d)
    nd = d / "precomputations" / "Neighbor" / "Neighbor_6_10_6"
    md = d / "precomputations" / "M" / "M_6_10_6"
    # kappa 10.5 has the reference's key int(kappa) = 10, but not this cache's key file
    with pytest.raises(A.AifmmError, match="another problem"):
        A.LSProblem(3, 10.5, 8, 0, precomputations=str(d))
    # deeper tree than the cache holds
    with pytest.raises(A.AifmmError, match="could not open"):
        A.LSProblem(4, 10.0, 8, 0, precomputations=str(d))
    # a ragged file
    p = nd / "fineNeighborInteraction_3_5"
    text = open(p).read()
    with open(p, "w") as f:
        f.write(text + "(1,2)\n")
    with pytest.raises(A.AifmmError, match="cannot form a matrix|expected 36 x 36"):
        A.LSProblem(3, 10.0, 8, 0, precomputations=str(d))
    with open(p, "w") as f:
        f.write(text)
    # without the key files (as the reference writes it), and with other
    # spacing and digits in one file: read by the same rule
    os.remove(nd / "aifmm_key")
    os.remove(md / "aifmm_key")
    q = nd / "colleagueNeighborInteraction_3_8"
    T = read_eigen(q)
    T2 = T * (1 + 1e-9)
    with open(q, "w") as f:
        for row in T2:
            f.write("\t".join(f"({z.real:.17g},{z.imag:.17g})" for z in row) + "\n")
    g = A.LSProblem(3, 10.5, 8, 0, precomputations=str(d))
    np.testing.assert_array_equal(g.tables()[1][8], T2)
    g.close()
