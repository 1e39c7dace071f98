"""The C ABI library loads and exports every symbol include/aifmm_ls.h
declares (no compute: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "aifmm_ls.h")).read()
    return sorted(set(re.findall(r"\b(aifmm_ls_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("aifmm_ls_create", "aifmm_ls_factor", "aifmm_ls_solve", "aifmm_ls_entries", "aifmm_ls_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import aifmm_ls_amd as A
    if not os.path.exists(A.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(A.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    import aifmm_ls_amd as A
    monkeypatch.setattr(A, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(A, "_lib", None)
    with pytest.raises(A.AifmmError):
        A.lib()


def test_set_tuning_rejects_unknown_modes():
    import aifmm_ls_amd as A
    if not os.path.exists(A.LIB_PATH):
        pytest.skip("library not built")
    for bad in (7, 1024 + 8 * 4096, 512 + 3 * 4096, 300):
        with pytest.raises(A.AifmmError):
            A.set_tuning(bad)
    A.set_tuning(1)  # the default
