"""The C-ABI library loads and exports every entry point include/mcg.h declares
(no GPU compute is run here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "mcg.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(mcg_\w+)\(", src, flags=re.M)))


def test_header_lists_entry_points():
    names = declared()
    assert "mcg_lloyd" in names and "mcg_assign_hard" in names and len(names) >= 18


def test_library_exports_every_declared_symbol():
    from minicore_amd import _lib
    L = ctypes.CDLL(_lib.LIBPATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == declared()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from minicore_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIBPATH", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError):
        _lib.load()


def test_error_path_without_device():
    """Invalid arguments come back as status codes with a message, not crashes."""
    from minicore_amd import _lib
    L = _lib.load()
    rc = L.mcg_ctx_create(12345, ctypes.byref(ctypes.c_void_p()))
    assert rc != 0
    assert L.mcg_last_error()
