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


def test_request_structs_match_header(tmp_path):
    """The ctypes mirrors of mcg_request / mcg_result have the C layout (gcc offsetof)."""
    import shutil
    import subprocess
    from minicore_amd import _lib
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = {"mcg_request": _lib.Request, "mcg_result": _lib.Result}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{ROOT}/include/mcg.h"', "int main(void) {"]
    for cname, py in fields.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in fields.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
