"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import glob
import os
import re

from conftest import ROOT


def declared_functions():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(([^;{}]*)\)\s*;", text):
            if m.group(1) not in ("if", "while", "for", "return", "sizeof"):
                names.add(m.group(1))
    return names


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for ref in ("has_topotoolbox", "fillsinks", "reconstruct", "excesstopography_fsm2d"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    from libtopotoolbox_amd import _lib

    L = _lib.lib()
    missing = [n for n in sorted(declared_functions()) if getattr(L, n, None) is None]
    assert not missing, f"declared but not exported: {missing}"


def test_python_signatures_cover_header():
    from libtopotoolbox_amd import _lib

    missing = sorted(declared_functions() - set(_lib.SIGNATURES))
    assert not missing, f"no ctypes signature for: {missing}"


def test_has_topotoolbox_without_gpu():
    from libtopotoolbox_amd import _lib

    assert _lib.lib().has_topotoolbox() == 1


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from libtopotoolbox_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    try:
        _lib.lib()
    except _lib.LibraryNotBuilt:
        pass
    else:
        raise AssertionError("loading an absent library must raise")
