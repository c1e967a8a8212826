"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import glob
import os
import re

import pytest

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


def test_stale_library_is_refused(monkeypatch):
    """The loader compares the source stamp built into the library with the
    sources beside it."""
    from libtopotoolbox_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "source_hash", lambda: "0" * 32)
    try:
        _lib.lib()
    except _lib.LibraryNotBuilt as e:
        assert "built from other sources" in str(e)
    else:
        raise AssertionError("a library built from other sources must be refused")
    monkeypatch.undo()
    _lib._lib = None
    assert _lib.lib().has_topotoolbox() == 1


def test_noalloc_rejects_oversized_grid_without_gpu():
    """The size check of the zero-allocation entry point runs before any
    device call: 600 x 600 tiles of 128 x 128 exceed its workspace."""
    from libtopotoolbox_amd import _lib

    L = _lib.lib()
    assert L.tt_fillsinks_noalloc_dev(None, None, None, 128 * 600, 128 * 600, None) == 2
    assert "at most 262128" in _lib.last_error()


def test_capped_entries_check_arguments_without_gpu():
    """max_visits outside [1, 2^32 - 1] and a NULL converged pointer are
    rejected before any device call."""
    import ctypes

    from libtopotoolbox_amd import _lib

    L = _lib.lib()
    conv = ctypes.c_int(7)
    for fn in ("tt_fillsinks_capped_dev", "tt_reconstruct_capped_dev"):
        f = getattr(L, fn)
        extra = (None,) if fn == "tt_fillsinks_capped_dev" else ()
        for cap in (0, -1, 2**32):
            assert f(None, None, *extra, 64, 64, cap, None, ctypes.byref(conv), None) == 1
            assert "max_visits" in _lib.last_error()
        assert f(None, None, *extra, 64, 64, 10, None, None, None) == 1
        assert "converged" in _lib.last_error()
        assert conv.value == 7


def test_fill_schedule_setter_without_gpu():
    """The schedule setter is host-only: known values are accepted, an
    unknown one is refused with a message, and the Python wrapper checks
    names before calling it."""
    import libtopotoolbox_amd as tt
    from libtopotoolbox_amd import _lib

    L = _lib.lib()
    for v in (1, 2, 0):
        assert L.tt_set_fill_schedule(v) == 0
    assert L.tt_set_fill_schedule(7) == 1
    assert "unknown schedule 7" in _lib.last_error()
    with pytest.raises(ValueError):
        tt.set_fill_schedule("lifo")
    tt.set_fill_schedule("auto")
