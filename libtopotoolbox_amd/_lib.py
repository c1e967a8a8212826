"""ctypes binding of libtopotoolbox_amd.so (the C ABI in include/topotoolbox_amd.h).

The shared library is built in-tree by ``libtopotoolbox_amd.build``. There is
no fallback: when the library is missing or cannot be loaded, every entry
point raises ``LibraryNotBuilt``.
"""
from __future__ import annotations

import ctypes
import os

from .build import LIB_PATH, source_hash


class LibraryNotBuilt(RuntimeError):
    pass


class TTError(RuntimeError):
    """A nonzero status from the C ABI."""


class Stats(ctypes.Structure):
    """Mirror of tt_stats."""

    _fields_ = [
        ("requeues", ctypes.c_int64),
        ("visits", ctypes.c_int64),
        ("local_iters", ctypes.c_int64),
        ("max_local", ctypes.c_int64),
        ("grid", ctypes.c_int64),
        ("tile_x", ctypes.c_int64),
        ("tile_y", ctypes.c_int64),
        ("error", ctypes.c_uint32),
        ("t_stage", ctypes.c_int64),
        ("t_relax", ctypes.c_int64),
        ("t_write", ctypes.c_int64),
        ("t_wait", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchItem(ctypes.Structure):
    """Mirror of tt_batch_item."""

    _fields_ = [
        ("nx", ctypes.c_ssize_t),
        ("ny", ctypes.c_ssize_t),
        ("offset", ctypes.c_ssize_t),
        ("cellsize", ctypes.c_float),
    ]


class BatchBuffers(ctypes.Structure):
    """Mirror of tt_batch_buffers (device pointers)."""

    _fields_ = [
        ("output", ctypes.c_void_p),
        ("spill", ctypes.c_void_p),
        ("dem", ctypes.c_void_p),
        ("bc", ctypes.c_void_p),
        ("slopes", ctypes.c_void_p),
        ("g_out", ctypes.c_void_p),
        ("g_dem", ctypes.c_void_p),
        ("g_slopes", ctypes.c_void_p),
    ]


BATCH_FILLSINKS = 0
BATCH_EXCESSTOPOGRAPHY = 1
BATCH_FORWARD_ONLY = 1

_vp = ctypes.c_void_p
_pd = ctypes.c_ssize_t  # ptrdiff_t
_dims_t = _pd * 2
_se_dims_t = _pd * 3

# name -> (restype, argtypes)
SIGNATURES = {
    "has_topotoolbox": (ctypes.c_int, []),
    "fillsinks": (None, [_vp, _vp, _vp, _dims_t]),
    "reconstruct": (None, [_vp, _vp, _dims_t]),
    "excesstopography_fsm2d": (None, [_vp, _vp, _vp, ctypes.c_float, _dims_t]),
    "fillsinks_hybrid": (None, [_vp, _vp, _vp, _vp, _dims_t]),
    "reconstruct_hybrid": (ctypes.c_int, [_vp, _vp, _vp, _dims_t]),
    "excesstopography_fmm2d": (None, [_vp, _vp, _vp, _vp, _vp, ctypes.c_float, _dims_t]),
    "excesstopography_fmm3d": (None, [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_float, _dims_t,
                                      _pd]),
    "tt_excesstopography_fmm3d_dev": (
        ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_float, _pd, _pd, _pd, _vp,
                       ctypes.POINTER(Stats)]),
    "identifyflats": (ctypes.c_ssize_t, [_vp, _vp, _dims_t]),
    "min_filter": (None, [_vp, _vp, _vp, _dims_t, _se_dims_t]),
    "max_filter": (None, [_vp, _vp, _vp, _dims_t, _se_dims_t]),
    "min_filter_square": (None, [_vp, _vp, _vp, ctypes.c_uint8, _dims_t]),
    "max_filter_square": (None, [_vp, _vp, _vp, ctypes.c_uint8, _dims_t]),
    "tt_fillsinks_dev": (ctypes.c_int, [_vp, _vp, _vp, _pd, _pd, _vp, ctypes.POINTER(Stats)]),
    "tt_fillsinks_noalloc_dev": (ctypes.c_int, [_vp, _vp, _vp, _pd, _pd, _vp]),
    "tt_reconstruct_dev": (ctypes.c_int, [_vp, _vp, _pd, _pd, _vp, ctypes.POINTER(Stats)]),
    "tt_fillsinks_capped_dev": (
        ctypes.c_int, [_vp, _vp, _vp, _pd, _pd, ctypes.c_int64, _vp,
                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(Stats)]),
    "tt_reconstruct_capped_dev": (
        ctypes.c_int, [_vp, _vp, _pd, _pd, ctypes.c_int64, _vp, ctypes.POINTER(ctypes.c_int),
                       ctypes.POINTER(Stats)]),
    "tt_excesstopography_fsm2d_dev": (
        ctypes.c_int, [_vp, _vp, _vp, ctypes.c_float, _pd, _pd, _vp, ctypes.POINTER(Stats)]),
    "tt_fillsinks_spill_dev": (
        ctypes.c_int, [_vp, _vp, _vp, _vp, _pd, _pd, _vp, ctypes.POINTER(Stats)]),
    "tt_fillsinks_backward_dev": (ctypes.c_int, [_vp, _vp, _vp, _pd, _pd, _vp]),
    "tt_excesstopography_fsm2d_backward_dev": (
        ctypes.c_int,
        [_vp, _vp, _vp, _vp, _vp, ctypes.c_float, _vp, _pd, _pd, _vp, ctypes.POINTER(Stats)]),
    "gwdt_computecosts": (None, [_vp, _vp, _vp, _vp, _vp, _dims_t]),
    "tt_gwdt_computecosts_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _pd, _pd, _vp]),
    "tt_identifyflats_dev": (
        ctypes.c_int, [_vp, _vp, _pd, _pd, _vp, ctypes.POINTER(ctypes.c_int64)]),
    "tt_value_filter_dev": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, _se_dims_t, _pd, _pd,
                                           _vp]),
    "tt_value_filter_square_dev": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, ctypes.c_int,
                                                  _pd, _pd, _vp]),
    "tt_batch_plan": (ctypes.c_int, [ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_size_t),
                                     ctypes.POINTER(ctypes.c_size_t)]),
    "tt_batch_plan_set_sweep_caps": (ctypes.c_int, [_vp, ctypes.c_int,
                                                     ctypes.POINTER(ctypes.c_uint32)]),
    "tt_batch_run_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_size_t, ctypes.c_uint,
                                        ctypes.c_int, _vp, ctypes.POINTER(Stats)]),
    "tt_set_tile_shape": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "tt_set_fill_schedule": (ctypes.c_int, [ctypes.c_int]),
    "tt_device_error": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]),
    "tt_set_wait_limit": (ctypes.c_int, [ctypes.c_uint64]),
    "tt_host_workspace_release": (None, []),
    "tt_build_hash": (ctypes.c_char_p, []),
    "tt_last_error": (ctypes.c_char_p, []),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load the HIP library (raises LibraryNotBuilt if it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(
                f"{LIB_PATH} is missing: build it with `python -m libtopotoolbox_amd.build`")
        try:
            L = ctypes.CDLL(LIB_PATH)
        except OSError as e:
            raise LibraryNotBuilt(f"cannot load {LIB_PATH}: {e}") from e
        stamp = getattr(L, "tt_build_hash", None)
        if stamp is None:
            raise LibraryNotBuilt(f"{LIB_PATH} carries no source stamp: rebuild it")
        stamp.restype = ctypes.c_char_p
        built, now = stamp().decode(), source_hash()
        if built != now:
            raise LibraryNotBuilt(
                f"{LIB_PATH} was built from other sources (stamp {built}, sources {now}): "
                "rebuild it with `python -m libtopotoolbox_amd.build`")
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list:
    """Names of the SIGNATURES entries the loaded library exports."""
    L = lib()
    return [n for n in SIGNATURES if getattr(L, n, None) is not None]


def last_error() -> str:
    msg = lib().tt_last_error()
    return msg.decode() if msg else ""


def check(status: int, what: str) -> None:
    if status != 0:
        raise TTError(f"{what} failed (status {status}): {last_error()}")


def device_error(clear: bool = True) -> int:
    """Error bits that persistent launches on the current GPU recorded since
    the last clear (synchronises the device). 1: a bounded queue wait
    expired, so some result is incomplete."""
    bits = ctypes.c_uint32(0)
    check(lib().tt_device_error(1 if clear else 0, ctypes.byref(bits)), "tt_device_error")
    return int(bits.value)


def check_device_error(what: str) -> None:
    """Raises TTError if a launch on the current GPU recorded an error."""
    bits = device_error(clear=True)
    if bits:
        raise TTError(f"{what}: device error bits 0x{bits:x} (1: a queue wait exceeded its "
                      "time limit; results are incomplete)")


def dims(nx: int, ny: int):
    return _dims_t(nx, ny)


def se_dims(sx: int, sy: int, slices: int):
    return _se_dims_t(sx, sy, slices)
