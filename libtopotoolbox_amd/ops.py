"""Host-side mirror of the reference interface for the hot path.

Two flavours of each reference entry point:

* ``fillsinks`` / ``reconstruct`` / ``excesstopography_fsm2d`` take torch
  tensors already on the GPU (2-D, float32, C-contiguous, shape (ny, nx) so
  that dims = (nx, ny) with dims[0] fastest, as in the reference) and run on
  the current torch stream through the device-pointer C ABI.
* ``*_host`` take numpy arrays and call the drop-in C symbols with the
  reference's exact signatures (host pointers, copies in and out).

Argument meaning follows include/topotoolbox.h of the reference:
fillsinks (:256-293), reconstruct (:1045-1056), excesstopography_fsm2d
(:536-609).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib


def _stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def _check_grid(t: torch.Tensor, name: str, dtype=torch.float32) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on the GPU")
    if t.dim() != 2:
        raise ValueError(f"{name} must be 2-D (ny, nx)")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be C-contiguous")


def _same_shape(ref: torch.Tensor, *others) -> None:
    for name, t in others:
        if t.shape != ref.shape:
            raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(ref.shape)}")
        if t.device != ref.device:
            raise ValueError(f"{name} is on {t.device}, expected {ref.device}")


def _stats_ptr(want: bool):
    return _lib.Stats() if want else None


def _ret(out, st):
    return (out, st.as_dict()) if st is not None else out


def set_tile_shape(qx: int, qy: int) -> None:
    """Tile shape (in 64x64 quadrants) for later calls on this thread."""
    _lib.check(_lib.lib().tt_set_tile_shape(qx, qy), "tt_set_tile_shape")


FILL_SCHEDULES = {"auto": 0, "fifo": 1, "priority": 2}


def set_fill_schedule(schedule: str) -> None:
    """Work-queue schedule of single-grid fillsinks and reconstruct for
    later calls on this thread: "auto" (default), "fifo" or "priority"
    (same results; "priority" speeds up reconstructions from sparse
    markers, which "auto" leaves on the FIFO queue)."""
    if schedule not in FILL_SCHEDULES:
        raise ValueError(f"unknown fill schedule {schedule!r}; one of {sorted(FILL_SCHEDULES)}")
    _lib.check(_lib.lib().tt_set_fill_schedule(FILL_SCHEDULES[schedule]), "tt_set_fill_schedule")


def fillsinks(dem: torch.Tensor, bc: torch.Tensor, out: torch.Tensor | None = None,
              stats: bool = False):
    """Filled DEM. ``bc == 1`` marks fixed cells (reference fillsinks)."""
    _check_grid(dem, "dem")
    _check_grid(bc, "bc", torch.uint8)
    _same_shape(dem, ("bc", bc))
    if out is None:
        out = torch.empty_like(dem)
    else:
        _check_grid(out, "out")
        _same_shape(dem, ("out", out))
    ny, nx = dem.shape
    st = _stats_ptr(stats)
    with torch.cuda.device(dem.device):
        rc = _lib.lib().tt_fillsinks_dev(out.data_ptr(), dem.data_ptr(), bc.data_ptr(), nx, ny,
                                         _stream_handle(), st)
    _lib.check(rc, "fillsinks")
    return _ret(out, st)


def reconstruct(marker: torch.Tensor, mask: torch.Tensor, inplace: bool = False,
                stats: bool = False):
    """Grayscale reconstruction by dilation of ``mask`` from ``marker``."""
    _check_grid(marker, "marker")
    _check_grid(mask, "mask")
    _same_shape(marker, ("mask", mask))
    m = marker if inplace else marker.clone()
    ny, nx = m.shape
    st = _stats_ptr(stats)
    with torch.cuda.device(m.device):
        rc = _lib.lib().tt_reconstruct_dev(m.data_ptr(), mask.data_ptr(), nx, ny,
                                           _stream_handle(), st)
    _lib.check(rc, "reconstruct")
    return _ret(m, st)


def fillsinks_capped(dem: torch.Tensor, bc: torch.Tensor, max_visits: int,
                     stats: bool = False):
    """fillsinks that stops after about ``max_visits`` tile visits (128 x 128
    tiles, each visited at least once), like the reference's iteration cap
    in reconstruct. Returns (output, converged): an unconverged output is a
    partial fill (full fill <= output, output >= dem) from which
    ``reconstruct(-output, -dem)`` resumes to the full result. Synchronises."""
    _check_grid(dem, "dem")
    _check_grid(bc, "bc", torch.uint8)
    _same_shape(dem, ("bc", bc))
    out = torch.empty_like(dem)
    ny, nx = dem.shape
    conv = ctypes.c_int(0)
    st = _stats_ptr(stats)
    with torch.cuda.device(dem.device):
        rc = _lib.lib().tt_fillsinks_capped_dev(out.data_ptr(), dem.data_ptr(), bc.data_ptr(),
                                                nx, ny, int(max_visits), _stream_handle(),
                                                ctypes.byref(conv), st)
    _lib.check(rc, "fillsinks_capped")
    return _ret((out, bool(conv.value)), st)


def reconstruct_capped(marker: torch.Tensor, mask: torch.Tensor, max_visits: int,
                       stats: bool = False):
    """reconstruct that stops after about ``max_visits`` tile visits (the
    reference stops at 1000 iterations). Returns (marker, converged); an
    unconverged marker resumes with ``reconstruct``. Synchronises."""
    _check_grid(marker, "marker")
    _check_grid(mask, "mask")
    _same_shape(marker, ("mask", mask))
    m = marker.clone()
    ny, nx = m.shape
    conv = ctypes.c_int(0)
    st = _stats_ptr(stats)
    with torch.cuda.device(m.device):
        rc = _lib.lib().tt_reconstruct_capped_dev(m.data_ptr(), mask.data_ptr(), nx, ny,
                                                  int(max_visits), _stream_handle(),
                                                  ctypes.byref(conv), st)
    _lib.check(rc, "reconstruct_capped")
    return _ret((m, bool(conv.value)), st)


def excesstopography_fsm2d(dem: torch.Tensor, threshold_slopes: torch.Tensor, cellsize: float,
                           out: torch.Tensor | None = None, stats: bool = False):
    """Solution of the constrained eikonal equation (subtract from dem for
    the excess topography), reference excesstopography_fsm2d."""
    _check_grid(dem, "dem")
    _check_grid(threshold_slopes, "threshold_slopes")
    _same_shape(dem, ("threshold_slopes", threshold_slopes))
    if out is None:
        out = torch.empty_like(dem)
    else:
        _check_grid(out, "out")
        _same_shape(dem, ("out", out))
    ny, nx = dem.shape
    st = _stats_ptr(stats)
    with torch.cuda.device(dem.device):
        rc = _lib.lib().tt_excesstopography_fsm2d_dev(out.data_ptr(), dem.data_ptr(),
                                                      threshold_slopes.data_ptr(),
                                                      float(cellsize), nx, ny,
                                                      _stream_handle(), st)
    _lib.check(rc, "excesstopography_fsm2d")
    return _ret(out, st)


def fillsinks_spill(dem: torch.Tensor, bc: torch.Tensor, stats: bool = False):
    """Filled DEM and, per cell, the int32 index (p = j * nx + i) of the cell
    whose DEM value the output copies; -1 on NaN and unreached cells."""
    _check_grid(dem, "dem")
    _check_grid(bc, "bc", torch.uint8)
    _same_shape(dem, ("bc", bc))
    out = torch.empty_like(dem)
    spill = torch.empty(dem.shape, dtype=torch.int32, device=dem.device)
    ny, nx = dem.shape
    st = _stats_ptr(stats)
    with torch.cuda.device(dem.device):
        rc = _lib.lib().tt_fillsinks_spill_dev(out.data_ptr(), spill.data_ptr(), dem.data_ptr(),
                                               bc.data_ptr(), nx, ny, _stream_handle(), st)
    _lib.check(rc, "fillsinks_spill")
    return _ret((out, spill), st)


def fillsinks_backward(spill: torch.Tensor, g_out: torch.Tensor) -> torch.Tensor:
    """g_dem[q] = sum of g_out[p] over p with spill[p] == q (deterministic)."""
    _check_grid(spill, "spill", torch.int32)
    _check_grid(g_out, "g_out")
    _same_shape(spill, ("g_out", g_out))
    g_dem = torch.empty_like(g_out)
    ny, nx = g_out.shape
    with torch.cuda.device(g_out.device):
        rc = _lib.lib().tt_fillsinks_backward_dev(g_dem.data_ptr(), spill.data_ptr(),
                                                  g_out.data_ptr(), nx, ny, _stream_handle())
    _lib.check(rc, "fillsinks_backward")
    return g_dem


def excesstopography_fsm2d_backward(excess: torch.Tensor, dem: torch.Tensor,
                                    threshold_slopes: torch.Tensor, cellsize: float,
                                    g_out: torch.Tensor, stats: bool = False):
    """(dL/ddem, dL/dslopes) for L = sum(g_out * excess) at the converged
    output `excess` of excesstopography_fsm2d."""
    for name, t in (("excess", excess), ("dem", dem), ("threshold_slopes", threshold_slopes),
                    ("g_out", g_out)):
        _check_grid(t, name)
    _same_shape(excess, ("dem", dem), ("threshold_slopes", threshold_slopes), ("g_out", g_out))
    g_dem = torch.empty_like(dem)
    g_s = torch.empty_like(dem)
    ny, nx = dem.shape
    st = _stats_ptr(stats)
    with torch.cuda.device(dem.device):
        rc = _lib.lib().tt_excesstopography_fsm2d_backward_dev(
            g_dem.data_ptr(), g_s.data_ptr(), excess.data_ptr(), dem.data_ptr(),
            threshold_slopes.data_ptr(), float(cellsize), g_out.data_ptr(), nx, ny,
            _stream_handle(), st)
    _lib.check(rc, "excesstopography_fsm2d_backward")
    return _ret((g_dem, g_s), st)


# ------------------------------------------------------------------ stencils
def identifyflats(dem: torch.Tensor, count: bool = True):
    """Flat / sill / presill classification (reference identifyflats,
    include/topotoolbox.h:363-405): int32 flags with bit 1 flat, 2 sill,
    4 presill. Returns (flags, number of flats), or only flags with
    ``count=False`` (then the call does not synchronise)."""
    _check_grid(dem, "dem")
    out = torch.empty(dem.shape, dtype=torch.int32, device=dem.device)
    ny, nx = dem.shape
    n = ctypes.c_int64(0)
    with torch.cuda.device(dem.device):
        rc = _lib.lib().tt_identifyflats_dev(out.data_ptr(), dem.data_ptr(), nx, ny,
                                             _stream_handle(),
                                             ctypes.byref(n) if count else None)
    _lib.check(rc, "identifyflats")
    return (out, int(n.value)) if count else out


def gwdt_computecosts(flats: torch.Tensor, original_dem: torch.Tensor,
                      filled_dem: torch.Tensor):
    """Costs of the gray-weighted distance transform and the flats'
    connected-component labels (reference gwdt_computecosts,
    include/topotoolbox.h:407-466): (costs float32, conncomps int64).
    Synchronises once."""
    _check_grid(flats, "flats", torch.int32)
    _check_grid(original_dem, "original_dem")
    _check_grid(filled_dem, "filled_dem")
    _same_shape(flats, ("original_dem", original_dem), ("filled_dem", filled_dem))
    costs = torch.empty(flats.shape, dtype=torch.float32, device=flats.device)
    cc = torch.empty(flats.shape, dtype=torch.int64, device=flats.device)
    ny, nx = flats.shape
    with torch.cuda.device(flats.device):
        rc = _lib.lib().tt_gwdt_computecosts_dev(costs.data_ptr(), cc.data_ptr(),
                                                 flats.data_ptr(), original_dem.data_ptr(),
                                                 filled_dem.data_ptr(), nx, ny,
                                                 _stream_handle())
    _lib.check(rc, "gwdt_computecosts")
    return costs, cc


def _se_dims(se) -> tuple:
    """se_dims (sx, sy, slices) of a structuring element shaped (sy, sx) or
    (slices, sy, sx), so that se_dims[0] is the fastest dimension."""
    if se.ndim == 2:
        return (se.shape[1], se.shape[0], 1)
    if se.ndim == 3:
        return (se.shape[2], se.shape[1], se.shape[0])
    raise ValueError("structuring element must be 2-D (sy, sx) or 3-D (slices, sy, sx)")


def _value_filter(op: int, dem: torch.Tensor, se: torch.Tensor, out):
    _check_grid(dem, "dem")
    if not isinstance(se, torch.Tensor) or se.dtype != torch.uint8:
        raise TypeError("structuring element must be a uint8 torch.Tensor")
    se = se.to(dem.device).contiguous()
    if out is None:
        out = torch.empty_like(dem)
    else:
        _check_grid(out, "out")
        _same_shape(dem, ("out", out))
    ny, nx = dem.shape
    with torch.cuda.device(dem.device):
        rc = _lib.lib().tt_value_filter_dev(op, out.data_ptr(), dem.data_ptr(), se.data_ptr(),
                                            _lib.se_dims(*_se_dims(se)), nx, ny,
                                            _stream_handle())
    _lib.check(rc, "max_filter" if op > 0 else "min_filter")
    return out


def min_filter(dem: torch.Tensor, se: torch.Tensor, out: torch.Tensor | None = None):
    """Minimum over a structuring element (reference min_filter,
    include/topotoolbox.h:58-102); NaN cells stay NaN."""
    return _value_filter(-1, dem, se, out)


def max_filter(dem: torch.Tensor, se: torch.Tensor, out: torch.Tensor | None = None):
    """Maximum over a structuring element (reference max_filter,
    include/topotoolbox.h:104-156); NaN cells stay NaN."""
    return _value_filter(1, dem, se, out)


def _value_filter_square(op: int, dem: torch.Tensor, width: int, out, tmp):
    _check_grid(dem, "dem")
    if not 0 <= int(width) <= 255:
        raise ValueError("width must be in 0..255 (uint8_t in the reference)")
    out = torch.empty_like(dem) if out is None else out
    tmp = torch.empty_like(dem) if tmp is None else tmp
    _check_grid(out, "out")
    _check_grid(tmp, "tmp")
    _same_shape(dem, ("out", out), ("tmp", tmp))
    ny, nx = dem.shape
    with torch.cuda.device(dem.device):
        rc = _lib.lib().tt_value_filter_square_dev(op, out.data_ptr(), dem.data_ptr(),
                                                   tmp.data_ptr(), int(width), nx, ny,
                                                   _stream_handle())
    _lib.check(rc, "max_filter_square" if op > 0 else "min_filter_square")
    return out


def min_filter_square(dem: torch.Tensor, width: int, out: torch.Tensor | None = None,
                      tmp: torch.Tensor | None = None):
    """Minimum over a width x width square (reference min_filter_square,
    include/topotoolbox.h:158-205), separably through ``tmp``."""
    return _value_filter_square(-1, dem, width, out, tmp)


def max_filter_square(dem: torch.Tensor, width: int, out: torch.Tensor | None = None,
                      tmp: torch.Tensor | None = None):
    """Maximum over a width x width square (reference max_filter_square,
    include/topotoolbox.h:207-254), separably through ``tmp``."""
    return _value_filter_square(1, dem, width, out, tmp)


# ------------------------------------------------------------------ host drop-ins
def _host(a, dtype) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.ndim != 2:
        raise ValueError("arrays must be 2-D (ny, nx)")
    return a


def fillsinks_host(dem: np.ndarray, bc: np.ndarray) -> np.ndarray:
    """Reference-signature fillsinks on host arrays (copies through the GPU)."""
    dem = _host(dem, np.float32)
    bc = _host(bc, np.uint8)
    out = np.empty_like(dem)
    ny, nx = dem.shape
    _lib.lib().tt_last_error()
    _lib.lib().fillsinks(out.ctypes.data, dem.ctypes.data, bc.ctypes.data, _lib.dims(nx, ny))
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return out


def reconstruct_host(marker: np.ndarray, mask: np.ndarray) -> np.ndarray:
    m = _host(marker, np.float32).copy()
    mask = _host(mask, np.float32)
    ny, nx = m.shape
    _lib.lib().reconstruct(m.ctypes.data, mask.ctypes.data, _lib.dims(nx, ny))
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return m


def excesstopography_fsm2d_host(dem: np.ndarray, slopes: np.ndarray,
                                cellsize: float) -> np.ndarray:
    dem = _host(dem, np.float32)
    slopes = _host(slopes, np.float32)
    out = np.empty_like(dem)
    ny, nx = dem.shape
    _lib.lib().excesstopography_fsm2d(out.ctypes.data, dem.ctypes.data, slopes.ctypes.data,
                                      float(cellsize), _lib.dims(nx, ny))
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return out


def fillsinks_hybrid_host(dem: np.ndarray, bc: np.ndarray) -> np.ndarray:
    """Reference-signature fillsinks_hybrid (the queue argument is scratch
    the GPU path does not need; same output as fillsinks)."""
    dem = _host(dem, np.float32)
    bc = _host(bc, np.uint8)
    out = np.empty_like(dem)
    queue = np.empty(dem.shape, dtype=np.intp)
    ny, nx = dem.shape
    _lib.lib().fillsinks_hybrid(out.ctypes.data, queue.ctypes.data, dem.ctypes.data,
                                bc.ctypes.data, _lib.dims(nx, ny))
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return out


def reconstruct_hybrid_host(marker: np.ndarray, mask: np.ndarray):
    """Reference-signature reconstruct_hybrid: (marker after reconstruction,
    return code: 0 when converged)."""
    m = _host(marker, np.float32).copy()
    mask = _host(mask, np.float32)
    queue = np.empty(m.shape, dtype=np.intp)
    ny, nx = m.shape
    rc = _lib.lib().reconstruct_hybrid(m.ctypes.data, queue.ctypes.data, mask.ctypes.data,
                                       _lib.dims(nx, ny))
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return m, rc


def excesstopography_fmm2d_host(dem: np.ndarray, slopes: np.ndarray,
                                cellsize: float) -> np.ndarray:
    """Reference-signature excesstopography_fmm2d (heap / back are scratch of
    the reference's priority queue; the GPU fast sweeping kernel solves the
    same equation)."""
    dem = _host(dem, np.float32)
    slopes = _host(slopes, np.float32)
    out = np.empty_like(dem)
    heap = np.empty(dem.shape, dtype=np.intp)
    back = np.empty(dem.shape, dtype=np.intp)
    ny, nx = dem.shape
    _lib.lib().excesstopography_fmm2d(out.ctypes.data, heap.ctypes.data, back.ctypes.data,
                                      dem.ctypes.data, slopes.ctypes.data, float(cellsize),
                                      _lib.dims(nx, ny))
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return out


def excesstopography_fmm3d_host(dem: np.ndarray, lithstack: np.ndarray,
                                threshold_slopes: np.ndarray, cellsize: float) -> np.ndarray:
    """Reference-signature excesstopography_fmm3d. ``lithstack`` has shape
    (ny, nx, nlayers) (layer tops, bottom up, cell-major as in the
    reference); ``threshold_slopes`` one slope per layer."""
    dem = _host(dem, np.float32)
    lith = np.ascontiguousarray(lithstack, dtype=np.float32)
    s = np.ascontiguousarray(np.asarray(threshold_slopes).reshape(-1), dtype=np.float32)
    ny, nx = dem.shape
    if lith.shape != (ny, nx, s.size):
        raise ValueError(f"lithstack has shape {lith.shape}, expected {(ny, nx, s.size)}")
    out = np.empty_like(dem)
    heap = np.empty(dem.shape, dtype=np.intp)
    back = np.empty(dem.shape, dtype=np.intp)
    _lib.lib().excesstopography_fmm3d(out.ctypes.data, heap.ctypes.data, back.ctypes.data,
                                      dem.ctypes.data, lith.ctypes.data, s.ctypes.data,
                                      float(cellsize), _lib.dims(nx, ny), s.size)
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return out


def excesstopography_fmm3d(dem: torch.Tensor, lithstack: torch.Tensor,
                           threshold_slopes: torch.Tensor, cellsize: float,
                           out: torch.Tensor | None = None, stats: bool = False):
    """Layered excess topography on the GPU (reference excesstopography_fmm3d):
    ``lithstack`` (ny, nx, nlayers) float32 layer tops, ``threshold_slopes``
    (nlayers,) float32, both on the DEM's device."""
    _check_grid(dem, "dem")
    ny, nx = dem.shape
    if (not isinstance(lithstack, torch.Tensor) or lithstack.dim() != 3
            or lithstack.dtype != torch.float32 or not lithstack.is_contiguous()
            or lithstack.shape[:2] != dem.shape or lithstack.device != dem.device):
        raise ValueError("lithstack must be a contiguous float32 (ny, nx, nlayers) tensor on "
                         "the DEM's device")
    nl = lithstack.shape[2]
    if (not isinstance(threshold_slopes, torch.Tensor) or threshold_slopes.numel() != nl
            or threshold_slopes.dtype != torch.float32 or threshold_slopes.device != dem.device):
        raise ValueError("threshold_slopes must be float32 with one slope per layer, on the "
                         "DEM's device")
    s = threshold_slopes.contiguous()
    if out is None:
        out = torch.empty_like(dem)
    st = _stats_ptr(stats)
    with torch.cuda.device(dem.device):
        rc = _lib.lib().tt_excesstopography_fmm3d_dev(out.data_ptr(), dem.data_ptr(),
                                                      lithstack.data_ptr(), s.data_ptr(),
                                                      float(cellsize), nx, ny, nl,
                                                      _stream_handle(), st)
    _lib.check(rc, "excesstopography_fmm3d")
    return _ret(out, st)


def identifyflats_host(dem: np.ndarray):
    """Reference-signature identifyflats: (int32 flags, number of flats)."""
    dem = _host(dem, np.float32)
    out = np.empty(dem.shape, dtype=np.int32)
    ny, nx = dem.shape
    n = _lib.lib().identifyflats(out.ctypes.data, dem.ctypes.data, _lib.dims(nx, ny))
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return out, int(n)


def _value_filter_host(name: str, dem: np.ndarray, se: np.ndarray) -> np.ndarray:
    dem = _host(dem, np.float32)
    se = np.ascontiguousarray(se, dtype=np.uint8)
    out = np.empty_like(dem)
    ny, nx = dem.shape
    getattr(_lib.lib(), name)(out.ctypes.data, dem.ctypes.data, se.ctypes.data,
                              _lib.dims(nx, ny), _lib.se_dims(*_se_dims(se)))
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return out


def min_filter_host(dem: np.ndarray, se: np.ndarray) -> np.ndarray:
    """Reference-signature min_filter; ``se`` shaped (sy, sx) or
    (slices, sy, sx)."""
    return _value_filter_host("min_filter", dem, se)


def max_filter_host(dem: np.ndarray, se: np.ndarray) -> np.ndarray:
    """Reference-signature max_filter; ``se`` shaped (sy, sx) or
    (slices, sy, sx)."""
    return _value_filter_host("max_filter", dem, se)


def _value_filter_square_host(name: str, dem: np.ndarray, width: int):
    dem = _host(dem, np.float32)
    out = np.empty_like(dem)
    tmp = np.empty_like(dem)
    ny, nx = dem.shape
    getattr(_lib.lib(), name)(out.ctypes.data, dem.ctypes.data, tmp.ctypes.data, int(width),
                              _lib.dims(nx, ny))
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return out, tmp


def min_filter_square_host(dem: np.ndarray, width: int):
    """Reference-signature min_filter_square: (output, tmp)."""
    return _value_filter_square_host("min_filter_square", dem, width)


def max_filter_square_host(dem: np.ndarray, width: int):
    """Reference-signature max_filter_square: (output, tmp)."""
    return _value_filter_square_host("max_filter_square", dem, width)


def gwdt_computecosts_host(flats: np.ndarray, original_dem: np.ndarray,
                           filled_dem: np.ndarray):
    """Reference-signature gwdt_computecosts: (costs, conncomps)."""
    flats = _host(flats, np.int32)
    original_dem = _host(original_dem, np.float32)
    filled_dem = _host(filled_dem, np.float32)
    costs = np.empty(flats.shape, np.float32)
    cc = np.empty(flats.shape, np.int64)
    ny, nx = flats.shape
    _lib.lib().gwdt_computecosts(costs.ctypes.data, cc.ctypes.data, flats.ctypes.data,
                                 original_dem.ctypes.data, filled_dem.ctypes.data,
                                 _lib.dims(nx, ny))
    err = _lib.last_error()
    if err:
        raise _lib.TTError(err)
    return costs, cc
