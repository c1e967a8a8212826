"""CPU oracle for fillsinks / reconstruct / excesstopography_fsm2d.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the libtopotoolbox_amd product.

Two libraries are wrapped here:

* ``liboracle.so`` built from ``tt_oracle.c`` (the C restatement of the
  reference path plus independent cross-checks; see that file's header for
  the reference lines each function follows), and
* ``_ref/libtopotoolbox_ref.so``, the reference hot path compiled from its
  own sources by ``make ref`` where those sources exist. It is optional:
  ``ref()`` returns None when it has not been built.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libtopotoolbox_ref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_dims_t = ctypes.c_int64 * 2
_se_dims_t = ctypes.c_int64 * 3
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build(ref: bool = False) -> None:
    """Compile the oracle (and the reference library when ``ref``)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None
_ref = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.oracle_pcg4d.restype = ctypes.c_float
        L.oracle_pcg4d.argtypes = [ctypes.c_uint32] * 4
        L.oracle_grid_random.argtypes = [_f32p, ctypes.c_uint64, _dims_t]
        L.oracle_reconstruct.restype = ctypes.c_int64
        L.oracle_reconstruct.argtypes = [_f32p, _f32p, _dims_t, ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_int32)]
        L.oracle_fillsinks.restype = ctypes.c_int64
        L.oracle_fillsinks.argtypes = [_f32p, _f32p, _u8p, _dims_t, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_int32)]
        L.oracle_fillsinks_spill.restype = ctypes.c_int
        L.oracle_fillsinks_spill.argtypes = [_f32p, ctypes.c_void_p, _f32p, _u8p, _dims_t]
        L.oracle_fillsinks_adjoint_f64.argtypes = [_f64p, _i64p, _f64p, _dims_t]
        L.oracle_fsm2d_f32.restype = ctypes.c_int64
        L.oracle_fsm2d_f32.argtypes = [_f32p, _f32p, _f32p, ctypes.c_float, _dims_t,
                                       ctypes.c_int64]
        L.oracle_fsm2d_f64.restype = ctypes.c_int64
        L.oracle_fsm2d_f64.argtypes = [_f64p, _f64p, _f64p, ctypes.c_double, _dims_t,
                                       ctypes.c_int64]
        L.oracle_fmm2d_f64.restype = ctypes.c_int
        L.oracle_fmm2d_f64.argtypes = [_f64p, _f64p, _f64p, ctypes.c_double, _dims_t]
        L.oracle_identifyflats.restype = ctypes.c_ssize_t
        L.oracle_identifyflats.argtypes = [_i32p, _f32p, _dims_t]
        L.oracle_value_filter.restype = None
        L.oracle_value_filter.argtypes = [_f32p, _f32p, _u8p, _dims_t, _se_dims_t, ctypes.c_int]
        L.oracle_value_filter_square.restype = None
        L.oracle_value_filter_square.argtypes = [_f32p, _f32p, _f32p, ctypes.c_int, _dims_t,
                                                 ctypes.c_int]
        L.oracle_gwdt_computecosts.restype = None
        L.oracle_gwdt_computecosts.argtypes = [_f32p, _i64p, _i32p, _f32p, _f32p, _dims_t]
        L.oracle_fsm2d_adjoint_f64.restype = ctypes.c_int
        L.oracle_fsm2d_adjoint_f64.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p,
                                               ctypes.c_double, _f64p, _dims_t]
        _lib = L
    return _lib


def ref():
    """The reference library compiled from its sources, or None."""
    global _ref
    if _ref is None and os.path.exists(_REF):
        R = ctypes.CDLL(_REF)
        R.fillsinks.restype = None
        R.fillsinks.argtypes = [_f32p, _f32p, _u8p, _dims_t]
        R.reconstruct.restype = None
        R.reconstruct.argtypes = [_f32p, _f32p, _dims_t]
        R.excesstopography_fsm2d.restype = None
        R.excesstopography_fsm2d.argtypes = [_f32p, _f32p, _f32p, ctypes.c_float, _dims_t]
        R.excesstopography_fmm2d.restype = None
        R.excesstopography_fmm2d.argtypes = [_f32p, _i64p, _i64p, _f32p, _f32p,
                                             ctypes.c_float, _dims_t]
        R.excesstopography_fmm3d.restype = None
        R.excesstopography_fmm3d.argtypes = [_f32p, _i64p, _i64p, _f32p, _f32p, _f32p,
                                             ctypes.c_float, _dims_t, ctypes.c_ssize_t]
        R.fillsinks_hybrid.restype = None
        R.fillsinks_hybrid.argtypes = [_f32p, _i64p, _f32p, _u8p, _dims_t]
        R.pcg4d.restype = ctypes.c_float
        R.pcg4d.argtypes = [ctypes.c_uint32] * 4
        R.has_topotoolbox.restype = ctypes.c_int
        R.identifyflats.restype = ctypes.c_ssize_t
        R.identifyflats.argtypes = [_i32p, _f32p, _dims_t]
        for name in ("min_filter", "max_filter"):
            getattr(R, name).restype = None
            getattr(R, name).argtypes = [_f32p, _f32p, _u8p, _dims_t, _se_dims_t]
        for name in ("min_filter_square", "max_filter_square"):
            getattr(R, name).restype = None
            getattr(R, name).argtypes = [_f32p, _f32p, _f32p, ctypes.c_uint8, _dims_t]
        R.gwdt_computecosts.restype = None
        R.gwdt_computecosts.argtypes = [_f32p, _i64p, _i32p, _f32p, _f32p, _dims_t]
        _ref = R
    return _ref


def _dims(a: np.ndarray) -> "ctypes.Array":
    # numpy (rows, cols) C-order array: dims[0] = cols (fastest), dims[1] = rows
    ny, nx = a.shape
    return _dims_t(nx, ny)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------- generators
def pcg4d(a: int, b: int, c: int, d: int) -> float:
    return lib().oracle_pcg4d(a, b, c, d)


def grid_random(seed: int, shape) -> np.ndarray:
    out = np.empty(shape, np.float32)
    lib().oracle_grid_random(out, seed, _dims(out))
    return out


# ---------------------------------------------------------------- fill
def fillsinks(dem: np.ndarray, bc: np.ndarray, max_iterations: int = 1000):
    """Restated reference fillsinks. Returns (filled, iterations, converged)."""
    dem = _c(dem, np.float32).copy()
    bc = _c(bc, np.uint8)
    out = np.empty_like(dem)
    conv = ctypes.c_int32(0)
    it = lib().oracle_fillsinks(out, dem, bc, _dims(dem), max_iterations, ctypes.byref(conv))
    return out, int(it), bool(conv.value)


def reconstruct(marker: np.ndarray, mask: np.ndarray, max_iterations: int = 1000):
    """Restated reference reconstruct. Returns (marker, iterations, converged)."""
    m = _c(marker, np.float32).copy()
    mask = _c(mask, np.float32)
    conv = ctypes.c_int32(0)
    it = lib().oracle_reconstruct(m, mask, _dims(m), max_iterations, ctypes.byref(conv))
    return m, int(it), bool(conv.value)


def fillsinks_pf(dem: np.ndarray, bc: np.ndarray) -> np.ndarray:
    """Priority-Flood fill (textbook cross-check)."""
    dem = _c(dem, np.float32)
    bc = _c(bc, np.uint8)
    out = np.empty_like(dem)
    if lib().oracle_fillsinks_spill(out, None, dem, bc, _dims(dem)):
        raise MemoryError("oracle priority-flood allocation failed")
    return out


def fillsinks_spill(dem: np.ndarray, bc: np.ndarray):
    """Fill level and lexicographic spill-cell index (int64, -1 = none)."""
    dem = _c(dem, np.float32)
    bc = _c(bc, np.uint8)
    out = np.empty_like(dem)
    spill = np.empty(dem.shape, np.int64)
    if lib().oracle_fillsinks_spill(out, spill.ctypes.data, dem, bc, _dims(dem)):
        raise MemoryError("oracle priority-flood allocation failed")
    return out, spill


def fillsinks_adjoint_f64(spill: np.ndarray, g_out: np.ndarray) -> np.ndarray:
    spill = _c(spill, np.int64)
    g = _c(g_out, np.float64)
    g_dem = np.empty(spill.shape, np.float64)
    lib().oracle_fillsinks_adjoint_f64(g_dem, spill, g, _dims(g))
    return g_dem


# ---------------------------------------------------------------- excess topography
def fsm2d_f32(dem, slopes, cellsize, max_iterations: int = 0):
    """Restated reference FSM in float32. Returns (u, iterations)."""
    dem = _c(dem, np.float32)
    s = _c(slopes, np.float32)
    u = np.empty_like(dem)
    it = lib().oracle_fsm2d_f32(u, dem, s, np.float32(cellsize), _dims(dem), max_iterations)
    return u, int(it)


def fsm2d_f64(dem, slopes, cellsize, max_iterations: int = 0):
    """FP64 FSM oracle. Inputs are promoted to float64. Returns (u, iterations)."""
    dem = _c(dem, np.float64)
    s = _c(slopes, np.float64)
    u = np.empty_like(dem)
    it = lib().oracle_fsm2d_f64(u, dem, s, float(cellsize), _dims(dem), max_iterations)
    return u, int(it)


def fmm2d_f64(dem, slopes, cellsize):
    dem = _c(dem, np.float64)
    s = _c(slopes, np.float64)
    u = np.empty_like(dem)
    if lib().oracle_fmm2d_f64(u, dem, s, float(cellsize), _dims(dem)):
        raise MemoryError("oracle fmm allocation failed")
    return u


def fsm2d_adjoint_f64(u, dem, slopes, cellsize, g_out):
    """FP64 adjoint at converged state u. Returns (g_dem, g_slopes)."""
    u = _c(u, np.float64)
    dem = _c(dem, np.float64)
    s = _c(slopes, np.float64)
    g = _c(g_out, np.float64)
    g_dem = np.empty_like(u)
    g_s = np.empty_like(u)
    if lib().oracle_fsm2d_adjoint_f64(g_dem, g_s, u, dem, s, float(cellsize), g, _dims(u)):
        raise MemoryError("oracle adjoint allocation failed")
    return g_dem, g_s


# ---------------------------------------------------------------- reference (optional)
def ref_fillsinks(dem, bc):
    R = ref()
    dem = _c(dem, np.float32).copy()
    bc = _c(bc, np.uint8)
    out = np.empty_like(dem)
    R.fillsinks(out, dem, bc, _dims(dem))
    return out


def ref_fillsinks_hybrid(dem, bc):
    R = ref()
    dem = _c(dem, np.float32).copy()
    bc = _c(bc, np.uint8)
    out = np.empty_like(dem)
    q = np.empty(dem.shape, np.int64)
    R.fillsinks_hybrid(out, q, dem, bc, _dims(dem))
    return out


def ref_reconstruct(marker, mask):
    R = ref()
    m = _c(marker, np.float32).copy()
    R.reconstruct(m, _c(mask, np.float32), _dims(m))
    return m


def ref_fsm2d(dem, slopes, cellsize):
    R = ref()
    dem = _c(dem, np.float32)
    u = np.empty_like(dem)
    R.excesstopography_fsm2d(u, dem, _c(slopes, np.float32), np.float32(cellsize), _dims(dem))
    return u


def ref_fmm3d(dem, lithstack, slopes, cellsize):
    """The compiled reference's excesstopography_fmm3d; lithstack (ny, nx,
    nlayers), slopes (nlayers,)."""
    R = ref()
    dem = _c(dem, np.float32)
    lith = _c(lithstack, np.float32)
    s = _c(np.asarray(slopes).reshape(-1), np.float32)
    u = np.empty_like(dem)
    heap = np.empty(dem.shape, np.int64)
    back = np.empty(dem.shape, np.int64)
    R.excesstopography_fmm3d(u, heap, back, dem, lith, s, np.float32(cellsize), _dims(dem),
                             s.size)
    return u


def ref_fmm2d(dem, slopes, cellsize):
    R = ref()
    dem = _c(dem, np.float32)
    u = np.empty_like(dem)
    heap = np.empty(dem.shape, np.int64)
    back = np.empty(dem.shape, np.int64)
    R.excesstopography_fmm2d(u, heap, back, dem, _c(slopes, np.float32),
                             np.float32(cellsize), _dims(dem))
    return u


# ------------------------------------------------------------------ flats, value filters
def _se_dims(se: np.ndarray):
    """(ny, nx) or (slices, ny, nx) numpy SE -> se_dims = (nx, ny, slices)."""
    se = np.asarray(se)
    if se.ndim == 2:
        se = se[None]
    return np.ascontiguousarray(se, np.uint8), _se_dims_t(se.shape[2], se.shape[1], se.shape[0])


def identifyflats(dem):
    """(int32 bit field, number of flats): bit 0 flat, 1 sill, 2 presill."""
    dem = _c(dem, np.float32)
    out = np.empty(dem.shape, np.int32)
    n = lib().oracle_identifyflats(out, dem, _dims(dem))
    return out, int(n)


def value_filter(dem, se, op: str):
    """min_filter (op "min") / max_filter (op "max") over structuring element se."""
    dem = _c(dem, np.float32)
    se, sd = _se_dims(se)
    out = np.empty_like(dem)
    lib().oracle_value_filter(out, dem, se, _dims(dem), sd, -1 if op == "min" else 1)
    return out


def value_filter_square(dem, width: int, op: str):
    dem = _c(dem, np.float32)
    out = np.empty_like(dem)
    tmp = np.empty_like(dem)
    lib().oracle_value_filter_square(out, dem, tmp, int(width), _dims(dem),
                                     -1 if op == "min" else 1)
    return out


def ref_identifyflats(dem):
    dem = _c(dem, np.float32).copy()
    out = np.empty(dem.shape, np.int32)
    n = ref().identifyflats(out, dem, _dims(dem))
    return out, int(n)


def ref_value_filter(dem, se, op: str):
    dem = _c(dem, np.float32).copy()
    se, sd = _se_dims(se)
    out = np.empty_like(dem)
    getattr(ref(), f"{op}_filter")(out, dem, se, _dims(dem), sd)
    return out


def ref_value_filter_square(dem, width: int, op: str):
    dem = _c(dem, np.float32).copy()
    out = np.empty_like(dem)
    tmp = np.empty_like(dem)
    getattr(ref(), f"{op}_filter_square")(out, dem, tmp, int(width), _dims(dem))
    return out



def gwdt_computecosts(flats, original, filled):
    """(costs float32, conncomps int64) of the gray-weighted distance
    transform (src/gwdt.c:166-274)."""
    flats = _c(flats, np.int32)
    original = _c(original, np.float32)
    filled = _c(filled, np.float32)
    costs = np.empty(flats.shape, np.float32)
    cc = np.empty(flats.shape, np.int64)
    lib().oracle_gwdt_computecosts(costs, cc, flats, original, filled, _dims(flats))
    return costs, cc


def ref_gwdt_computecosts(flats, original, filled):
    flats = _c(flats, np.int32).copy()
    original = _c(original, np.float32).copy()
    filled = _c(filled, np.float32).copy()
    costs = np.empty(flats.shape, np.float32)
    cc = np.empty(flats.shape, np.int64)
    ref().gwdt_computecosts(costs, cc, flats, original, filled, _dims(flats))
    return costs, cc
