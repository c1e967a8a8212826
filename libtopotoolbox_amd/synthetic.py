"""Seeded synthetic DEMs and threshold-slope fields for tests and benchmarks.

All arrays are numpy, shape (ny, nx), C order: dims = (nx, ny) with nx the
fastest dimension. Generators emit +0.0 only (never -0.0).

* ``fractal``          spectral surface, amplitude ~ k^(-beta/2), random
                       phases, DC = 0, min-max rescaled to [0, zmax]
* ``add_pits``         lower a seeded fraction of cells by U(0, depth)
* ``add_depressions``  carve paraboloid depressions, some flat-bottomed
* ``add_nan_blob``     NaN disk centred on an edge cell
* ``voronoi_slopes``   piecewise-constant tan(slope) over Voronoi cells
* ``edge_bc``          bc = 1 on the grid edge
* ``uneven_tile_shapes`` side lengths for the batched workload
"""
from __future__ import annotations

import hashlib

import numpy as np


def fractal(ny: int, nx: int, beta: float, seed: int, zmax: float) -> np.ndarray:
    rng = np.random.default_rng(seed)
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    k2 = ky * ky + kx * kx
    k2[0, 0] = 1.0
    amp = k2 ** (-beta / 4.0)  # |k|^(-beta/2)
    amp[0, 0] = 0.0
    phase = rng.uniform(0.0, 2.0 * np.pi, size=amp.shape)
    spec = amp * np.exp(1j * phase)
    del amp, phase, k2
    z = np.fft.irfft2(spec, s=(ny, nx))
    del spec
    zmin, zmaxv = z.min(), z.max()
    z = (z - zmin) * (zmax / (zmaxv - zmin)) if zmaxv > zmin else np.zeros_like(z)
    z = z.astype(np.float32)
    z += np.float32(0.0)  # -0.0 + 0.0 = +0.0
    return z


def add_pits(z: np.ndarray, frac: float, depth: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    n = z.size
    k = int(round(frac * n))
    idx = rng.choice(n, size=k, replace=False)
    drop = rng.uniform(0.0, depth, size=k).astype(np.float32)
    flat = z.reshape(-1)
    flat[idx] = flat[idx] - drop
    flat += np.float32(0.0)
    return z


def add_depressions(z: np.ndarray, count: int, rmin: float, rmax: float, dmin: float,
                    dmax: float, flat_frac: float, seed: int) -> np.ndarray:
    """z = min(z, zc - d (1 - (r/R)^2)) inside each disk; flat-bottomed ones
    clamp the bowl at zc - d (1 - f^2), f ~ U(0.3, 0.7)."""
    rng = np.random.default_rng(seed)
    ny, nx = z.shape
    for _ in range(count):
        cy = rng.integers(0, ny)
        cx = rng.integers(0, nx)
        R = rng.uniform(rmin, rmax)
        d = rng.uniform(dmin, dmax)
        flat = rng.uniform() < flat_frac
        f = rng.uniform(0.3, 0.7)
        zc = float(z[cy, cx])
        r = int(np.ceil(R))
        y0, y1 = max(0, cy - r), min(ny, cy + r + 1)
        x0, x1 = max(0, cx - r), min(nx, cx + r + 1)
        yy = np.arange(y0, y1, dtype=np.float64)[:, None] - cy
        xx = np.arange(x0, x1, dtype=np.float64)[None, :] - cx
        rr = (yy * yy + xx * xx) / (R * R)
        bowl = zc - d * (1.0 - rr)
        if flat:
            bowl = np.maximum(bowl, zc - d * (1.0 - f * f))
        inside = rr < 1.0
        sub = z[y0:y1, x0:x1]
        cand = bowl.astype(np.float32)
        np.copyto(sub, np.minimum(sub, cand), where=inside)
    z += np.float32(0.0)
    return z


def add_nan_blob(z: np.ndarray, frac: float, seed: int) -> np.ndarray:
    """NaN disk centred on a random cell of the left edge covering about
    `frac` of the grid (half a disk lies inside)."""
    rng = np.random.default_rng(seed)
    ny, nx = z.shape
    r = np.sqrt(2.0 * frac * nx * ny / np.pi)
    cy = rng.integers(0, ny)
    yy = np.arange(ny, dtype=np.float64)[:, None] - cy
    xx = np.arange(nx, dtype=np.float64)[None, :]
    z[(yy * yy + xx * xx) < r * r] = np.nan
    return z


def voronoi_slopes(ny: int, nx: int, nsites: int, lo: float, hi: float, seed: int,
                   device=None) -> np.ndarray:
    """tan(threshold slope) constant over the Voronoi cell of each of
    `nsites` integer sites (nearest in Euclidean distance, lowest site index
    on ties; exact integer arithmetic)."""
    rng = np.random.default_rng(seed)
    sy = rng.integers(0, ny, size=nsites)
    sx = rng.integers(0, nx, size=nsites)
    val = rng.uniform(lo, hi, size=nsites).astype(np.float32)
    try:
        import torch
        dev = torch.device(device) if device is not None else torch.device("cpu")
        Y = torch.arange(ny, device=dev, dtype=torch.int64)[:, None]
        X = torch.arange(nx, device=dev, dtype=torch.int64)[None, :]
        best = torch.full((ny, nx), np.iinfo(np.int64).max, dtype=torch.int64, device=dev)
        arg = torch.zeros((ny, nx), dtype=torch.int64, device=dev)
        for s in range(nsites):
            dy = (Y - int(sy[s])) ** 2
            dx = (X - int(sx[s])) ** 2
            d2 = dy + dx
            closer = d2 < best
            best = torch.where(closer, d2, best)
            arg = torch.where(closer, torch.full_like(arg, s), arg)
        idx = arg.cpu().numpy()
    except ImportError:
        Y = np.arange(ny, dtype=np.int64)[:, None]
        X = np.arange(nx, dtype=np.int64)[None, :]
        best = np.full((ny, nx), np.iinfo(np.int64).max, dtype=np.int64)
        idx = np.zeros((ny, nx), dtype=np.int64)
        for s in range(nsites):
            d2 = (Y - sy[s]) ** 2 + (X - sx[s]) ** 2
            closer = d2 < best
            best[closer] = d2[closer]
            idx[closer] = s
    return val[idx]


def edge_bc(ny: int, nx: int) -> np.ndarray:
    bc = np.zeros((ny, nx), np.uint8)
    bc[0, :] = 1
    bc[-1, :] = 1
    bc[:, 0] = 1
    bc[:, -1] = 1
    return bc


def nan_to_bc(dem: np.ndarray, bc: np.ndarray):
    """Reference caller convention for NaN cells (include/topotoolbox.h:
    279-281 and test/grid.c:505-529): bc = 1 and dem = -INF there. Returns
    (dem', bc', nan_mask); restore NaN in the output afterwards."""
    nanm = np.isnan(dem)
    dem2 = dem.copy()
    bc2 = bc.copy()
    dem2[nanm] = -np.inf
    bc2[nanm] = 1
    return dem2, bc2, nanm


def uneven_tile_shapes(count: int, smin: int, smax: int, seed: int):
    """(ny, nx) pairs: side s = exp(U(ln smin, ln smax)), aspect a ~ U(0.5, 2),
    nx = round(s sqrt(a)), ny = round(s / sqrt(a))."""
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(np.log(smin), np.log(smax), size=count))
    a = rng.uniform(0.5, 2.0, size=count)
    nx = np.maximum(1, np.rint(s * np.sqrt(a))).astype(int)
    ny = np.maximum(1, np.rint(s / np.sqrt(a))).astype(int)
    return list(zip(ny.tolist(), nx.tolist()))


# ---------------------------------------------------------------- named workloads
def dem_pits(ny: int, nx: int, seed: int) -> np.ndarray:
    """Fractal (beta 2.6, i.e. Hurst 0.8), 0-1000 m, 5% single-cell pits
    lowered by U(0, 20) m."""
    z = fractal(ny, nx, 2.6, seed, 1000.0)
    return add_pits(z, 0.05, 20.0, seed + 1_000_003)


def dem_depressions(ny: int, nx: int, seed: int, nan_blob: bool = True) -> np.ndarray:
    """Fractal (beta 2.4, Hurst 0.7), 0-3000 m, 200 paraboloid depressions
    (R ~ U(10, 300) cells, depth U(1, 50) m, 30% flat-bottomed), and a NaN
    blob of 2% of the cells touching the left edge."""
    z = fractal(ny, nx, 2.4, seed, 3000.0)
    z = add_depressions(z, 200, 10.0, 300.0, 1.0, 50.0, 0.3, seed + 7)
    if nan_blob:
        z = add_nan_blob(z, 0.02, seed + 11)
    return z


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def config5_items(op: str, count: int = 512, seed: int = 5, workers: int = 16):
    """The batched workload of uneven tiles: side exp(U(ln 128, ln 2048)),
    aspect U(0.5, 2); per-tile fractal (Hurst 0.8, 0-1000 m, 5% pits) with
    seed 5000 + t; edge bc (op "fillsinks") or 8-site Voronoi slopes
    U(0.2, 0.8) with seed 9000 + t (op "excesstopography"); upstream
    gradient N(0, 1) with seed 7000 + t. Returns (shapes, dems, aux, g)."""
    from concurrent.futures import ThreadPoolExecutor

    shapes = uneven_tile_shapes(count, 128, 2048, seed)

    def one(t):
        ny, nx = shapes[t]
        d = dem_pits(ny, nx, 5000 + t)
        g = np.random.default_rng(7000 + t).standard_normal((ny, nx), dtype=np.float32)
        aux = (edge_bc(ny, nx) if op == "fillsinks"
               else voronoi_slopes(ny, nx, 8, 0.2, 0.8, 9000 + t))
        return d, aux, g

    with ThreadPoolExecutor(max(1, workers)) as ex:
        res = list(ex.map(one, range(count)))
    return shapes, [r[0] for r in res], [r[1] for r in res], [r[2] for r in res]

