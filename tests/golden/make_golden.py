"""Write the golden fixtures in this folder from the reference library.

The reference hot path is compiled from its own sources (``make -C oracle
ref`` -> oracle/_ref/libtopotoolbox_ref.so); this script runs it on the
inputs of the reference's own tests and stores inputs and outputs as .npz.
Only data is stored. Run from the repo root:

    python tests/golden/make_golden.py

Inputs:
* fillsinks: test/random_dem.cpp:629-647 + test/grid.c:136-168 (GridRandom,
  dims {100, 200}, edge bc), plus a NaN variant under the caller convention
  of test/grid.c:505-529 and a raw-NaN variant.
* reconstruct: random mask with a marker below it on 30% of cells.
* excesstopography: test/excesstopography.cpp:97-112 (dims {200, 100},
  row 0 NaN, cellsize 30) and the numerics case of :156-196.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def edge_bc(shape):
    bc = np.zeros(shape, np.uint8)
    bc[0, :] = bc[-1, :] = 1
    bc[:, 0] = bc[:, -1] = 1
    return bc


def pcg_grid(shape, c, d, scale=1.0):
    ny, nx = shape
    out = np.empty(shape, np.float32)
    for j in range(ny):
        for i in range(nx):
            out[j, i] = scale * oracle.pcg4d(i, j, c, d)
    return out


def main():
    oracle.build(ref=True)
    R = oracle.ref()
    assert R is not None and R.has_topotoolbox() == 1
    cases = {}
    # --- pcg4d known answers
    rng0 = np.random.default_rng(0)
    args = rng0.integers(0, 2**32, size=(64, 4), dtype=np.uint64).astype(np.uint32)
    args[:8] = [[0, 0, 0, 0], [1, 2, 3, 4], [199, 99, 0, 1], [2**32 - 1] * 4,
                [5, 0, 0, 0], [0, 5, 0, 0], [0, 0, 5, 0], [0, 0, 0, 5]]
    vals = np.array([R.pcg4d(*[int(v) for v in a]) for a in args], np.float32)
    cases["pcg4d_kat"] = dict(args=args, values=vals)
    # --- fillsinks on GridRandom, dims {100, 200} -> numpy (200, 100)
    for seed in (0, 1, 2):
        dem = oracle.grid_random(seed, (200, 100))
        bc = edge_bc(dem.shape)
        cases[f"fillsinks_random_s{seed}"] = dict(dem=dem, bc=bc, out=oracle.ref_fillsinks(dem, bc))
    # --- NaN under the caller convention and raw NaN
    dem = oracle.grid_random(7, (200, 100))
    rng = np.random.default_rng(7)
    nanm = rng.uniform(size=dem.shape) < 0.05
    nanm[50:80, 20:60] = True
    dem_nan = dem.copy()
    dem_nan[nanm] = np.nan
    bc = edge_bc(dem.shape)
    cases["fillsinks_nan_raw"] = dict(dem=dem_nan, bc=bc, out=oracle.ref_fillsinks(dem_nan, bc))
    dem_c = dem_nan.copy()
    bc_c = bc.copy()
    dem_c[nanm] = -np.inf
    bc_c[nanm] = 1
    out_c = oracle.ref_fillsinks(dem_c, bc_c)
    cases["fillsinks_nan_convention"] = dict(dem=dem_c, bc=bc_c, out=out_c)
    # --- flats: integer-valued DEM with many exact ties
    dem_i = np.floor(oracle.grid_random(11, (150, 120)) / 10.0).astype(np.float32)
    bc_i = edge_bc(dem_i.shape)
    cases["fillsinks_flats"] = dict(dem=dem_i, bc=bc_i, out=oracle.ref_fillsinks(dem_i, bc_i))
    # --- reconstruct
    mask = oracle.grid_random(21, (120, 160))
    marker = np.where(rng.uniform(size=mask.shape) < 0.3, mask - 5.0, -np.inf).astype(np.float32)
    cases["reconstruct_random"] = dict(marker=marker, mask=mask,
                                       out=oracle.ref_reconstruct(marker, mask))
    # --- excess topography, test/excesstopography.cpp inputs (dims {200, 100})
    for seed in (0, 1, 2):
        dem = pcg_grid((100, 200), seed, 0, 100.0)
        dem[:, 0] = np.nan  # row index i == 0 is NaN (column 0 in numpy layout)
        slopes = pcg_grid((100, 200), seed, 1)
        cases[f"fsm_random_s{seed}"] = dict(
            dem=dem, slopes=slopes, cellsize=np.float32(30.0),
            fsm=oracle.ref_fsm2d(dem, slopes, 30.0), fmm=oracle.ref_fmm2d(dem, slopes, 30.0))
    # --- eikonal numerics: boundary 0, interior +INF
    dem = np.zeros((100, 200), np.float32)
    dem[1:-1, 1:-1] = np.inf
    slopes = pcg_grid((100, 200), 3, 0)
    cases["fsm_numerics"] = dict(dem=dem, slopes=slopes, cellsize=np.float32(30.0),
                                 fsm=oracle.ref_fsm2d(dem, slopes, 30.0),
                                 fmm=oracle.ref_fmm2d(dem, slopes, 30.0))
    for name, arrays in cases.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **arrays)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
