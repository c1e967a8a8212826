"""GPU parity: the HIP path (through the C ABI) against the oracle.

fillsinks / reconstruct: bit-exact (the reconstruction fixed point is
unique and made of input values). excesstopography_fsm2d: within
max |u - u64| <= 1e-4 * max(|u64|, 1 m) of the FP64 oracle, and within
the same bound of the reference-order float restatement.
"""
import numpy as np
import pytest
import torch

from conftest import assert_bitwise, load_golden

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 37), (37, 1), (2, 3), (64, 64), (65, 129), (200, 100), (257, 300),
          (130, 513)]
TILES = [(2, 2), (2, 1), (1, 2), (1, 1)]
FSM_REL = 1e-4


def edge_bc(shape):
    bc = np.zeros(shape, np.uint8)
    bc[0, :] = bc[-1, :] = 1
    bc[:, 0] = bc[:, -1] = 1
    return bc


def gpu_fill(dem, bc, tile=(2, 2)):
    import libtopotoolbox_amd as tt

    tt.set_tile_shape(*tile)
    d = torch.from_numpy(np.ascontiguousarray(dem)).cuda()
    b = torch.from_numpy(np.ascontiguousarray(bc)).cuda()
    out, st = tt.fillsinks(d, b, stats=True)
    assert st["error"] == 0, st
    return out.cpu().numpy(), st


def gpu_fsm(dem, slopes, cs, tile=(2, 2)):
    import libtopotoolbox_amd as tt

    tt.set_tile_shape(*tile)
    d = torch.from_numpy(np.ascontiguousarray(dem, np.float32)).cuda()
    s = torch.from_numpy(np.ascontiguousarray(slopes, np.float32)).cuda()
    out, st = tt.excesstopography_fsm2d(d, s, cs, stats=True)
    assert st["error"] == 0, st
    return out.cpu().numpy(), st


def fsm_close(u, u64, what):
    nan32, nan64 = np.isnan(u), np.isnan(u64)
    np.testing.assert_array_equal(nan32, nan64, err_msg=what + " NaN pattern")
    inf = np.isinf(u64)
    np.testing.assert_array_equal(u[inf], u64[inf].astype(np.float32), err_msg=what + " INF")
    ok = ~nan64 & ~inf
    err = np.abs(u[ok].astype(np.float64) - u64[ok]) / np.maximum(np.abs(u64[ok]), 1.0)
    assert err.size == 0 or err.max() <= FSM_REL, f"{what}: max rel err {err.max():.3e}"
    return 0.0 if err.size == 0 else float(err.max())


# ------------------------------------------------------------------ fill
@pytest.mark.parametrize("name", ["fillsinks_random_s0", "fillsinks_random_s1",
                                  "fillsinks_random_s2", "fillsinks_nan_raw",
                                  "fillsinks_nan_convention", "fillsinks_flats"])
def test_fillsinks_golden(cuda, name):
    g = load_golden(name)
    for tile in TILES:
        out, _ = gpu_fill(g["dem"], g["bc"], tile)
        assert_bitwise(out, g["out"], f"{name} tile {tile}")


@pytest.mark.parametrize("shape", SHAPES)
def test_fillsinks_shapes(cuda, oracle_lib, shape):
    rng = np.random.default_rng(hash(shape) % 2**32)
    dem = oracle_lib.grid_random(shape[0] * 1000 + shape[1], shape)
    dem[rng.uniform(size=shape) < 0.02] = np.nan
    bc = edge_bc(shape)
    want = oracle_lib.fillsinks_pf(dem, bc)
    for tile in TILES:
        out, _ = gpu_fill(dem, bc, tile)
        assert_bitwise(out, want, f"shape {shape} tile {tile}")


def test_results_without_statistics(cuda, oracle_lib):
    """The default path collects no visit statistics (the visits skip their
    counter atomics). Its fill and spill outputs are bit-identical to the
    path with statistics and to the oracle; excess topography without
    statistics stays within the FP64 tolerance."""
    import libtopotoolbox_amd as tt

    shape = (389, 517)  # partial tiles on both axes, nx % 4 != 0
    rng = np.random.default_rng(23)
    dem = oracle_lib.grid_random(7, shape)
    dem[rng.uniform(size=shape) < 0.02] = np.nan
    bc = edge_bc(shape)
    want = oracle_lib.fillsinks_pf(dem, bc)
    d = torch.from_numpy(np.ascontiguousarray(dem)).cuda()
    b = torch.from_numpy(bc).cuda()
    for tile in TILES:
        tt.set_tile_shape(*tile)
        off = tt.fillsinks(d, b).cpu().numpy()
        assert_bitwise(off, want, f"fill without statistics, tile {tile}")
        (lev, sp) = tt.ops.fillsinks_spill(d, b)
        (lev_s, sp_s), st = tt.ops.fillsinks_spill(d, b, stats=True)
        assert st["error"] == 0, st
        assert_bitwise(lev.cpu().numpy(), want, f"spill levels, tile {tile}")
        np.testing.assert_array_equal(sp.cpu().numpy(), sp_s.cpu().numpy(),
                                      err_msg=f"spill cells with / without statistics, {tile}")
    slopes = rng.uniform(0.05, 0.9, size=shape).astype(np.float32)
    u64, _ = oracle_lib.fsm2d_f64(dem, slopes, 30.0)
    s = torch.from_numpy(slopes).cuda()
    for tile in TILES:
        tt.set_tile_shape(*tile)
        u = tt.excesstopography_fsm2d(d, s, 30.0).cpu().numpy()
        fsm_close(u, u64, f"fsm without statistics, tile {tile}")


def test_fillsinks_special_values(cuda, oracle_lib):
    rng = np.random.default_rng(5)
    dem = np.floor(rng.uniform(0, 8, size=(190, 170))).astype(np.float32)  # flats
    dem[rng.uniform(size=dem.shape) < 0.02] = -np.inf
    dem[rng.uniform(size=dem.shape) < 0.02] = np.inf
    dem[rng.uniform(size=dem.shape) < 0.05] = np.nan
    for bc in (edge_bc(dem.shape), (rng.uniform(size=dem.shape) < 0.01).astype(np.uint8),
               np.zeros(dem.shape, np.uint8), np.full(dem.shape, 2, np.uint8)):
        want, _, conv = oracle_lib.fillsinks(dem, bc)
        assert conv
        out, _ = gpu_fill(dem, bc)
        assert_bitwise(out, want, "special values")


def test_fillsinks_pits_1024(cuda, oracle_lib):
    from libtopotoolbox_amd import synthetic as S

    dem = S.dem_pits(1024, 1024, 1)
    bc = edge_bc(dem.shape)
    want, it, conv = oracle_lib.fillsinks(dem, bc)
    assert conv and it < 1000
    out, st = gpu_fill(dem, bc)
    assert_bitwise(out, want, "config-1 1024^2")
    out2, _ = gpu_fill(dem, bc)
    assert_bitwise(out2, out, "determinism")
    print("stats", st)


def test_reconstruct(cuda, oracle_lib):
    import libtopotoolbox_amd as tt

    g = load_golden("reconstruct_random")
    m = torch.from_numpy(g["marker"]).cuda()
    k = torch.from_numpy(g["mask"]).cuda()
    out = tt.reconstruct(m, k).cpu().numpy()
    assert_bitwise(out, g["out"], "reconstruct golden")
    rng = np.random.default_rng(9)
    mask = oracle_lib.grid_random(77, (300, 260))
    mask[rng.uniform(size=mask.shape) < 0.03] = np.nan
    marker = np.where(rng.uniform(size=mask.shape) < 0.1, mask + 7.0, -np.inf).astype(np.float32)
    want, _, conv = oracle_lib.reconstruct(marker, mask)
    assert conv
    out = tt.reconstruct(torch.from_numpy(marker).cuda(), torch.from_numpy(mask).cuda())
    assert_bitwise(out.cpu().numpy(), want, "reconstruct random")


def test_hybrid_and_fmm_dropins(cuda, oracle_lib):
    """fillsinks_hybrid / reconstruct_hybrid (reference src/fillsinks.c:50-77,
    src/reconstruct.c:291-312) give the same bits as the oracle's fillsinks /
    reconstruct; excesstopography_fmm2d is within the reference's own
    FSM-vs-FMM bound (1e-4, test/excesstopography.cpp:69-82) of the
    fast-marching oracle, relative with denominator max(|u|, 1 m)."""
    import libtopotoolbox_amd as tt

    g = load_golden("fillsinks_nan_convention")
    assert_bitwise(tt.fillsinks_hybrid_host(g["dem"], g["bc"]), g["out"], "fillsinks_hybrid")
    rng = np.random.default_rng(21)
    mask = oracle_lib.grid_random(79, (120, 170))
    marker = np.where(rng.uniform(size=mask.shape) < 0.1, mask - 3.0, -np.inf).astype(np.float32)
    want, _, conv = oracle_lib.reconstruct(marker, mask)
    assert conv
    got, rc = tt.reconstruct_hybrid_host(marker, mask)
    assert rc == 0
    assert_bitwise(got, want, "reconstruct_hybrid")
    dem = (rng.uniform(size=(150, 130)) * 300).astype(np.float32)
    slopes = rng.uniform(0.1, 0.9, size=dem.shape).astype(np.float32)
    u = tt.excesstopography_fmm2d_host(dem, slopes, 30.0)
    fsm_close(u, oracle_lib.fmm2d_f64(dem, slopes, 30.0), "fmm2d drop-in vs FP64 fast marching")
    if oracle_lib.ref() is not None:
        fsm_close(u, oracle_lib.ref_fmm2d(dem, slopes, 30.0).astype(np.float64),
                  "fmm2d drop-in vs the compiled reference")


def test_host_dropins(cuda, oracle_lib):
    import libtopotoolbox_amd as tt

    g = load_golden("fillsinks_nan_raw")
    assert_bitwise(tt.fillsinks_host(g["dem"], g["bc"]), g["out"], "fillsinks host")
    g = load_golden("reconstruct_random")
    assert_bitwise(tt.reconstruct_host(g["marker"], g["mask"]), g["out"], "reconstruct host")
    g = load_golden("fsm_random_s0")
    u = tt.excesstopography_fsm2d_host(g["dem"], g["slopes"], float(g["cellsize"]))
    u64, _ = oracle_lib.fsm2d_f64(g["dem"], g["slopes"], float(g["cellsize"]))
    fsm_close(u, u64, "fsm host")


# ------------------------------------------------------------------ excess topography
@pytest.mark.parametrize("name", ["fsm_random_s0", "fsm_random_s1", "fsm_random_s2",
                                  "fsm_numerics"])
def test_fsm_golden(cuda, oracle_lib, name):
    g = load_golden(name)
    cs = float(g["cellsize"])
    u64, _ = oracle_lib.fsm2d_f64(g["dem"], g["slopes"], cs)
    for tile in TILES:
        u, _ = gpu_fsm(g["dem"], g["slopes"], cs, tile)
        fsm_close(u, u64, f"{name} tile {tile}")
        # the reference's own FSM==FMM criterion against its float output
        fin = np.isfinite(g["fsm"])
        tol = 1e-4 if name != "fsm_numerics" else 1e-6 * np.maximum(1.0, np.abs(g["fsm"][fin]))
        assert np.all(np.abs(u[fin] - g["fsm"][fin]) <= tol)
        assert np.all((u <= g["dem"]) | np.isnan(g["dem"]))


@pytest.mark.parametrize("shape", SHAPES)
def test_fsm_shapes(cuda, oracle_lib, shape):
    rng = np.random.default_rng(hash(shape) % 2**32 + 1)
    dem = (rng.uniform(0, 400, size=shape)).astype(np.float32)
    dem[rng.uniform(size=shape) < 0.02] = np.nan
    slopes = rng.uniform(0.05, 0.9, size=shape).astype(np.float32)
    u64, _ = oracle_lib.fsm2d_f64(dem, slopes, 30.0)
    for tile in TILES:
        u, _ = gpu_fsm(dem, slopes, 30.0, tile)
        fsm_close(u, u64, f"shape {shape} tile {tile}")


def test_fsm_closed_forms(cuda):
    big = np.float32(1.0e6)
    dem = np.full((20, 130), big, np.float32)
    dem[0, :] = 0.0
    u, _ = gpu_fsm(dem, np.full(dem.shape, 0.5, np.float32), 30.0)
    np.testing.assert_array_equal(u, (15.0 * np.arange(20, dtype=np.float32))[:, None]
                                  * np.ones((1, 130), np.float32))
    dem = np.full((70, 70), big, np.float32)
    dem[0, 0] = 0.0
    u, _ = gpu_fsm(dem, np.full(dem.shape, 0.5, np.float32), 30.0)
    assert u[0, 3] == 45.0 and u[3, 0] == 45.0
    assert abs(float(u[1, 1]) - 15.0 * (1 + 1 / np.sqrt(2))) < 1e-5


def test_fsm_pits_1024(cuda, oracle_lib):
    from libtopotoolbox_amd import synthetic as S

    dem = S.dem_pits(1024, 1024, 3)
    slopes = S.voronoi_slopes(1024, 1024, 64, 0.2, 0.8, 33)
    u64, _ = oracle_lib.fsm2d_f64(dem, slopes, 30.0)
    u, st = gpu_fsm(dem, slopes, 30.0)
    err = fsm_close(u, u64, "config-3 generator at 1024^2")
    u32, _ = oracle_lib.fsm2d_f32(dem, slopes, 30.0)
    fsm_close(u, u32.astype(np.float64), "vs float restatement")
    print("max rel err", err, "stats", st)


# ------------------------------------------------------------------ full-size workloads
@pytest.mark.slow
def test_fillsinks_8192_depressions(cuda, oracle_lib):
    """8192^2 fractal (Hurst 0.7, 0-3000 m) with 200 paraboloid depressions
    (30% flat-bottomed) and a 2% NaN blob on an edge, NaN handled by the
    reference caller convention: bit-exact against the CPU fillsinks (the
    compiled reference where it is built, else the C restatement), plus
    fill >= dem and no 8-neighbour pits among unfixed cells."""
    from libtopotoolbox_amd import synthetic as S

    n = 8192
    dem = S.dem_depressions(n, n, 2)
    dem, bc, nanm = S.nan_to_bc(dem, S.edge_bc(n, n))
    out, st = gpu_fill(dem, bc)
    assert st["error"] == 0
    if oracle_lib.ref() is not None:
        want = oracle_lib.ref_fillsinks(dem, bc)
        what = "compiled reference"
    else:
        want, it, conv = oracle_lib.fillsinks(dem, bc)
        assert conv
        what = "C restatement"
    assert_bitwise(out, want, f"8192^2 vs {what}")
    assert bool((out >= dem).all())
    f = out  # NaN cells are bc drains at -INF under the convention
    pad = np.pad(f, 1, constant_values=np.inf)
    lowest = np.min([pad[1 + dy:1 + dy + n, 1 + dx:1 + dx + n]
                     for dy in (-1, 0, 1) for dx in (-1, 0, 1) if dy or dx], axis=0)
    free = bc == 0
    assert not bool((f[free] < lowest[free]).any()), "a pit remains"
    print(f"8192^2 fillsinks bit-exact vs {what}; stats {st}")


@pytest.mark.slow
def test_fsm_8192_voronoi(cuda, oracle_lib):
    """8192^2 fractal (Hurst 0.8, 0-1000 m, 5% pits), cellsize 30 m, slopes
    from 64 Voronoi cells U(0.2, 0.8): max-abs and max-rel error against the
    FP64 FSM oracle (relative denominator max(|u64|, 1 m)), bound 1e-4."""
    from libtopotoolbox_amd import synthetic as S

    n = 8192
    dem = S.dem_pits(n, n, 1)
    slopes = S.voronoi_slopes(n, n, 64, 0.2, 0.8, 33, device="cuda")
    u, st = gpu_fsm(dem, slopes, 30.0)
    assert st["error"] == 0
    u64, it = oracle_lib.fsm2d_f64(dem, slopes, 30.0)
    rel = fsm_close(u, u64, "8192^2")
    ok = np.isfinite(u64)
    max_abs = float(np.abs(u[ok] - u64[ok]).max())
    print(f"8192^2 excess topography vs FP64: max abs {max_abs:.3e} m, max rel {rel:.3e}, "
          f"oracle iterations {it}; stats {st}")
