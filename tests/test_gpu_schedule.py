"""GPU parity of both schedules of single-grid sink filling and
reconstruction: the lowest-level-first queue and the FIFO queue each reach
the oracle's fixed point bit for bit, on every tile shape, whatever the
automatic choice would have been (tt_set_fill_schedule forces one)."""
import numpy as np
import pytest
import torch

from conftest import assert_bitwise, load_golden

pytestmark = pytest.mark.gpu

TILES = [(2, 2), (2, 1), (1, 2), (1, 1)]


@pytest.fixture(params=["priority", "fifo"])
def schedule(request):
    import libtopotoolbox_amd as tt

    tt.set_fill_schedule(request.param)
    yield request.param
    tt.set_fill_schedule("auto")


def edge_bc(shape):
    bc = np.zeros(shape, np.uint8)
    bc[0, :] = bc[-1, :] = 1
    bc[:, 0] = bc[:, -1] = 1
    return bc


def fill(dem, bc, tile):
    import libtopotoolbox_amd as tt

    tt.set_tile_shape(*tile)
    d = torch.from_numpy(np.ascontiguousarray(dem)).cuda()
    b = torch.from_numpy(np.ascontiguousarray(bc)).cuda()
    out, st = tt.fillsinks(d, b, stats=True)
    assert st["error"] == 0, st
    return out.cpu().numpy(), st


@pytest.mark.parametrize("name", ["fillsinks_random_s0", "fillsinks_nan_raw",
                                  "fillsinks_nan_convention", "fillsinks_flats"])
def test_golden(cuda, schedule, name):
    g = load_golden(name)
    for tile in TILES:
        out, _ = fill(g["dem"], g["bc"], tile)
        assert_bitwise(out, g["out"], f"{name} {schedule} tile {tile}")


@pytest.mark.parametrize("shape", [(1, 1), (37, 1), (65, 129), (257, 300), (700, 900)])
def test_shapes(cuda, oracle_lib, schedule, shape):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    dem = oracle_lib.grid_random(shape[0] * 1000 + shape[1], shape)
    dem[rng.uniform(size=shape) < 0.02] = np.nan
    bc = edge_bc(shape)
    want = oracle_lib.fillsinks_pf(dem, bc)
    for tile in TILES:
        out, _ = fill(dem, bc, tile)
        assert_bitwise(out, want, f"shape {shape} {schedule} tile {tile}")


def test_degenerate_levels(cuda, oracle_lib, schedule):
    """Level ranges the bucket map must survive: one level everywhere (an
    empty range), huge finite magnitudes (a range that overflows), +-INF and
    NaN cells, no seed at all, seeds only in the middle, every cell a seed."""
    rng = np.random.default_rng(11)
    shape = (300, 410)
    cases = {
        "constant": np.full(shape, 5.0, np.float32),
        "huge": rng.choice(np.float32([-3e38, 3e38, 0.0, 1e30]), size=shape).astype(np.float32),
        "specials": np.where(rng.uniform(size=shape) < 0.1,
                             rng.choice(np.float32([np.inf, -np.inf, np.nan]), size=shape),
                             rng.uniform(0, 50, size=shape)).astype(np.float32),
    }
    middle = np.zeros(shape, np.uint8)
    middle[150, 200] = middle[20:25, 300] = 1
    bcs = {"edge": edge_bc(shape), "none": np.zeros(shape, np.uint8), "middle": middle,
           "all": np.ones(shape, np.uint8)}
    for dname, dem in cases.items():
        for bname, bc in bcs.items():
            want, _, conv = oracle_lib.fillsinks(dem, bc)
            assert conv
            for tile in ((2, 2), (1, 1)):
                out, _ = fill(dem, bc, tile)
                assert_bitwise(out, want, f"{dname} / bc {bname} {schedule} tile {tile}")


def test_depressions_1500(cuda, oracle_lib, schedule):
    """Many tiles per workgroup's worth of queue traffic: 1500 x 1700 with
    depressions, flat bottoms and a NaN blob (NaN drains to the edge)."""
    from libtopotoolbox_amd import synthetic as S

    dem = S.dem_depressions(1500, 1700, 72)
    dem, bc, _ = S.nan_to_bc(dem, S.edge_bc(1500, 1700))
    want = oracle_lib.fillsinks_pf(dem, bc)
    for tile in TILES:
        out, st = fill(dem, bc, tile)
        assert_bitwise(out, want, f"depressions 1500x1700 {schedule} tile {tile}")


def test_more_tiles_than_one_bitmap_stretch(cuda, oracle_lib, schedule):
    """4160 x 4170 cells in 64 x 64 tiles: 4290 tiles, more than the 4096
    one stretch of 64 bitmap words holds, so pops rank and pick across
    several stretches of a bucket (the 16384^2 fill takes that path with
    128 x 128 tiles)."""
    from libtopotoolbox_amd import synthetic as S

    dem = S.dem_depressions(4160, 4170, 31)
    dem, bc, _ = S.nan_to_bc(dem, S.edge_bc(4160, 4170))
    want = oracle_lib.fillsinks_pf(dem, bc)
    out, st = fill(dem, bc, (1, 1))
    assert st["tile_x"] == 64 and st["tile_y"] == 64
    assert_bitwise(out, want, f"4160x4170 {schedule} tile (1, 1)")


@pytest.mark.parametrize("kind", ["sparse", "dense"])
def test_reconstruct_schedules(cuda, oracle_lib, schedule, kind):
    """Reconstruction by dilation on both queues and every tile shape,
    against the oracle's reconstruct: a sparse marker (the mask on the
    edge, -INF elsewhere: fillsinks posed as a reconstruction) and a dense
    one (mask - 5 everywhere, with raised cells), over a mask with NaN
    holes."""
    import libtopotoolbox_amd as tt

    rng = np.random.default_rng(41)
    shape = (700, 650)
    mask = oracle_lib.grid_random(88, shape)
    mask[rng.uniform(size=shape) < 0.02] = np.nan
    if kind == "sparse":
        marker = np.where(edge_bc(shape) == 1, mask, -np.inf).astype(np.float32)
    else:
        marker = (mask - 5.0).astype(np.float32)
        raised = rng.uniform(size=shape) < 0.01
        marker[raised] = mask[raised] + 3.0
    want, _, conv = oracle_lib.reconstruct(marker, mask)
    assert conv
    m = torch.from_numpy(marker).cuda()
    k = torch.from_numpy(mask).cuda()
    for tile in TILES:
        tt.set_tile_shape(*tile)
        out, st = tt.reconstruct(m, k, stats=True)
        assert st["error"] == 0, st
        assert_bitwise(out.cpu().numpy(), want, f"reconstruct {kind} {schedule} tile {tile}")


def test_random_configurations(cuda, oracle_lib, schedule):
    """Twelve seeded random configurations: sides in [60, 1500], every tile
    shape, edge / scattered / interior-only boundary cells, NaN holes, -INF
    and +INF cells, quantised levels (plateaus), against the oracle's
    restatement of the reference fillsinks."""
    rng = np.random.default_rng(2027)
    for k in range(12):
        ny, nx = (int(v) for v in rng.integers(60, 1500, size=2))
        dem = oracle_lib.grid_random(500 + k, (ny, nx))
        if k % 3 == 0:
            dem = np.floor(dem * 4.0) / 4.0  # plateaus
        u = rng.uniform(size=dem.shape)
        dem[u < 0.01] = np.nan
        dem[(u >= 0.01) & (u < 0.012)] = -np.inf
        dem[(u >= 0.012) & (u < 0.014)] = np.inf
        dem = dem.astype(np.float32)
        kind = k % 3
        if kind == 0:
            bc = edge_bc(dem.shape)
        elif kind == 1:
            bc = (rng.uniform(size=dem.shape) < 0.002).astype(np.uint8)
        else:
            bc = np.zeros(dem.shape, np.uint8)
            bc[ny // 3:ny // 3 + 3, nx // 2] = 1
        want, _, conv = oracle_lib.fillsinks(dem, bc)
        assert conv
        tile = TILES[k % len(TILES)]
        out, _ = fill(dem, bc, tile)
        assert_bitwise(out, want, f"config {k} ({ny}x{nx}, bc kind {kind}) {schedule} tile {tile}")

