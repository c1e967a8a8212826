"""GPU parity of gwdt_computecosts against the oracle: costs bit for bit
and conncomps equal (the same deepest pixel as the reference's scan picks,
including components whose deepest depth is shared by many pixels)."""
import numpy as np
import pytest
import torch

from conftest import assert_bitwise

pytestmark = pytest.mark.gpu


def gwdt_inputs(rng, shape, levels=4, zero_depth=False):
    flats = (rng.uniform(size=shape) < rng.uniform(0.3, 0.95)).astype(np.int32)
    flats[0, :] = flats[-1, :] = 0
    flats[:, 0] = flats[:, -1] = 0
    original = rng.uniform(0, 10, size=shape).astype(np.float32)
    depth = rng.integers(0, levels, size=shape).astype(np.float32)
    if zero_depth:
        depth[:] = 0
    return flats, original, (original + depth).astype(np.float32)


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def gpu(flats, orig, filled):
    import libtopotoolbox_amd as tt

    costs, cc = tt.gwdt_computecosts(cuda(flats), cuda(orig), cuda(filled))
    return costs.cpu().numpy(), cc.cpu().numpy()


def assert_costs(got, want, what):
    """Bitwise, except that NaN costs (from NaN depths) only have to be NaN:
    their payload comes from the hardware's NaN arithmetic."""
    nan = np.isnan(want)
    np.testing.assert_array_equal(np.isnan(got), nan, err_msg=what + " NaN costs")
    assert_bitwise(got[~nan], want[~nan], what + " costs")


def check(flats, orig, filled, oracle_lib, what):
    want_c, want_cc = oracle_lib.gwdt_computecosts(flats, orig, filled)
    got_c, got_cc = gpu(flats, orig, filled)
    np.testing.assert_array_equal(got_cc, want_cc, err_msg=what + " conncomps")
    assert_costs(got_c, want_c, what)


@pytest.mark.parametrize("seed", range(6))
def test_random_ties(oracle_lib, seed):
    rng = np.random.default_rng(900 + seed)
    for trial in range(40):
        shape = tuple(int(v) for v in rng.integers(3, 70, size=2))
        f, o, fl = gwdt_inputs(rng, shape, int(rng.integers(1, 5)), trial % 4 == 0)
        check(f, o, fl, oracle_lib, f"seed {seed} trial {trial} {shape}")


def test_empty_and_no_flats(oracle_lib):
    import libtopotoolbox_amd as tt

    for shape in [(1, 1), (2, 5), (6, 1)]:
        f = np.zeros(shape, np.int32)
        o = np.zeros(shape, np.float32)
        check(f, o, o, oracle_lib, f"no flats {shape}")
    c, cc = tt.gwdt_computecosts(torch.zeros((0, 4), dtype=torch.int32, device="cuda"),
                                 torch.zeros((0, 4), device="cuda"),
                                 torch.zeros((0, 4), device="cuda"))
    assert c.numel() == 0 and cc.numel() == 0


def test_filled_dem_pipeline(oracle_lib):
    """fillsinks -> identifyflats -> gwdt_computecosts on the GPU, against
    the oracle on the same filled DEM and flats, on quantised DEMs (natural
    flats of equal depth 0 and filled basins)."""
    import libtopotoolbox_amd as tt

    rng = np.random.default_rng(44)
    for k, shape in enumerate([(64, 64), (200, 300), (513, 257)]):
        dem = np.round(rng.uniform(0, 6, size=shape)).astype(np.float32)
        if k == 1:
            dem = np.round(np.cumsum(rng.uniform(-1, 1, size=shape), axis=1) / 2).astype(np.float32)
        bc = np.zeros(shape, np.uint8)
        bc[0, :] = bc[-1, :] = 1
        bc[:, 0] = bc[:, -1] = 1
        d = cuda(dem)
        filled = tt.fillsinks(d, cuda(bc))
        flats, _ = tt.identifyflats(filled)
        costs, cc = tt.gwdt_computecosts(flats, d, filled)
        f_np, fl_np = flats.cpu().numpy(), filled.cpu().numpy()
        want_c, want_cc = oracle_lib.gwdt_computecosts(f_np, dem, fl_np)
        np.testing.assert_array_equal(cc.cpu().numpy(), want_cc, err_msg=str(shape))
        assert_bitwise(costs.cpu().numpy(), want_c, str(shape))


def test_large_plateaus_and_basins(oracle_lib):
    """1500 x 2100: one large natural plateau (every pixel of it a deepest
    one, so the label comes from the replayed scan) next to basins with a
    unique deepest pixel and basins with tied depths."""
    rng = np.random.default_rng(7)
    ny, nx = 1500, 2100
    flats = np.zeros((ny, nx), np.int32)
    y, x = np.mgrid[0:ny, 0:nx]
    plateau = (np.hypot(y - 700, x - 600) < 500 + 80 * np.sin(x / 17.0) * np.cos(y / 23.0))
    flats[plateau] = 1
    holes = rng.uniform(size=(ny, nx)) < 0.08
    flats[holes] = 0
    basins = (np.hypot(y - 400, x - 1600) < 300) | (np.hypot(y - 1200, x - 1500) < 250)
    flats[basins] = 1
    flats[0, :] = flats[-1, :] = 0
    flats[:, 0] = flats[:, -1] = 0
    orig = rng.uniform(0, 100, size=(ny, nx)).astype(np.float32)
    depth = np.zeros((ny, nx), np.float32)
    depth[basins] = np.round(rng.uniform(0, 3, size=basins.sum())).astype(np.float32)
    filled = (orig + depth).astype(np.float32)
    check(flats, orig, filled, oracle_lib, "large")


@pytest.mark.parametrize("mode", ["scattered", "blocks"])
def test_nan_depths(oracle_lib, mode):
    """Flats of +-INF cells have NaN depth (INF - INF): components made of
    them only, and components mixing them with numbers, as the reference's
    comparisons treat them."""
    rng = np.random.default_rng(12 if mode == "scattered" else 13)
    for trial in range(30):
        shape = tuple(int(v) for v in rng.integers(3, 60, size=2))
        f, o, fl = gwdt_inputs(rng, shape, int(rng.integers(1, 4)), trial % 4 == 0)
        if mode == "scattered":
            m = rng.uniform(size=shape) < 0.15
        else:
            y, x = np.mgrid[0:shape[0], 0:shape[1]]
            m = ((x // 7 + y // 5) % 3) == 0
        o[m] = fl[m] = np.inf if trial % 2 else -np.inf
        check(f, o, fl, oracle_lib, f"{mode} trial {trial} {shape}")


def test_host_dropin(oracle_lib):
    import libtopotoolbox_amd as tt

    rng = np.random.default_rng(3)
    f, o, fl = gwdt_inputs(rng, (50, 60), 2)
    want_c, want_cc = oracle_lib.gwdt_computecosts(f, o, fl)
    got_c, got_cc = tt.gwdt_computecosts_host(f, o, fl)
    np.testing.assert_array_equal(got_cc, want_cc)
    assert_costs(got_c, want_c, "host drop-in")


@pytest.mark.parametrize("pattern", ["diag", "anti", "checker", "hstripes", "vstripes", "stairs"])
def test_tile_border_patterns(oracle_lib, pattern):
    """Flats that cross the 64 x 16 labelling tiles in every direction
    (single-pixel diagonal chains, checkerboards, stripes, staircases), on a
    grid that is no multiple of the tile, with tied and distinct depths."""
    rng = np.random.default_rng(hash(pattern) % 1000)
    ny, nx = 211, 301
    y, x = np.mgrid[0:ny, 0:nx]
    flats = {
        "diag": (x - y) % 9 == 0,
        "anti": (x + y) % 7 == 0,
        "checker": (x + y) % 2 == 0,
        "hstripes": (y % 17) < 15,
        "vstripes": (x % 65) < 62,
        "stairs": ((x // 5 + y // 3) % 4) != 0,
    }[pattern].astype(np.int32)
    flats[0, :] = flats[-1, :] = 0
    flats[:, 0] = flats[:, -1] = 0
    orig = rng.uniform(0, 10, size=(ny, nx)).astype(np.float32)
    for levels in (1, 3, 50):
        depth = rng.integers(0, levels, size=(ny, nx)).astype(np.float32)
        check(flats, orig, (orig + depth).astype(np.float32), oracle_lib,
              f"{pattern} levels {levels}")


def test_many_tile_ties_and_nan_flats(oracle_lib):
    """Components spanning many tiles whose deepest depth is shared (two
    depth levels), next to a large all-NaN flat (INF - INF depths) and
    components mixing NaN and numbers: the tie replay over many tiles."""
    rng = np.random.default_rng(77)
    ny, nx = 700, 900
    flats = (rng.uniform(size=(ny, nx)) < 0.9).astype(np.int32)
    flats[0, :] = flats[-1, :] = 0
    flats[:, 0] = flats[:, -1] = 0
    orig = rng.uniform(0, 10, size=(ny, nx)).astype(np.float32)
    filled = (orig + rng.integers(0, 2, size=(ny, nx))).astype(np.float32)
    y, x = np.mgrid[0:ny, 0:nx]
    blob = np.hypot(y - 350, x - 600) < 180
    orig[blob] = filled[blob] = np.inf
    mixed = (np.hypot(y - 150, x - 150) < 60) & (rng.uniform(size=(ny, nx)) < 0.3)
    orig[mixed] = filled[mixed] = -np.inf
    check(flats, orig, filled, oracle_lib, "many-tile ties")
