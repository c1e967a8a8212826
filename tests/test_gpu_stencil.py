"""GPU parity of the stencils next to the hot path: identifyflats and the
min / max value filters (general structuring element and square), against
the oracle, bit for bit (both are selections of input values, so the sign
of zeros and the NaN pattern must match too).
"""
import numpy as np
import pytest
import torch

from conftest import assert_bitwise

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 37), (37, 1), (2, 3), (3, 3), (8, 64), (9, 65), (57, 83), (130, 513),
          (1, 4), (4, 4), (67, 260), (130, 1028)]  # nx % 4 == 0: the streaming kernel


def stencil_dem(rng, shape, flats=True):
    """Random DEM with plateaus, NaN, +-INF and -0.0 cells."""
    dem = rng.uniform(0, 50, size=shape).astype(np.float32)
    if flats:
        dem = np.round(dem / 5) * 5
    dem[rng.uniform(size=shape) < 0.03] = np.nan
    dem[rng.uniform(size=shape) < 0.01] = np.inf
    dem[rng.uniform(size=shape) < 0.01] = -np.inf
    zeros = rng.uniform(size=shape) < 0.05
    dem[zeros] = np.where(rng.uniform(size=zeros.sum()) < 0.5, -0.0, 0.0)
    return dem.astype(np.float32)


def cuda(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("shape", SHAPES)
def test_identifyflats(oracle_lib, shape):
    import libtopotoolbox_amd as tt

    rng = np.random.default_rng(hash(shape) % 2**32)
    for flats in (True, False):
        dem = stencil_dem(rng, shape, flats)
        want, wn = oracle_lib.identifyflats(dem)
        got, n = tt.identifyflats(cuda(dem))
        np.testing.assert_array_equal(got.cpu().numpy(), want)
        assert n == wn


def test_identifyflats_plateau_and_host_dropin(oracle_lib):
    import libtopotoolbox_amd as tt

    # one large plateau with a single outlet: flats everywhere inside, the
    # outlet's equal neighbours are sills, the flats touching them presills
    dem = np.full((300, 260), 10.0, np.float32)
    dem[0, :] = dem[-1, :] = dem[:, 0] = dem[:, -1] = 20.0
    dem[150, 0] = 10.0
    dem[40:60, 100:130] = 12.0
    want, wn = oracle_lib.identifyflats(dem)
    got, n = tt.identifyflats_host(dem)
    np.testing.assert_array_equal(got, want)
    assert n == wn
    got_d, n_d = tt.identifyflats(cuda(dem))
    np.testing.assert_array_equal(got_d.cpu().numpy(), want)
    assert n_d == wn
    flags = tt.identifyflats(cuda(dem), count=False)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(flags.cpu().numpy(), want)


SES = {
    "3x3": np.ones((3, 3), np.uint8),
    "cross": np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]], np.uint8),
    "5x3": (np.arange(15).reshape(3, 5) % 3 != 1).astype(np.uint8),
    "2x2": np.ones((2, 2), np.uint8),
    "1x1": np.ones((1, 1), np.uint8),
    "7x1": np.ones((1, 7), np.uint8),
    "1x6": np.ones((6, 1), np.uint8),
    "slices": np.stack([np.eye(3, dtype=np.uint8), np.ones((3, 3), np.uint8)[::-1],
                        np.array([[1, 0, 1], [0, 0, 0], [1, 0, 1]], np.uint8)]),
    "empty": np.zeros((3, 3), np.uint8),
    "9x9": np.ones((9, 9), np.uint8),
    # past the LDS path's limits (33 cells per axis, 1024 entries): the
    # direct kernel
    "35x1": np.ones((1, 35), np.uint8),
    "3x3x120": (np.arange(1080).reshape(120, 3, 3) % 7 != 3).astype(np.uint8),
}


@pytest.mark.parametrize("se_name", sorted(SES))
@pytest.mark.parametrize("op", ["min", "max"])
def test_value_filter(oracle_lib, se_name, op):
    import libtopotoolbox_amd as tt

    se = SES[se_name]
    fn = tt.min_filter if op == "min" else tt.max_filter
    rng = np.random.default_rng(17 + len(se_name))
    # nx % 4 == 0 shapes take the register-streaming 3x3 kernel for 3x3
    # elements (partial wavefronts and strips included)
    for shape in [(1, 1), (2, 3), (37, 1), (57, 83), (130, 257), (5, 4), (131, 300), (70, 1024)]:
        dem = stencil_dem(rng, shape)
        want = oracle_lib.value_filter(dem, se, op)
        got = fn(cuda(dem), cuda(se)).cpu().numpy()
        assert_bitwise(got, want, f"{op}_filter {se_name} {shape}")


@pytest.mark.parametrize("width", [0, 1, 2, 3, 4, 5, 8, 31, 255])
@pytest.mark.parametrize("op", ["min", "max"])
def test_value_filter_square(oracle_lib, width, op):
    import libtopotoolbox_amd as tt

    fn = tt.min_filter_square if op == "min" else tt.max_filter_square
    rng = np.random.default_rng(300 + width)
    # nx % 4 == 0 shapes take the streaming kernel for widths 1..5
    for shape in [(1, 1), (3, 2), (1, 300), (300, 1), (57, 83), (260, 129), (2, 4), (67, 260),
                  (130, 1028)]:
        dem = stencil_dem(rng, shape)
        want = oracle_lib.value_filter_square(dem, width, op)
        tmp = torch.empty(shape, dtype=torch.float32, device="cuda")
        got = fn(cuda(dem), width, tmp=tmp).cpu().numpy()
        assert_bitwise(got, want, f"{op}_filter_square w={width} {shape}")


def test_value_filter_host_dropins(oracle_lib):
    import libtopotoolbox_amd as tt

    rng = np.random.default_rng(5)
    dem = stencil_dem(rng, (70, 90))
    se = SES["slices"]
    assert_bitwise(tt.min_filter_host(dem, se), oracle_lib.value_filter(dem, se, "min"), "min")
    assert_bitwise(tt.max_filter_host(dem, se), oracle_lib.value_filter(dem, se, "max"), "max")
    for op, fn in (("min", tt.min_filter_square_host), ("max", tt.max_filter_square_host)):
        out, tmp = fn(dem, 5)
        assert_bitwise(out, oracle_lib.value_filter_square(dem, 5, op), op + " square")
        # tmp holds the pass along dims[0]: the square filter of width 1 x 5
        want_tmp = oracle_lib.value_filter(dem, np.ones((1, 5), np.uint8), op)
        finite = ~np.isnan(dem)
        assert_bitwise(tmp[finite], want_tmp[finite], op + " square tmp")


def test_stencils_large(oracle_lib):
    """4096 x 3000 (tail tiles in both directions), against the oracle."""
    import libtopotoolbox_amd as tt

    rng = np.random.default_rng(9)
    dem = stencil_dem(rng, (3000, 4100))
    d = cuda(dem)
    want, wn = oracle_lib.identifyflats(dem)
    got, n = tt.identifyflats(d)
    np.testing.assert_array_equal(got.cpu().numpy(), want)
    assert n == wn
    se = SES["cross"]
    assert_bitwise(tt.max_filter(d, cuda(se)).cpu().numpy(),
                   oracle_lib.value_filter(dem, se, "max"), "max cross")
    assert_bitwise(tt.min_filter_square(d, 7).cpu().numpy(),
                   oracle_lib.value_filter_square(dem, 7, "min"), "min square 7")


@pytest.mark.slow
def test_stencils_and_costs_8192(oracle_lib):
    """The 8192^2 fill workload (fractal with depressions and a -INF blob of
    bc cells): fill on the GPU, then identifyflats, min_filter 3x3,
    max_filter_square 5 and gwdt_computecosts on the filled DEM, each
    against the oracle on the same arrays (bitwise; NaN costs, from INF -
    INF depths in the blob, as NaN)."""
    import libtopotoolbox_amd as tt
    from libtopotoolbox_amd import synthetic as S

    n = 8192
    dem = S.dem_depressions(n, n, 2)
    dem, bc, _ = S.nan_to_bc(dem, S.edge_bc(n, n))
    d = cuda(dem)
    filled = tt.fillsinks(d, cuda(bc))
    flags, nflats = tt.identifyflats(filled)
    f_np = filled.cpu().numpy()
    want, wn = oracle_lib.identifyflats(f_np)
    np.testing.assert_array_equal(flags.cpu().numpy(), want)
    assert nflats == wn
    se = np.ones((3, 3), np.uint8)
    assert_bitwise(tt.min_filter(filled, cuda(se)).cpu().numpy(),
                   oracle_lib.value_filter(f_np, se, "min"), "min_filter 3x3")
    assert_bitwise(tt.max_filter_square(filled, 5).cpu().numpy(),
                   oracle_lib.value_filter_square(f_np, 5, "max"), "max_filter_square 5")
    costs, cc = tt.gwdt_computecosts(flags, d, filled)
    want_c, want_cc = oracle_lib.gwdt_computecosts(want, dem, f_np)
    np.testing.assert_array_equal(cc.cpu().numpy(), want_cc)
    got_c = costs.cpu().numpy()
    nan = np.isnan(want_c)
    np.testing.assert_array_equal(np.isnan(got_c), nan)
    assert_bitwise(got_c[~nan], want_c[~nan], "gwdt costs")
    print(f"8192^2: {nflats} flats; stencils and costs bit-exact vs the oracle")
