import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")
    config.addinivalue_line("markers", "slow: full-size workload (tens of seconds)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test without a GPU: run with -m 'not gpu' here")
    return torch.device("cuda:0")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize])


def assert_bitwise(a, b, what=""):
    ba, bb = bits(a), bits(b)
    if not np.array_equal(ba, bb):
        bad = np.argwhere(ba != bb)
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} cells differ; first at {i}: {a[i]!r} vs {b[i]!r}")


# ------------------------------------------------------------------ fillsinks adjoint scale cases
def scatter_scale_cases(oracle_lib):
    """(name, dem, bc, g_out) cases for the fillsinks adjoint's fixed-point
    scale: each fails a fixed 2^-32 scale with a +-2^31 range.

    * g_out = 1/N on a non-power-of-two grid (1/N is not a multiple of
      2^-32: ~1e-4 relative rounding per term at a fixed scale);
    * one lake of ~1e6 cells with g_out ~ 1e6 (a sum of ~1e12 > 2^31);
    * g_out ~ 1e-30 (below 2^-32: lost at a fixed scale);
    * +INF, NaN, and +INF with -INF on the contributors of three spill
      cells, and +INF on a NaN DEM cell (spill -1: contributes nothing).
    """
    from libtopotoolbox_amd import synthetic as S

    def edge(shape):
        bc = np.zeros(shape, np.uint8)
        bc[0, :] = bc[-1, :] = 1
        bc[:, 0] = bc[:, -1] = 1
        return bc

    out = []
    dem = S.dem_pits(1000, 777, 11)
    out.append(("g = 1/N", dem, edge(dem.shape), np.full(dem.shape, 1.0 / dem.size, np.float32)))

    rng = np.random.default_rng(21)
    lake = (50 + 10 * rng.uniform(size=(1100, 900))).astype(np.float32)
    lake[0, :] = lake[-1, :] = lake[:, 0] = lake[:, -1] = 100.0
    out.append(("one lake, g ~ 1e6", lake, edge(lake.shape),
                (1e6 * rng.uniform(0.5, 1.5, size=lake.shape)).astype(np.float32)))

    dem = S.dem_pits(517, 389, 12)
    out.append(("g ~ 1e-30", dem, edge(dem.shape),
                (1e-30 * rng.standard_normal(dem.shape)).astype(np.float32)))

    dem = S.dem_pits(300, 257, 13)
    dem[5, 5] = np.nan
    bc = edge(dem.shape)
    g = rng.standard_normal(dem.shape).astype(np.float32)
    _, spill = oracle_lib.fillsinks_spill(dem, bc)
    s = spill.reshape(-1)
    targets, counts = np.unique(s[s >= 0], return_counts=True)
    many = targets[np.argsort(-counts)][:3]  # the three largest spill sets
    gf = g.reshape(-1)
    c0 = np.flatnonzero(s == many[0])
    gf[c0[0]], gf[c0[1]] = np.inf, -np.inf
    gf[np.flatnonzero(s == many[1])[0]] = np.inf
    gf[np.flatnonzero(s == many[2])[0]] = np.nan
    g[5, 5] = np.inf
    out.append(("NaN / INF", dem, bc, g))
    return out


def assert_scatter_close(got, want, g, what):
    """got (float32) against the FP64 scatter: the same NaN and INF cells,
    finite cells within 1e-6 relative (plus 1e-12 max|g| absolute)."""
    nan_g, nan_w = np.isnan(got), np.isnan(want)
    np.testing.assert_array_equal(nan_g, nan_w, err_msg=f"{what}: NaN cells")
    inf_w = np.isinf(want)
    np.testing.assert_array_equal(np.isinf(got), inf_w, err_msg=f"{what}: INF cells")
    np.testing.assert_array_equal(got[inf_w], want[inf_w], err_msg=f"{what}: INF signs")
    ok = ~nan_w & ~inf_w
    gmax = np.abs(g[np.isfinite(g)]).max()
    err = np.abs(got[ok] - want[ok]) - (1e-6 * np.abs(want[ok]) + 1e-12 * gmax)
    assert err.max() <= 0, (what, float(err.max()))
