"""GPU adjoints against the FP64 oracle and FP64 central finite differences.

fillsinks: the spill index is compared exactly with the oracle's
lexicographic Priority-Flood; the backward against the FP64 scatter
(relative 1e-6; the fixed-point scale is chosen per call, pinned by
g_out = 1/N, a lake of ~1e6 cells with g_out ~ 1e6, g_out ~ 1e-30 and
NaN / INF entries); finite differences
(L = sum(g * fill), eps = 1e-3 applied in float32 and divided by the
actual float difference) are exact because outputs copy inputs.

excesstopography_fsm2d: the backward against the FP64 adjoint evaluated at
the GPU's own converged state (relative 1e-4), and against FP64 central
finite differences of the FP64 oracle (eps = 1e-3; relative 1e-3 + 1e-4
absolute; where a perturbation of 1e-3 crosses a branch switch of the
Godunov update, detected by disagreement with eps = 1e-4, the 1e-4
difference is used). At the 4096^2 size the finite differences run on a 96x96 crop
centred on each sampled cell: on the pit-dominated test terrain a cell's
influence stays within a few cells, so the crop sees the same local
problem; cells whose constrained/eikonal classification differs between
float32 and float64 are reported and excluded (non-differentiable there).
"""
import numpy as np
import pytest
import torch

from conftest import assert_bitwise, assert_scatter_close, load_golden, scatter_scale_cases

pytestmark = pytest.mark.gpu


def edge_bc(shape):
    bc = np.zeros(shape, np.uint8)
    bc[0, :] = bc[-1, :] = 1
    bc[:, 0] = bc[:, -1] = 1
    return bc


def to_gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


# ------------------------------------------------------------------ fill
def fill_cases(oracle_lib):
    out = []
    for name in ("fillsinks_random_s0", "fillsinks_nan_raw", "fillsinks_nan_convention",
                 "fillsinks_flats"):
        g = load_golden(name)
        out.append((name, g["dem"], g["bc"]))
    rng = np.random.default_rng(12)
    dem = oracle_lib.grid_random(3, (300, 257))
    dem[rng.uniform(size=dem.shape) < 0.02] = np.nan
    out.append(("random 300x257", dem, (rng.uniform(size=dem.shape) < 0.01).astype(np.uint8)
                | edge_bc(dem.shape)))
    return out


@pytest.mark.parametrize("tile", [(2, 1), (1, 1), (1, 2)])
def test_spill_index_matches_oracle(cuda, oracle_lib, tile):
    import libtopotoolbox_amd as tt
    from libtopotoolbox_amd import ops

    tt.set_tile_shape(*tile)
    try:
        for name, dem, bc in fill_cases(oracle_lib):
            want, wspill = oracle_lib.fillsinks_spill(dem, bc)
            (out, spill), st = ops.fillsinks_spill(to_gpu(dem), to_gpu(bc), stats=True)
            assert st["error"] == 0
            out = out.cpu().numpy()
            np.testing.assert_array_equal(out.view(np.uint32), want.view(np.uint32), err_msg=name)
            np.testing.assert_array_equal(spill.cpu().numpy().astype(np.int64), wspill,
                                          err_msg=name)
            # the plain forward gives the same levels
            plain = tt.fillsinks(to_gpu(dem), to_gpu(bc)).cpu().numpy()
            np.testing.assert_array_equal(plain.view(np.uint32), want.view(np.uint32))
    finally:
        tt.set_tile_shape(2, 2)


def test_spill_index_seeded_tile_boundaries(cuda, oracle_lib):
    """Regression: a 1681 x 936 fractal with 5% pits (seed 5096) on 128 x 64
    tiles; spill cells seeded on a tile's first visit must still queue the
    neighbours that staged the tile's unseeded values (three cells on tile
    boundary rows differed from the oracle when they did not)."""
    import libtopotoolbox_amd as tt
    from libtopotoolbox_amd import ops, synthetic as S

    dem = S.dem_pits(1681, 936, 5096)
    bc = edge_bc(dem.shape)
    want, wspill = oracle_lib.fillsinks_spill(dem, bc)
    try:
        for tile in [(2, 1), (1, 2), (1, 1)]:
            tt.set_tile_shape(*tile)
            (out, spill), st = ops.fillsinks_spill(to_gpu(dem), to_gpu(bc), stats=True)
            assert st["error"] == 0
            np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), want.view(np.uint32))
            np.testing.assert_array_equal(spill.cpu().numpy().astype(np.int64), wspill,
                                          err_msg=f"tile {tile}")
    finally:
        tt.set_tile_shape(2, 2)


def test_fill_backward_vs_fp64(cuda, oracle_lib):
    from libtopotoolbox_amd import ops

    rng = np.random.default_rng(3)
    for name, dem, bc in fill_cases(oracle_lib):
        g = rng.standard_normal(dem.shape).astype(np.float32)
        _, spill = oracle_lib.fillsinks_spill(dem, bc)
        want = oracle_lib.fillsinks_adjoint_f64(spill, g.astype(np.float64))
        got = ops.fillsinks_backward(to_gpu(spill.astype(np.int32)), to_gpu(g)).cpu().numpy()
        err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
        assert err.max() <= 1e-6, (name, err.max())
        # determinism: a second run gives the same bits
        again = ops.fillsinks_backward(to_gpu(spill.astype(np.int32)), to_gpu(g)).cpu().numpy()
        np.testing.assert_array_equal(again.view(np.uint32), got.view(np.uint32))


def test_fill_backward_scale_and_nonfinite(cuda, oracle_lib):
    from libtopotoolbox_amd import ops

    for name, dem, bc, g in scatter_scale_cases(oracle_lib):
        _, spill = oracle_lib.fillsinks_spill(dem, bc)
        want = oracle_lib.fillsinks_adjoint_f64(spill, g.astype(np.float64))
        sp = to_gpu(spill.astype(np.int32))
        got = ops.fillsinks_backward(sp, to_gpu(g)).cpu().numpy()
        assert_scatter_close(got, want, g, name)
        again = ops.fillsinks_backward(sp, to_gpu(g)).cpu().numpy()
        assert_bitwise(again, got, f"{name}: second run")


def fill_fd(dem_t, bc_t, g_t, cells, eps=1e-3):
    """FP64 central differences of L = sum(g * fillsinks(dem)) on the GPU,
    divided by the float32 difference actually applied."""
    import libtopotoolbox_amd as tt

    nx = dem_t.shape[1]
    out = []
    g64 = g_t.double()
    for p in cells:
        j, i = divmod(int(p), nx)
        x = float(dem_t[j, i])
        up = np.float32(x + eps)
        dn = np.float32(x - eps)
        d = dem_t.clone()
        d[j, i] = float(up)
        lu = (g64 * tt.fillsinks(d, bc_t).double()).sum().item()
        d[j, i] = float(dn)
        ld = (g64 * tt.fillsinks(d, bc_t).double()).sum().item()
        out.append((lu - ld) / (float(up) - float(dn)))
    return np.array(out)


def test_fill_autograd_fd_small(cuda, oracle_lib):
    from libtopotoolbox_amd.autograd import fillsinks as dfill

    rng = np.random.default_rng(4)
    dem = (rng.uniform(size=(70, 90)) * 300).astype(np.float32)
    bc = edge_bc(dem.shape)
    g = rng.standard_normal(dem.shape).astype(np.float32)
    d = to_gpu(dem).requires_grad_(True)
    out = dfill(d, to_gpu(bc))
    (out.double() * to_gpu(g).double()).sum().backward()
    grad = d.grad.cpu().numpy()
    cells = rng.choice(dem.size, size=64, replace=False)
    fd = fill_fd(to_gpu(dem), to_gpu(bc), to_gpu(g), cells)
    np.testing.assert_allclose(grad.reshape(-1)[cells], fd, rtol=1e-5, atol=1e-5)
    assert abs(grad.astype(np.float64).sum() - g.astype(np.float64).sum()) < 1e-3


# ------------------------------------------------------------------ excess topography
def fsm_case(rng, shape):
    dem = (rng.uniform(size=shape) * 300).astype(np.float32)
    slopes = rng.uniform(0.1, 0.9, size=shape).astype(np.float32)
    return dem, slopes


def test_fsm_backward_vs_fp64_adjoint(cuda, oracle_lib):
    from libtopotoolbox_amd import ops

    rng = np.random.default_rng(5)
    for shape in [(64, 64), (130, 257), (200, 100)]:
        dem, slopes = fsm_case(rng, shape)
        dem[rng.uniform(size=shape) < 0.02] = np.nan
        g = rng.standard_normal(shape).astype(np.float32)
        u = ops.excesstopography_fsm2d(to_gpu(dem), to_gpu(slopes), 30.0)
        (gd, gs), st = ops.excesstopography_fsm2d_backward(u, to_gpu(dem), to_gpu(slopes), 30.0,
                                                           to_gpu(g), stats=True)
        assert st["error"] == 0
        wd, ws = oracle_lib.fsm2d_adjoint_f64(u.cpu().numpy(), dem, slopes, 30.0, g)
        for got, want, what in ((gd.cpu().numpy(), wd, "dem"), (gs.cpu().numpy(), ws, "slope")):
            err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
            assert err.max() <= 1e-4, (shape, what, err.max())


def fsm_fd_oracle(oracle_lib, dem, slopes, g, j, i, eps=1e-3, cs=30.0):
    """FP64 central differences of L = sum(g * u) with respect to dem[j, i]
    and slopes[j, i]. Where the eps = 1e-3 difference straddles a branch
    switch of the Godunov update (it disagrees with eps / 10 by more than
    1e-3 relative), the eps / 10 value is used. Returns (fd_dem, fd_slope,
    number of kinks met)."""
    def L(d, s):
        return (g * oracle_lib.fsm2d_f64(d, s, cs)[0]).sum()

    d64 = dem.astype(np.float64)
    s64 = slopes.astype(np.float64)
    out, kinks = [], 0
    for wrt in ("dem", "slope"):
        vals = []
        for h in (eps, eps / 10):
            e = np.zeros(d64.shape)
            e[j, i] = h
            if wrt == "dem":
                vals.append((L(d64 + e, s64) - L(d64 - e, s64)) / (2 * h))
            else:
                vals.append((L(d64, s64 + e) - L(d64, s64 - e)) / (2 * h))
        if abs(vals[0] - vals[1]) > 1e-3 * abs(vals[1]) + 1e-4:
            kinks += 1
            out.append(vals[1])
        else:
            out.append(vals[0])
    return out[0], out[1], kinks


def test_fsm_autograd_fd_small(cuda, oracle_lib):
    from libtopotoolbox_amd.autograd import excesstopography_fsm2d as dfsm

    rng = np.random.default_rng(6)
    dem, slopes = fsm_case(rng, (48, 56))
    g = rng.standard_normal(dem.shape)
    d = to_gpu(dem).requires_grad_(True)
    s = to_gpu(slopes).requires_grad_(True)
    u = dfsm(d, s, 30.0)
    (u.double() * to_gpu(g)).sum().backward()
    gd, gs = d.grad.cpu().numpy(), s.grad.cpu().numpy()
    u32 = u.detach().cpu().numpy()
    u64, _ = oracle_lib.fsm2d_f64(dem, slopes, 30.0)
    checked = 0
    for p in rng.choice(dem.size, size=60, replace=False):
        j, i = divmod(int(p), dem.shape[1])
        if (u32[j, i] < dem[j, i]) != (u64[j, i] < dem[j, i]):
            continue
        fd_d, fd_s, _ = fsm_fd_oracle(oracle_lib, dem, slopes, g, j, i)
        assert abs(fd_d - gd[j, i]) <= 1e-3 * abs(fd_d) + 1e-4, (j, i, fd_d, gd[j, i])
        assert abs(fd_s - gs[j, i]) <= 1e-3 * abs(fd_s) + 1e-4, (j, i, fd_s, gs[j, i])
        checked += 1
    assert checked >= 50


# ------------------------------------------------------------------ 4096^2 adjoint workload
@pytest.mark.slow
def test_adjoint_4096(cuda, oracle_lib):
    """Fractal (Hurst 0.8, 0-1000 m, 5% pits) 4096^2 DEM, Voronoi threshold
    slopes U(0.2, 0.8), upstream gradient N(0, 1) (seed 44), 256 sampled
    cells (seed 444)."""
    from libtopotoolbox_amd import synthetic as S
    from libtopotoolbox_amd.autograd import excesstopography_fsm2d as dfsm
    from libtopotoolbox_amd.autograd import fillsinks as dfill

    n = 4096
    dem = S.dem_pits(n, n, 1)
    slopes = S.voronoi_slopes(n, n, 64, 0.2, 0.8, 33, device="cuda")
    g = np.random.default_rng(44).standard_normal((n, n)).astype(np.float32)
    cells = np.random.default_rng(444).choice(n * n, size=256, replace=False)
    bc = edge_bc(dem.shape)
    dem_t, bc_t, g_t = to_gpu(dem), to_gpu(bc), to_gpu(g)

    # fillsinks: autograd gradient vs FP64 oracle adjoint and GPU FD
    d = dem_t.clone().requires_grad_(True)
    (dfill(d, bc_t).double() * g_t.double()).sum().backward()
    grad = d.grad.cpu().numpy()
    _, spill = oracle_lib.fillsinks_spill(dem, bc)
    want = oracle_lib.fillsinks_adjoint_f64(spill, g.astype(np.float64))
    err = np.abs(grad - want) / np.maximum(1.0, np.abs(want))
    assert err.max() <= 1e-6, err.max()
    fd = fill_fd(dem_t, bc_t, g_t, cells)
    np.testing.assert_allclose(grad.reshape(-1)[cells], fd, rtol=1e-5, atol=1e-5)

    # excess topography: autograd vs FP64 adjoint at the GPU state, FD on crops
    d = dem_t.clone().requires_grad_(True)
    s = to_gpu(slopes).requires_grad_(True)
    u = dfsm(d, s, 30.0)
    (u.double() * g_t.double()).sum().backward()
    gd, gs = d.grad.cpu().numpy(), s.grad.cpu().numpy()
    u32 = u.detach().cpu().numpy()
    wd, ws = oracle_lib.fsm2d_adjoint_f64(u32, dem, slopes, 30.0, g)
    for got, w, what in ((gd, wd, "dem"), (gs, ws, "slope")):
        e = np.abs(got - w) / np.maximum(1.0, np.abs(w))
        assert e.max() <= 1e-4, (what, e.max())
    half = 48
    checked = skipped = kinks = 0
    for p in cells:
        j, i = divmod(int(p), n)
        j0, j1 = max(0, j - half), min(n, j + half)
        i0, i1 = max(0, i - half), min(n, i + half)
        cd, cs_, cg = dem[j0:j1, i0:i1], slopes[j0:j1, i0:i1], g[j0:j1, i0:i1].astype(np.float64)
        u64c, _ = oracle_lib.fsm2d_f64(cd, cs_, 30.0)
        if (u64c[j - j0, i - i0] < cd[j - j0, i - i0]) != (u32[j, i] < dem[j, i]):
            skipped += 1
            continue
        fd_d, fd_s, k = fsm_fd_oracle(oracle_lib, cd, cs_, cg, j - j0, i - i0)
        kinks += k
        assert abs(fd_d - gd[j, i]) <= 1e-3 * abs(fd_d) + 1e-4, (j, i, fd_d, gd[j, i])
        assert abs(fd_s - gs[j, i]) <= 1e-3 * abs(fd_s) + 1e-4, (j, i, fd_s, gs[j, i])
        checked += 1
    print(f"fsm FD: {checked} cells checked, {skipped} skipped (classification differs), "
          f"{kinks} derivatives taken at eps 1e-4 (branch switch within 1e-3)")
    assert checked >= 250


def spill_stress_cases():
    """Grids whose submerged components and flat plateaus cross many tiles:
    a fractal quantised to 5 m steps (large plateaus at their own level:
    flat cells, settled by the spill relaxation) and the depression
    generator at 1500 x 1700 (flat-bottomed bowls spanning tiles, a NaN blob
    under the reference caller convention)."""
    from libtopotoolbox_amd import synthetic as S

    z = S.fractal(700, 900, 2.4, 71, 300.0)
    z = (np.round(z / 5.0) * 5.0).astype(np.float32) + np.float32(0.0)
    out = [("quantised 700x900", z, edge_bc(z.shape))]
    d = S.dem_depressions(1500, 1700, 72)
    d, bc, _ = S.nan_to_bc(d, S.edge_bc(1500, 1700))
    out.append(("depressions 1500x1700", d, bc))
    return out


@pytest.mark.parametrize("tile", [(2, 2), (2, 1), (1, 1)])
def test_spill_labelling_stress(cuda, oracle_lib, tile):
    """Spill index bit-exact to the oracle's lexicographic Priority-Flood on
    grids where components and plateaus cross label and relaxation tiles."""
    import libtopotoolbox_amd as tt
    from libtopotoolbox_amd import ops

    tt.set_tile_shape(*tile)
    try:
        for name, dem, bc in spill_stress_cases():
            want, wspill = oracle_lib.fillsinks_spill(dem, bc)
            (out, spill), st = ops.fillsinks_spill(to_gpu(dem), to_gpu(bc), stats=True)
            assert st["error"] == 0
            np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), want.view(np.uint32),
                                          err_msg=name)
            np.testing.assert_array_equal(spill.cpu().numpy().astype(np.int64), wspill,
                                          err_msg=f"{name} tile {tile}")
    finally:
        tt.set_tile_shape(2, 2)


def test_spill_labelling_stress_batch(cuda, oracle_lib):
    """The same grids as items of one batch (tile-by-tile labelling, the
    relaxation settling components that cross tile edges)."""
    from libtopotoolbox_amd.batched import Batch

    cases = spill_stress_cases()
    shapes = [d.shape for _, d, _ in cases]
    rng = np.random.default_rng(73)
    gs = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    b = Batch("fillsinks", shapes, tile=(1, 2))
    b.load([d for _, d, _ in cases], bc=[bc for _, _, bc in cases], g_out=gs)
    assert b.run(stats=True)["error"] == 0
    r = b.outputs()
    for k, (name, dem, bc) in enumerate(cases):
        want, wspill = oracle_lib.fillsinks_spill(dem, bc)
        np.testing.assert_array_equal(r["output"][k].cpu().numpy().view(np.uint32),
                                      want.view(np.uint32), err_msg=name)
        np.testing.assert_array_equal(r["spill"][k].cpu().numpy().astype(np.int64), wspill,
                                      err_msg=name)
        w = oracle_lib.fillsinks_adjoint_f64(wspill, gs[k].astype(np.float64))
        assert_scatter_close(r["g_dem"][k].cpu().numpy(), w, gs[k], name)


@pytest.mark.slow
def test_spill_8192_depressions(cuda, oracle_lib):
    """Spill index at 8192^2 on the fill workload's DEM (about 1.2M submerged
    components, 170K flat cells, lakes across many tiles): bit-exact to the
    oracle's lexicographic Priority-Flood."""
    from libtopotoolbox_amd import ops, synthetic as S

    n = 8192
    dem = S.dem_depressions(n, n, 2)
    dem, bc, _ = S.nan_to_bc(dem, S.edge_bc(n, n))
    (out, spill), st = ops.fillsinks_spill(to_gpu(dem), to_gpu(bc), stats=True)
    assert st["error"] == 0
    out = out.cpu().numpy()
    spill = spill.cpu().numpy()
    want, wspill = oracle_lib.fillsinks_spill(dem, bc)
    assert_bitwise(out, want, "8192^2 levels")
    bad = np.flatnonzero(spill.reshape(-1).astype(np.int64) != wspill.reshape(-1))
    assert bad.size == 0, f"{bad.size} spill indices differ, first at {bad[:5]}"
    print(f"8192^2 spill index bit-exact; relaxation stats {st}")
