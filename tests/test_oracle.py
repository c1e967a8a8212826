"""Pins the CPU oracle before it is trusted as the checker.

Which check pins which oracle function:

* oracle_pcg4d                 known answers from the reference's pcg4d
                               (test/utils.c:6-28), fixture pcg4d_kat.npz
* oracle_fillsinks /
  oracle_reconstruct           golden fixtures written by the reference
                               library; live comparison with oracle/_ref
                               where it is built; a numpy brute-force fixed
                               point on tiny grids; closed forms (single pit,
                               notched bowl, nested bowls); invariants
* oracle_fillsinks_spill / _pf bitwise against the restatement; spill
                               cells carry the output value; lexicographic
                               brute force on tiny grids
* oracle_fillsinks_adjoint_f64 exact finite differences (outputs are copies
                               of inputs) and conservation of the sum
* oracle_fsm2d_f32             golden fixtures (bitwise) and oracle/_ref
* oracle_fsm2d_f64             reference FMM fixtures within 1e-4 (the
                               reference's own FSM==FMM test), closed forms
                               (1-D ramp, plane, single source diagonal)
* oracle_fmm2d_f64             equals FSM f64 within 1e-4; reference FMM
* oracle_fsm2d_adjoint_f64     FP64 central finite differences
"""
import numpy as np
import pytest

from conftest import assert_bitwise, load_golden


# ------------------------------------------------------------------ helpers
def brute_fill(dem, bc):
    """Jacobi iteration of f = max(d, min(f, 8 neighbours)) from f0 = dem on
    bc==1 cells and +INF elsewhere; NaN cells are impassable."""
    nanm = np.isnan(dem)
    d = np.where(nanm, np.inf, dem).astype(np.float64)
    f = np.where((bc == 1) & ~nanm, d, np.inf)
    ny, nx = dem.shape
    while True:
        pad = np.full((ny + 2, nx + 2), np.inf)
        pad[1:-1, 1:-1] = f
        m = f.copy()
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                m = np.minimum(m, pad[1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx])
        g = np.where(nanm, np.inf, np.maximum(d, m))
        if np.array_equal(g, f):
            break
        f = g
    out = f.astype(np.float32)
    out[nanm] = dem[nanm]
    return out


def edge_bc(shape):
    bc = np.zeros(shape, np.uint8)
    bc[0, :] = bc[-1, :] = 1
    bc[:, 0] = bc[:, -1] = 1
    return bc


# ------------------------------------------------------------------ pcg4d
def test_pcg4d_known_answers(oracle_lib):
    kat = load_golden("pcg4d_kat")
    for args, val in zip(kat["args"], kat["values"]):
        got = np.float32(oracle_lib.pcg4d(*[int(a) for a in args]))
        assert got == val, (args, got, val)


# ------------------------------------------------------------------ fillsinks
@pytest.mark.parametrize("name", ["fillsinks_random_s0", "fillsinks_random_s1",
                                  "fillsinks_random_s2", "fillsinks_nan_raw",
                                  "fillsinks_nan_convention", "fillsinks_flats"])
def test_fillsinks_golden(oracle_lib, name):
    g = load_golden(name)
    out, it, conv = oracle_lib.fillsinks(g["dem"], g["bc"])
    assert conv and it < 1000
    assert_bitwise(out, g["out"], name)
    assert_bitwise(oracle_lib.fillsinks_pf(g["dem"], g["bc"]), g["out"], name + " (pf)")


def test_reconstruct_golden(oracle_lib):
    g = load_golden("reconstruct_random")
    out, it, conv = oracle_lib.reconstruct(g["marker"], g["mask"])
    assert conv
    assert_bitwise(out, g["out"], "reconstruct")


def test_fillsinks_vs_reference_library(oracle_lib):
    if oracle_lib.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(3)
    for seed in range(6):
        ny, nx = rng.integers(1, 90, size=2)
        dem = oracle_lib.grid_random(seed, (int(ny), int(nx)))
        dem[rng.uniform(size=dem.shape) < 0.03] = np.nan
        bc = (rng.uniform(size=dem.shape) < 0.02).astype(np.uint8)
        bc |= edge_bc(dem.shape)
        out, _, conv = oracle_lib.fillsinks(dem, bc)
        assert conv
        assert_bitwise(out, oracle_lib.ref_fillsinks(dem, bc), f"seed {seed}")
        marker = np.where(rng.uniform(size=dem.shape) < 0.2, dem + 3.0, -np.inf)
        marker = marker.astype(np.float32)
        r, _, _ = oracle_lib.reconstruct(marker, dem)
        assert_bitwise(r, oracle_lib.ref_reconstruct(marker, dem), f"recon {seed}")


@pytest.mark.parametrize("seed", range(8))
def test_fillsinks_brute_force(oracle_lib, seed):
    rng = np.random.default_rng(seed)
    ny, nx = rng.integers(1, 12, size=2)
    dem = rng.integers(0, 6, size=(ny, nx)).astype(np.float32)  # many ties
    if seed % 2:
        dem[rng.uniform(size=dem.shape) < 0.15] = np.nan
    bc = (rng.uniform(size=dem.shape) < 0.25).astype(np.uint8)
    want = brute_fill(dem, bc)
    out, _, conv = oracle_lib.fillsinks(dem, bc)
    assert conv
    assert_bitwise(out, want, "restatement")
    assert_bitwise(oracle_lib.fillsinks_pf(dem, bc), want, "priority-flood")


def test_fillsinks_closed_forms(oracle_lib):
    # single pit in a plane fills to the plane
    dem = np.full((9, 9), 10.0, np.float32)
    dem[4, 4] = 5.0
    out, _, _ = oracle_lib.fillsinks(dem, edge_bc(dem.shape))
    assert np.all(out == 10.0)
    # cone-shaped bowl with a notch in the rim fills to the notch height
    n = 41
    y, x = np.mgrid[0:n, 0:n]
    r = np.hypot(x - 20, y - 20)
    dem = np.where(r < 15, r, 30.0 - r).astype(np.float32)
    dem[20, 5:] = np.minimum(dem[20, 5:], 12.5)  # notch along a row to the edge
    out, _, _ = oracle_lib.fillsinks(dem, edge_bc(dem.shape))
    inside = r < 12
    assert np.all(out[inside] == np.float32(12.5))
    # nested square bowls (Chebyshev rings, no diagonal leaks): inner pit,
    # inner rim 10, middle basin 5, outer rim 20 notched to 15 on one side
    rc = np.maximum(np.abs(x - 20), np.abs(y - 20))
    dem2 = np.select([rc < 4, rc < 6, rc < 12, rc < 14], [1.0, 10.0, 5.0, 20.0], 0.0)
    dem2 = dem2.astype(np.float32)
    dem2[20, 32:34] = 15.0
    out2, _, _ = oracle_lib.fillsinks(dem2, edge_bc(dem2.shape))
    assert np.all(out2[rc < 12] == 15.0)
    assert np.all(out2[rc >= 14] == 0.0)


def test_fillsinks_invariants(oracle_lib):
    dem = oracle_lib.grid_random(99, (64, 80))
    bc = edge_bc(dem.shape)
    out, _, _ = oracle_lib.fillsinks(dem, bc)
    assert np.all(out >= dem)
    assert_bitwise(out[bc == 1], dem[bc == 1], "bc cells")
    again, _, _ = oracle_lib.fillsinks(out, bc)
    assert_bitwise(again, out, "idempotence")
    # no interior pit: every interior cell has a neighbour not strictly higher
    pad = np.pad(out, 1, constant_values=np.inf)
    higher = np.ones_like(out, bool)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dy or dx:
                higher &= pad[1 + dy:1 + dy + out.shape[0], 1 + dx:1 + dx + out.shape[1]] > out
    assert not higher[1:-1, 1:-1].any()
    # every value is a copy of some input value
    assert np.isin(out, dem).all()


def test_reconstruct_cap_and_resume(oracle_lib):
    dem = oracle_lib.grid_random(5, (60, 60))
    bc = edge_bc(dem.shape)
    full, it, conv = oracle_lib.fillsinks(dem, bc)
    assert conv and it > 1
    part, it1, conv1 = oracle_lib.fillsinks(dem, bc, max_iterations=1)
    assert it1 == 1 and not conv1
    assert not np.array_equal(part, full)
    # the partial state is an upper bound; resuming from it (reconstruct on
    # the negated arrays) reaches the same fixed point
    marker = -part
    res, _, conv2 = oracle_lib.reconstruct(marker, -dem)
    assert conv2
    assert_bitwise(-res, full, "resume")


# ------------------------------------------------------------------ spill index
def brute_spill(dem, bc):
    """Lexicographic (value, index) fixed point on a tiny grid."""
    ny, nx = dem.shape
    n = ny * nx
    nanm = np.isnan(dem).reshape(-1)
    d = dem.reshape(-1).astype(np.float64)
    key = [(np.inf, n) for _ in range(n)]
    for p in range(n):
        if bc.reshape(-1)[p] == 1 and not nanm[p]:
            key[p] = (d[p], p)
    changed = True
    while changed:
        changed = False
        for p in range(n):
            if nanm[p] or bc.reshape(-1)[p] == 1:
                continue
            j, i = divmod(p, nx)
            best = key[p]
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    qj, qi = j + dj, i + di
                    if (dj or di) and 0 <= qj < ny and 0 <= qi < nx:
                        q = qj * nx + qi
                        if not nanm[q] and key[q] < best:
                            best = key[q]
            new = max((d[p], p), best) if best[0] < np.inf else best
            if new < key[p]:
                key[p] = new
                changed = True
    return np.array([k[1] if k[0] < np.inf else -1 for k in key]).reshape(ny, nx)


@pytest.mark.parametrize("seed", range(6))
def test_spill_index(oracle_lib, seed):
    rng = np.random.default_rng(100 + seed)
    ny, nx = rng.integers(2, 10, size=2)
    dem = rng.integers(0, 5, size=(ny, nx)).astype(np.float32)
    if seed % 3 == 0:
        dem[rng.uniform(size=dem.shape) < 0.15] = np.nan
    bc = edge_bc(dem.shape) if seed % 2 else (rng.uniform(size=dem.shape) < 0.3).astype(np.uint8)
    out, spill = oracle_lib.fillsinks_spill(dem, bc)
    ok = spill >= 0
    assert_bitwise(out[ok], dem.reshape(-1)[spill[ok]], "spill carries the value")
    np.testing.assert_array_equal(spill, brute_spill(dem, bc))


def test_fill_adjoint_exact_fd(oracle_lib):
    rng = np.random.default_rng(8)
    dem = (rng.uniform(size=(12, 14)) * 100).astype(np.float32)
    bc = edge_bc(dem.shape)
    g_out = rng.standard_normal(dem.shape)
    _, spill = oracle_lib.fillsinks_spill(dem, bc)
    g_dem = oracle_lib.fillsinks_adjoint_f64(spill, g_out)
    assert abs(g_dem.sum() - g_out.sum()) < 1e-12 * max(1.0, np.abs(g_out).sum())
    for p in rng.choice(dem.size, size=20, replace=False):
        j, i = divmod(int(p), dem.shape[1])
        up, dn = dem.copy(), dem.copy()
        up[j, i] = np.float32(dem[j, i] + np.float32(1e-3))
        dn[j, i] = np.float32(dem[j, i] - np.float32(1e-3))
        fu = oracle_lib.fillsinks_pf(up, bc).astype(np.float64)
        fd = oracle_lib.fillsinks_pf(dn, bc).astype(np.float64)
        fdv = ((g_out * fu).sum() - (g_out * fd).sum()) / (float(up[j, i]) - float(dn[j, i]))
        assert abs(fdv - g_dem[j, i]) < 1e-6 * max(1.0, abs(g_dem[j, i])), (j, i, fdv, g_dem[j, i])


# ------------------------------------------------------------------ excess topography
@pytest.mark.parametrize("name", ["fsm_random_s0", "fsm_random_s1", "fsm_random_s2",
                                  "fsm_numerics"])
def test_fsm_golden(oracle_lib, name):
    g = load_golden(name)
    u, it = oracle_lib.fsm2d_f32(g["dem"], g["slopes"], float(g["cellsize"]))
    assert_bitwise(u, g["fsm"], name)
    # FP64 FSM and FP64 FMM agree with the reference's FMM within the
    # reference's own FSM==FMM tolerance (test/excesstopography.cpp:69-82)
    u64, _ = oracle_lib.fsm2d_f64(g["dem"], g["slopes"], float(g["cellsize"]))
    m64 = oracle_lib.fmm2d_f64(g["dem"], g["slopes"], float(g["cellsize"]))
    fin = np.isfinite(g["fmm"])
    # absolute 1e-4 for the reference's random DEMs (values below 100 m);
    # the numerics case reaches ~1e3 m, where the reference's float FMM
    # rounds at ~1e-4 m, so it is held to 1e-6 relative instead
    scale = 1.0 if name != "fsm_numerics" else 1e-2 * np.maximum(1.0, np.abs(u64[fin]))
    assert np.all(np.abs(u64[fin] - g["fmm"][fin]) <= 1e-4 * scale)
    assert np.nanmax(np.abs(m64 - u64)) <= 1e-4
    np.testing.assert_array_equal(np.isnan(u64), np.isnan(g["fsm"]))


def test_fsm_vs_reference_library(oracle_lib):
    if oracle_lib.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(4)
    for seed in range(4):
        ny, nx = rng.integers(1, 70, size=2)
        dem = (rng.uniform(size=(ny, nx)) * 300).astype(np.float32)
        dem[rng.uniform(size=dem.shape) < 0.05] = np.nan
        slopes = rng.uniform(0.05, 1.0, size=dem.shape).astype(np.float32)
        u, _ = oracle_lib.fsm2d_f32(dem, slopes, 10.0)
        assert_bitwise(u, oracle_lib.ref_fsm2d(dem, slopes, 10.0), f"seed {seed}")


def test_fsm_closed_forms(oracle_lib):
    big = 1.0e6
    # 1-D strip: u_i = i * fi (one-sided branch only, exact in double)
    dem = np.full((1, 50), big)
    dem[0, 0] = 0.0
    s = np.full(dem.shape, 0.5)
    u, _ = oracle_lib.fsm2d_f64(dem, s, 30.0)
    np.testing.assert_array_equal(u[0], 15.0 * np.arange(50))
    # plane: zero row, interior high -> u = k * fi on every column
    dem = np.full((20, 30), big)
    dem[0, :] = 0.0
    s = np.full(dem.shape, 0.5)
    u, _ = oracle_lib.fsm2d_f64(dem, s, 30.0)
    np.testing.assert_array_equal(u, 15.0 * np.arange(20)[:, None] * np.ones((1, 30)))
    # single source: axes k*fi, first diagonal fi*(1 + 1/sqrt(2))
    dem = np.full((10, 10), big)
    dem[0, 0] = 0.0
    u, _ = oracle_lib.fsm2d_f64(dem, np.full(dem.shape, 0.5), 30.0)
    assert u[0, 3] == 45.0 and u[3, 0] == 45.0
    assert abs(u[1, 1] - 15.0 * (1 + 1 / np.sqrt(2))) < 1e-12
    # constraint u <= dem and upwind gradient <= slope (+1e-4 on interior)
    dem = (np.random.default_rng(1).uniform(size=(30, 40)) * 200)
    s = np.random.default_rng(2).uniform(0.1, 0.9, size=dem.shape)
    u, _ = oracle_lib.fsm2d_f64(dem, s, 30.0)
    assert np.all(u <= dem)
    gx = np.maximum(0, np.maximum((u[:, 1:-1] - u[:, :-2]), (u[:, 1:-1] - u[:, 2:]))) / 30
    gy = np.maximum(0, np.maximum((u[1:-1, :] - u[:-2, :]), (u[1:-1, :] - u[2:, :]))) / 30
    g = np.hypot(gx[1:-1, :], gy[:, 1:-1])
    assert np.all(g - s[1:-1, 1:-1] < 1e-4)


@pytest.mark.parametrize("seed", range(3))
def test_fsm_adjoint_fd(oracle_lib, seed):
    rng = np.random.default_rng(200 + seed)
    ny, nx = 14, 17
    dem = rng.uniform(0, 300, size=(ny, nx))
    s = rng.uniform(0.1, 0.9, size=(ny, nx))
    g_out = rng.standard_normal((ny, nx))
    cs = 30.0
    u, _ = oracle_lib.fsm2d_f64(dem, s, cs)
    g_dem, g_s = oracle_lib.fsm2d_adjoint_f64(u, dem, s, cs, g_out)
    eps = 1e-6

    def L(d, sl):
        return (g_out * oracle_lib.fsm2d_f64(d, sl, cs)[0]).sum()

    for p in rng.choice(ny * nx, size=25, replace=False):
        j, i = divmod(int(p), nx)
        e = np.zeros((ny, nx))
        e[j, i] = eps
        fd_d = (L(dem + e, s) - L(dem - e, s)) / (2 * eps)
        fd_s = (L(dem, s + e) - L(dem, s - e)) / (2 * eps)
        assert abs(fd_d - g_dem[j, i]) < 1e-5 * max(1.0, abs(fd_d)), ("dem", j, i, fd_d, g_dem[j, i])
        assert abs(fd_s - g_s[j, i]) < 1e-5 * max(1.0, abs(fd_s)), ("slope", j, i, fd_s, g_s[j, i])


# ------------------------------------------------------------------ flats and value filters
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
    return dem


def test_identifyflats_vs_reference_and_invariants(oracle_lib):
    rng = np.random.default_rng(61)
    for shape in [(1, 1), (3, 3), (1, 40), (40, 1), (57, 83)]:
        dem = stencil_dem(rng, shape)
        out, n = oracle_lib.identifyflats(dem)
        if oracle_lib.ref() is not None:
            want, wn = oracle_lib.ref_identifyflats(dem)
            np.testing.assert_array_equal(out, want)
            assert n == wn
        assert n == int((out & 1).sum())
        ny, nx = shape
        for j in range(ny):
            for i in range(nx):
                nb = [(j + a, i + b) for a in (-1, 0, 1) for b in (-1, 0, 1)
                      if (a or b) and 0 <= j + a < ny and 0 <= i + b < nx]
                if out[j, i] & 1:  # flat: interior, nothing strictly lower
                    assert 0 < j < ny - 1 and 0 < i < nx - 1
                    assert all(not (dem[q] < dem[j, i]) and not np.isnan(dem[q]) or
                               dem[j, i] == -np.inf for q in nb)
                if out[j, i] & 2:  # sill: not flat, touches an equal flat
                    assert not out[j, i] & 1
                    assert any(out[q] & 1 and dem[q] == dem[j, i] for q in nb)
                if out[j, i] & 4:  # presill: a flat touching an equal non-flat
                    assert out[j, i] & 1
                    assert any(not out[q] & 1 and dem[q] == dem[j, i] for q in nb)


def test_value_filters_vs_reference(oracle_lib):
    if oracle_lib.ref() is None:
        pytest.skip("reference library not built")
    rng = np.random.default_rng(62)
    ses = [np.ones((3, 3)), np.array([[0, 1, 0], [1, 1, 1], [0, 1, 0]]), np.ones((1, 1)),
           np.ones((2, 4)), rng.uniform(size=(2, 5, 3)) < 0.5]
    for shape in [(1, 1), (2, 7), (45, 38)]:
        dem = stencil_dem(rng, shape, flats=False)
        for se in ses:
            for op in ("min", "max"):
                assert_bitwise(oracle_lib.value_filter(dem, se, op),
                               oracle_lib.ref_value_filter(dem, se, op), f"{op} {se.shape}")
        for width in (1, 2, 3, 5, 8):
            for op in ("min", "max"):
                assert_bitwise(oracle_lib.value_filter_square(dem, width, op),
                               oracle_lib.ref_value_filter_square(dem, width, op),
                               f"square {op} {width}")


def test_value_filter_mathematics(oracle_lib):
    """Pins that do not need the reference: on NaN-free inputs without
    signed zeros, the square filter equals the general filter with a full
    square structuring element (separability), max is -min(-x) (duality),
    a 1x1 element is the identity and the result brackets the input."""
    rng = np.random.default_rng(63)
    dem = rng.uniform(-20, 50, size=(31, 44)).astype(np.float32) + np.float32(0.0)
    for width in (1, 3, 4, 7):
        se = np.ones((width, width), np.uint8)
        for op in ("min", "max"):
            assert_bitwise(oracle_lib.value_filter_square(dem, width, op),
                           oracle_lib.value_filter(dem, se, op), f"separable {op} {width}")
        assert_bitwise(oracle_lib.value_filter(dem, se, "max"),
                       -oracle_lib.value_filter(-dem, se, "min"), f"duality {width}")
    assert_bitwise(oracle_lib.value_filter(dem, np.ones((1, 1)), "min"), dem, "identity")
    lo = oracle_lib.value_filter(dem, np.ones((3, 3)), "min")
    hi = oracle_lib.value_filter(dem, np.ones((3, 3)), "max")
    assert (lo <= dem).all() and (hi >= dem).all()
    # brute force on a tiny grid: window minimum of the cells in range
    small = dem[:6, :7]
    got = oracle_lib.value_filter(small, np.ones((3, 3)), "min")
    for j in range(6):
        for i in range(7):
            assert got[j, i] == small[max(j - 1, 0):j + 2, max(i - 1, 0):i + 2].min()


# ------------------------------------------------------------------ gwdt costs
def gwdt_inputs(rng, shape, levels=4, zero_depth=False):
    """Random flats (none on the border) with depths from few levels, so
    that components often hold several deepest pixels."""
    flats = (rng.uniform(size=shape) < rng.uniform(0.3, 0.95)).astype(np.int32)
    flats[0, :] = flats[-1, :] = 0
    flats[:, 0] = flats[:, -1] = 0
    original = rng.uniform(0, 10, size=shape).astype(np.float32)
    depth = rng.integers(0, levels, size=shape).astype(np.float32)
    if zero_depth:
        depth[:] = 0
    return flats, original, (original + depth).astype(np.float32)


def test_gwdt_computecosts_vs_reference_and_mathematics(oracle_lib):
    """The oracle against the compiled reference (bitwise), and against what
    the definition fixes: labels constant on the 8-connected components of
    the flats (scipy.ndimage.label), each label a deepest pixel of its
    component, cost = (max depth - depth) + 0.1 rounded from double, zero
    on non-flats."""
    from scipy import ndimage

    rng = np.random.default_rng(71)
    for trial in range(60):
        shape = tuple(int(v) for v in rng.integers(3, 45, size=2))
        flats, orig, filled = gwdt_inputs(rng, shape, int(rng.integers(1, 5)), trial % 5 == 0)
        costs, cc = oracle_lib.gwdt_computecosts(flats, orig, filled)
        if oracle_lib.ref() is not None:
            rc, rcc = oracle_lib.ref_gwdt_computecosts(flats, orig, filled)
            np.testing.assert_array_equal(cc, rcc)
            np.testing.assert_array_equal(costs.view(np.uint32), rc.view(np.uint32))
        if trial % 3 == 1 and oracle_lib.ref() is not None:
            # NaN depths (INF - INF) scattered among the flats
            o2, f2 = orig.copy(), filled.copy()
            m = rng.uniform(size=shape) < 0.2
            o2[m] = f2[m] = np.inf
            c2, cc2 = oracle_lib.gwdt_computecosts(flats, o2, f2)
            rc2, rcc2 = oracle_lib.ref_gwdt_computecosts(flats, o2, f2)
            np.testing.assert_array_equal(cc2, rcc2)
            np.testing.assert_array_equal(np.isnan(c2), np.isnan(rc2))
            ok = ~np.isnan(c2)
            np.testing.assert_array_equal(c2[ok].view(np.uint32), rc2[ok].view(np.uint32))
        f = (flats & 1) > 0
        assert np.all(cc[~f] == 0) and np.all(costs[~f] == 0)
        lab, n = ndimage.label(f, structure=np.ones((3, 3)))
        depth = filled - orig
        for k in range(1, n + 1):
            m = lab == k
            labels = np.unique(cc[m])
            assert labels.size == 1
            root = int(labels[0])
            assert m.flat[root]
            dmax = depth[m].max()
            assert depth.flat[root] == dmax
            want = ((dmax - depth[m]).astype(np.float64) + 0.1).astype(np.float32)
            np.testing.assert_array_equal(costs[m].view(np.uint32), want.view(np.uint32))
            deepest = np.flatnonzero(m & (depth == dmax))
            if deepest.size == 1:
                assert root == deepest[0]
