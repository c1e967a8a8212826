"""Batched persistent launch: every item of a batch against the oracle.

fillsinks items: output and spill index bit-exact to the oracle's
lexicographic Priority-Flood, g_dem bit-exact to the single-grid backward on
the same spill index and within 1e-6 relative of the FP64 scatter.
Excess-topography items: output within 1e-4 relative (denominator
max(|u64|, 1 m)) of the FP64 oracle, gradients within 1e-4 relative of the
FP64 adjoint evaluated at the batch's own converged state.

Graph capture: a captured launch is replayed several times with new inputs
loaded and unrelated eager GPU work in between; every replay is checked
against an eager run on the same inputs.
"""
import numpy as np
import pytest
import torch

from conftest import assert_bitwise, assert_scatter_close, scatter_scale_cases

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (1, 300), (300, 1), (3, 3), (65, 129), (130, 257), (64, 64), (200, 100),
          (128, 128), (129, 127), (257, 65), (40, 700)]


def fill_items(oracle_lib, seed=0, shapes=SHAPES):
    rng = np.random.default_rng(seed)
    dems, bcs = [], []
    for k, (ny, nx) in enumerate(shapes):
        d = oracle_lib.grid_random(seed * 100 + k, (ny, nx))
        d[rng.uniform(size=d.shape) < 0.02] = np.nan
        if k % 3 == 0:
            d = np.round(d * 8) / 8  # ties and flats
        bc = (rng.uniform(size=d.shape) < 0.005).astype(np.uint8)
        bc[0, :] = bc[-1, :] = 1
        bc[:, 0] = bc[:, -1] = 1
        dems.append(d.astype(np.float32))
        bcs.append(bc)
    return dems, bcs


def fsm_items(seed=0, shapes=SHAPES):
    rng = np.random.default_rng(seed)
    dems, slopes = [], []
    for ny, nx in shapes:
        d = (rng.uniform(size=(ny, nx)) * 300).astype(np.float32)
        d[rng.uniform(size=d.shape) < 0.02] = np.nan
        dems.append(d)
        slopes.append(rng.uniform(0.1, 0.9, size=(ny, nx)).astype(np.float32))
    return dems, slopes


def grads(shapes, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(s).astype(np.float32) for s in shapes]


def check_fill(b, oracle_lib, dems, bcs, gs, what):
    from libtopotoolbox_amd import ops

    r = b.outputs()
    for k, (d, bc, g) in enumerate(zip(dems, bcs, gs)):
        want, wspill = oracle_lib.fillsinks_spill(d, bc)
        assert_bitwise(r["output"][k].cpu().numpy(), want, f"{what} item {k} output")
        spill = r["spill"][k].cpu().numpy()
        np.testing.assert_array_equal(spill.astype(np.int64), wspill,
                                      err_msg=f"{what} item {k} spill")
        gd = r["g_dem"][k].cpu().numpy()
        single = ops.fillsinks_backward(r["spill"][k].contiguous(),
                                        torch.from_numpy(g).cuda()).cpu().numpy()
        assert_bitwise(gd, single, f"{what} item {k} g_dem vs single grid")
        w = oracle_lib.fillsinks_adjoint_f64(wspill, g.astype(np.float64))
        err = np.abs(gd - w) / np.maximum(1.0, np.abs(w))
        assert err.max() <= 1e-6, (what, k, err.max())


def test_batch_fill_adjoint_scale_and_nonfinite(cuda, oracle_lib):
    """Per-item fixed-point scales and NaN / INF propagation: the cases of
    very different max |g_out| share one launch, each item equals the
    single-grid backward bit for bit and the FP64 scatter to 1e-6."""
    from libtopotoolbox_amd import ops
    from libtopotoolbox_amd.batched import Batch

    cases = scatter_scale_cases(oracle_lib)
    b = Batch("fillsinks", [c[1].shape for c in cases])
    b.load([c[1] for c in cases], bc=[c[2] for c in cases], g_out=[c[3] for c in cases])
    b.run()
    r = b.outputs()
    for k, (name, dem, bc, g) in enumerate(cases):
        want, wspill = oracle_lib.fillsinks_spill(dem, bc)
        assert_bitwise(r["output"][k].cpu().numpy(), want, f"{name}: output")
        np.testing.assert_array_equal(r["spill"][k].cpu().numpy().astype(np.int64), wspill)
        gd = r["g_dem"][k].cpu().numpy()
        single = ops.fillsinks_backward(r["spill"][k].contiguous(),
                                        torch.from_numpy(g).cuda()).cpu().numpy()
        assert_bitwise(gd, single, f"{name}: g_dem vs single grid")
        w = oracle_lib.fillsinks_adjoint_f64(wspill, g.astype(np.float64))
        assert_scatter_close(gd, w, g, name)


def fsm_rel(u, u64):
    nan32, nan64 = np.isnan(u), np.isnan(u64)
    np.testing.assert_array_equal(nan32, nan64)
    ok = ~nan64 & ~np.isinf(u64)
    np.testing.assert_array_equal(u[~ok & ~nan64], u64[~ok & ~nan64].astype(np.float32))
    if not ok.any():
        return 0.0
    return float((np.abs(u[ok] - u64[ok]) / np.maximum(np.abs(u64[ok]), 1.0)).max())


def check_fsm(b, oracle_lib, dems, slopes, gs, cellsize, what, adjoint=True):
    r = b.outputs()
    for k, (d, s, g) in enumerate(zip(dems, slopes, gs)):
        u = r["output"][k].cpu().numpy()
        u64, _ = oracle_lib.fsm2d_f64(d, s, cellsize)
        assert fsm_rel(u, u64) <= 1e-4, (what, k)
        if not adjoint:
            continue
        wd, ws = oracle_lib.fsm2d_adjoint_f64(u, d, s, cellsize, g)
        for got, want, name in ((r["g_dem"][k], wd, "dem"), (r["g_slopes"][k], ws, "slope")):
            got = got.cpu().numpy()
            err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
            assert err.max() <= 1e-4, (what, k, name, err.max())


@pytest.mark.parametrize("tile", [(2, 1), (1, 2), (1, 1)])
def test_batch_fillsinks_fwd_adjoint(cuda, oracle_lib, tile):
    from libtopotoolbox_amd.batched import Batch

    dems, bcs = fill_items(oracle_lib, 1)
    gs = grads(SHAPES, 2)
    b = Batch("fillsinks", SHAPES, tile=tile)
    b.load(dems, bc=bcs, g_out=gs)
    st = b.run(stats=True)
    assert st["error"] == 0
    check_fill(b, oracle_lib, dems, bcs, gs, f"tile {tile}")


@pytest.mark.parametrize("tile", [(2, 1), (1, 2), (1, 1)])
def test_batch_fsm_fwd_adjoint(cuda, oracle_lib, tile):
    from libtopotoolbox_amd.batched import Batch

    dems, slopes = fsm_items(3)
    gs = grads(SHAPES, 4)
    b = Batch("excesstopography", SHAPES, tile=tile, cellsize=30.0)
    b.load(dems, slopes=slopes, g_out=gs)
    st = b.run(stats=True)
    assert st["error"] == 0
    check_fsm(b, oracle_lib, dems, slopes, gs, 30.0, f"tile {tile}")


def test_batch_per_item_cellsize(cuda, oracle_lib):
    from libtopotoolbox_amd.batched import Batch

    shapes = SHAPES[4:8]
    dems, slopes = fsm_items(5, shapes)
    cs = [10.0, 30.0, 2.5, 90.0]
    b = Batch("excesstopography", shapes, cellsize=cs)
    b.load(dems, slopes=slopes)
    assert b.run(forward_only=True, stats=True)["error"] == 0
    r = b.outputs()["output"]
    for k in range(len(shapes)):
        u64, _ = oracle_lib.fsm2d_f64(dems[k], slopes[k], cs[k])
        assert fsm_rel(r[k].cpu().numpy(), u64) <= 1e-4


def test_batch_forward_only_and_sweep_cap(cuda, oracle_lib):
    """Forward-only leaves the gradient buffers alone; a tiny per-visit
    sweep cap (tiles requeue themselves) changes nothing in the results."""
    from libtopotoolbox_amd.batched import Batch

    dems, bcs = fill_items(oracle_lib, 6)
    b = Batch("fillsinks", SHAPES)
    b.load(dems, bc=bcs)
    b.g_dem.fill_(7.0)
    st = b.run(forward_only=True, max_local=2, stats=True)
    assert st["error"] == 0 and st["requeues"] > 0
    assert bool((b.g_dem == 7.0).all())
    r = b.outputs()
    for k, (d, bc) in enumerate(zip(dems, bcs)):
        want, wspill = oracle_lib.fillsinks_spill(d, bc)
        assert_bitwise(r["output"][k].cpu().numpy(), want, f"capped item {k}")
        np.testing.assert_array_equal(r["spill"][k].cpu().numpy().astype(np.int64), wspill)

    # per-item caps (1 sweep on every other item, the launch's cap of 3 on
    # the rest), forward + adjoint: the same results
    gs = grads(SHAPES, 9)
    b.load(dems, bc=bcs, g_out=gs)
    b.set_sweep_caps([1 if k % 2 else 0 for k in range(len(SHAPES))])
    st = b.run(max_local=3, stats=True)
    assert st["error"] == 0 and st["requeues"] > 0
    check_fill(b, oracle_lib, dems, bcs, gs, "per-item caps")

    dems, slopes = fsm_items(7)
    gs = grads(SHAPES, 8)
    f = Batch("excesstopography", SHAPES, cellsize=30.0)
    f.load(dems, slopes=slopes, g_out=gs)
    assert f.run(max_local=2, stats=True)["error"] == 0
    check_fsm(f, oracle_lib, dems, slopes, gs, 30.0, "capped fsm")
    f.set_sweep_caps([k % 3 for k in range(len(SHAPES))])
    assert f.run(max_local=5, stats=True)["error"] == 0
    check_fsm(f, oracle_lib, dems, slopes, gs, 30.0, "per-item capped fsm")


def test_batch_graph_replays(cuda, oracle_lib):
    """Replays of a captured launch, with new inputs and eager work between
    them, each checked against an eager launch on the same inputs."""
    import libtopotoolbox_amd as tt
    from libtopotoolbox_amd.batched import Batch

    sets = []
    for seed in (11, 12, 13):
        dems, bcs = fill_items(oracle_lib, seed)
        sets.append((dems, bcs, grads(SHAPES, seed + 50)))
    b = Batch("fillsinks", SHAPES)
    eager = []
    for dems, bcs, gs in sets:
        b.load(dems, bc=bcs, g_out=gs)
        b.run()
        torch.cuda.synchronize()
        eager.append({k: [v.clone() for v in vs] for k, vs in b.outputs().items()})
    b.load(sets[0][0], bc=sets[0][1], g_out=sets[0][2])
    g = b.capture()
    for rep, idx in enumerate((1, 0, 2, 2)):
        dems, bcs, gs = sets[idx]
        b.load(dems, bc=bcs, g_out=gs)
        b.output.fill_(-1.0)
        b.g_dem.fill_(-1.0)
        # unrelated eager work on the same device between replays
        x = torch.from_numpy(dems[5]).cuda()
        tt.fillsinks(x, torch.ones_like(x, dtype=torch.uint8))
        g.replay()
        torch.cuda.synchronize()
        got = b.outputs()
        for key in ("output", "spill", "g_dem"):
            for k in range(len(SHAPES)):
                assert_bitwise(got[key][k].cpu().numpy(), eager[idx][key][k].cpu().numpy(),
                               f"replay {rep} set {idx} {key} item {k}")


def test_batch_graph_replays_fsm(cuda, oracle_lib):
    from libtopotoolbox_amd.batched import Batch

    sets = []
    for seed in (21, 22):
        dems, slopes = fsm_items(seed)
        sets.append((dems, slopes, grads(SHAPES, seed + 50)))
    b = Batch("excesstopography", SHAPES, cellsize=30.0)
    b.load(sets[0][0], slopes=sets[0][1], g_out=sets[0][2])
    g = b.capture()
    for rep, idx in enumerate((1, 0, 1)):
        dems, slopes, gs = sets[idx]
        b.load(dems, slopes=slopes, g_out=gs)
        b.output.fill_(-1.0)
        g.replay()
        torch.cuda.synchronize()
        check_fsm(b, oracle_lib, dems, slopes, gs, 30.0, f"replay {rep}")


def test_autotune_small_batch(cuda, oracle_lib, tmp_path, monkeypatch):
    from libtopotoolbox_amd import autotune
    from libtopotoolbox_amd.batched import Batch

    monkeypatch.setenv("TT_AMD_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    dems, bcs = fill_items(oracle_lib, 31)
    gs = grads(SHAPES, 32)
    b = Batch("fillsinks", SHAPES)
    b.load(dems, bc=bcs, g_out=gs)
    res = autotune.tune(b, reps=1, caps=[0, 4])
    # 3 tiles x 2 caps, then the other cap for the one size class present
    assert len(res["table"]) == 7 and not res["cached"]
    assert res["classes"] == {"small": len(SHAPES), "medium": 0, "large": 0}
    assert tuple(res["tile"]) == b.tile
    b.run(max_local=res["max_local"])
    check_fill(b, oracle_lib, dems, bcs, gs, "tuned")
    again = autotune.tune(b, reps=1, caps=[0, 4])
    assert again["cached"] and again["tile"] == res["tile"]
    assert again["class_caps"] == res["class_caps"]
