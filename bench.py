"""Benchmark: fillsinks / excess topography on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload fill|fsm]

A step is one call of the hot path over one synthetic 8192 x 8192 float32
DEM already resident in HBM:

* fill (default): fillsinks on a fractal DEM (Hurst 0.7, 0-3000 m) with 200
  paraboloid depressions (30% flat-bottomed) and a 2% NaN blob on an edge,
  NaN cells handled by the reference caller convention (bc = 1, dem = -INF).
* fsm: excesstopography_fsm2d on a fractal DEM (Hurst 0.8, 0-1000 m, 5%
  pits), cellsize 30 m, threshold slopes from 64 Voronoi cells U(0.2, 0.8).

Prints one JSON line. `value` is Gcells/s of the selected workload;
`roofline` prices the persistent kernel's launch time against the
algorithmic bytes (fill 9 B/cell: dem + bc + output; fsm 12 B/cell: dem +
slopes + output) and 8 TB/s; `cpu_baseline` times the reference path on
the host (one core) on a bounded crop of the same DEM. The other workload
is measured too and reported under "extra", with the forward+adjoint
passes, reconstruct from a sparse and a dense marker on both work queues,
the batched launch and the stencils next to the path (identifyflats,
min / max filters).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = ("fillsinks/excess-topo Gcells/s & %HBM peak, 8192² fp32; "
          "fwd+adjoint vs plain")
HBM_PEAK_GBS = 8000.0
N = 8192
BYTES_PER_CELL = {"fill": 9, "fsm": 12}
# forward + adjoint, compulsory bytes per cell: fill dem 4 + bc 1 + output 4 +
# g_out 4 + g_dem 4 (the spill index kept out of the count, as if on chip);
# fsm dem, slopes, output, g_out, g_dem, g_slopes, 4 each
BYTES_FWD_ADJ = {"fill": 17, "fsm": 24}


def make_inputs(workload: str, n: int):
    from libtopotoolbox_amd import synthetic as S

    if workload == "fill":
        dem = S.dem_depressions(n, n, 2)
        dem, bc, nanm = S.nan_to_bc(dem, S.edge_bc(n, n))
        return {"dem": dem, "bc": bc}
    dem = S.dem_pits(n, n, 1)
    try:
        import torch
        dev = "cuda" if torch.cuda.is_available() else None
    except ImportError:
        dev = None
    slopes = S.voronoi_slopes(n, n, 64, 0.2, 0.8, 33, device=dev)
    return {"dem": dem, "slopes": slopes}


def gpu_run(workload: str, arrays: dict, steps: int, warmup: int):
    import torch

    import libtopotoolbox_amd as tt

    dev = torch.device("cuda:0")
    dem = torch.from_numpy(arrays["dem"]).to(dev)
    out = torch.empty_like(dem)
    if workload == "fill":
        bc = torch.from_numpy(arrays["bc"]).to(dev)

        def step(stats=False):
            return tt.fillsinks(dem, bc, out=out, stats=stats)
    else:
        slopes = torch.from_numpy(arrays["slopes"]).to(dev)

        def step(stats=False):
            return tt.excesstopography_fsm2d(dem, slopes, 30.0, out=out, stats=stats)

    _, st = step(stats=True)  # also a warmup; reports visits and phase times
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        step()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_launch_ms = [a.elapsed_time(b) for a, b in ev]
    return {"wall_s": wall, "launch_ms": per_launch_ms, "stats": st, "out": out}


def time_events(fn, steps: int):
    """Per-call milliseconds of fn() on the current stream (HIP events)."""
    import torch

    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall * 1e3 / steps, [a.elapsed_time(b) for a, b in ev]


def fwd_adjoint_run(workload: str, arrays: dict, steps: int, plain_ms: float):
    """Forward + adjoint at full size (fillsinks: spill forward + scatter;
    excess topography: forward + lambda relaxation), seed-44 N(0, 1)
    upstream gradient, against the plain forward's time."""
    import torch

    from libtopotoolbox_amd import ops

    dev = torch.device("cuda:0")
    dem = torch.from_numpy(arrays["dem"]).to(dev)
    g = torch.randn(dem.shape, generator=torch.Generator(device=dev).manual_seed(44),
                    device=dev)
    if workload == "fill":
        bc = torch.from_numpy(arrays["bc"]).to(dev)

        def step():
            _, spill = ops.fillsinks_spill(dem, bc)
            return ops.fillsinks_backward(spill, g)
    else:
        slopes = torch.from_numpy(arrays["slopes"]).to(dev)
        out = torch.empty_like(dem)

        def step():
            ops.excesstopography_fsm2d(dem, slopes, 30.0, out=out)
            return ops.excesstopography_fsm2d_backward(out, dem, slopes, 30.0, g)
    step()
    ms, _ = time_events(step, steps)
    cells = dem.numel()
    return {"ms_per_step": round(ms, 4), "gcells_per_s": round(cells / (ms * 1e-3) / 1e9, 4),
            "plain_forward_ms": round(plain_ms, 4),
            "fwd_adjoint_over_plain": round(ms / plain_ms, 3),
            "bytes_per_cell": BYTES_FWD_ADJ[workload]}


def batched_inputs(op: str, count: int, seed: int = 5):
    """The uneven-tile workload (libtopotoolbox_amd.synthetic.config5_items)."""
    from libtopotoolbox_amd import synthetic as S

    return S.config5_items(op, count, seed, workers=min(16, os.cpu_count() or 1))


def batched_run(op: str, count: int, steps: int, cpu: bool = True):
    """One graph-captured persistent launch over all tiles (forward +
    adjoint, autotuned tile shape and per-size-class sweep caps) against the same tiles run
    one at a time through the plain single-grid entry points."""
    import torch

    from libtopotoolbox_amd import autotune, ops
    from libtopotoolbox_amd.batched import Batch

    t0 = time.perf_counter()
    shapes, dems, aux, gs = batched_inputs(op, count)
    gen_s = time.perf_counter() - t0
    cells = sum(a * b for a, b in shapes)
    b = Batch(op, shapes, cellsize=30.0)
    if op == "fillsinks":
        b.load(dems, bc=aux, g_out=gs)
    else:
        b.load(dems, slopes=aux, g_out=gs)
    cpu_res = batched_cpu_baseline(op, dems, aux, min(16, os.cpu_count() or 1)) if cpu else None
    del dems, aux, gs
    tune = autotune.tune(b, reps=2, use_cache=False)
    graph = b.capture(max_local=tune["max_local"])
    graph.replay()
    ms, per = time_events(graph.replay, steps)
    st = b.run(max_local=tune["max_local"], stats=True)
    fwd_ms, _ = time_events(lambda: b.run(forward_only=True, max_local=tune["max_local"]),
                            max(2, steps // 2))

    # the same tiles one at a time through the plain kernels
    dv, ov = b.views(b.dem), b.views(b.output)
    gv = b.views(b.g_out)
    if op == "fillsinks":
        bcv = b.views(b.bc)

        def seq():
            for d, c, g in zip(dv, bcv, gv):
                _, spill = ops.fillsinks_spill(d, c)
                ops.fillsinks_backward(spill, g)

        def seq_fwd():
            for d, c, o in zip(dv, bcv, ov):
                ops.fillsinks(d, c)
    else:
        sv = b.views(b.slopes)

        def seq():
            for d, s, g in zip(dv, sv, gv):
                u = ops.excesstopography_fsm2d(d, s, 30.0)
                ops.excesstopography_fsm2d_backward(u, d, s, 30.0, g)

        def seq_fwd():
            for d, s in zip(dv, sv):
                ops.excesstopography_fsm2d(d, s, 30.0)
    seq()
    seq_ms, _ = time_events(seq, 2)
    seq_fwd_ms, _ = time_events(seq_fwd, 2)

    # spot parity: the batch's results equal the single-grid ones
    graph.replay()
    torch.cuda.synchronize()
    r = b.outputs()
    worst = 0.0
    for k in range(0, count, max(1, count // 16)):
        if op == "fillsinks":
            o, spill = ops.fillsinks_spill(dv[k], bcv[k])
            gd = ops.fillsinks_backward(spill, gv[k])
            same = (torch.equal(o.view(torch.int32), r["output"][k].view(torch.int32))
                    and torch.equal(spill, r["spill"][k])
                    and torch.equal(gd.view(torch.int32), r["g_dem"][k].view(torch.int32)))
            worst = max(worst, 0.0 if same else 1.0)
        else:
            u = ops.excesstopography_fsm2d(dv[k], sv[k], 30.0)
            ok = torch.isfinite(u)
            rel = ((u - r["output"][k]).abs() / u.abs().clamp_min(1.0))[ok]
            worst = max(worst, float(rel.max()) if rel.numel() else 0.0)
    bytes_cell = BYTES_FWD_ADJ["fill" if op == "fillsinks" else "fsm"]
    launch_ms = float(np.median(per))
    res = {
        "tiles": count, "cells": cells, "input_gen_s": round(gen_s, 1),
        "autotune": {k: tune[k] for k in ("tile", "max_local", "class_caps", "classes",
                                          "class_edges_cells", "global", "per_class_ms",
                                          "gain_vs_global", "table")},
        "graph_ms": round(ms, 3), "gcells_per_s": round(cells / (ms * 1e-3) / 1e9, 4),
        "forward_only_ms": round(fwd_ms, 3),
        "one_at_a_time_ms": round(seq_ms, 3), "one_at_a_time_forward_ms": round(seq_fwd_ms, 3),
        "one_at_a_time_gcells_per_s": round(cells / (seq_ms * 1e-3) / 1e9, 4),
        "speedup_vs_one_at_a_time": round(seq_ms / ms, 3),
        "cpu_baseline": cpu_res,
        "roofline": {"bound": "hbm", "bytes_per_cell": bytes_cell,
                     "achieved": round(bytes_cell * cells / (launch_ms * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(bytes_cell * cells / (launch_ms * 1e-3) / 1e9
                                   / HBM_PEAK_GBS, 4)},
        "kernel_stats": {k: st[k] for k in ("requeues", "visits", "grid", "tile_x", "tile_y")},
        "parity_spot": ("bitwise" if op == "fillsinks" and worst == 0.0 else
                        ("MISMATCH" if op == "fillsinks" else f"max rel {worst:.2e}")),
    }
    del b, graph
    torch.cuda.empty_cache()
    return res


# Stencils next to the path, bytes per cell: identifyflats reads the DEM and
# writes int32 flags; a value filter reads the DEM and writes the result; the
# square filter's contract also writes tmp (the pass along dims[0]).
# gwdt_computecosts reads flats, original and filled DEMs and writes float
# costs and int64 labels.
def reconstruct_run(arrays: dict, steps: int):
    """reconstruct (by dilation) at n x n on the fill workload's DEM, from a
    sparse marker (the negated DEM on the bc cells, -INF elsewhere: fillsinks
    posed as a reconstruction) and a dense one (marker = DEM - 50 under mask =
    DEM), each on the FIFO queue (what AUTO runs) and the lowest-level-first
    queue: ms per call from HIP events (the marker copy the call starts from
    is outside the timed region), the roofline against 12 B/cell (marker read
    and written, mask read), and whether the two queues' outputs are
    bit-identical."""
    import torch

    import libtopotoolbox_amd as tt

    dev = torch.device("cuda:0")
    dem = torch.from_numpy(arrays["dem"]).to(dev)
    bc = torch.from_numpy(arrays["bc"]).to(dev)
    n = dem.numel()
    cases = {
        "sparse_marker": (torch.where(bc == 1, -dem, torch.full_like(dem, -float("inf"))), -dem),
        "dense_marker": (dem - 50.0, dem),
    }
    res = {}
    for name, (marker, mask) in cases.items():
        work = torch.empty_like(marker)
        r, outs = {}, {}
        for sched in ("fifo", "priority"):
            tt.set_fill_schedule(sched)
            try:
                def call():
                    work.copy_(marker)
                    tt.reconstruct(work, mask, inplace=True)

                call()
                stream = torch.cuda.current_stream()
                ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                      for _ in range(steps)]
                for a, b in ev:
                    work.copy_(marker)
                    a.record(stream)
                    tt.reconstruct(work, mask, inplace=True)
                    b.record(stream)
                torch.cuda.synchronize()
                ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
                outs[sched] = work.clone()
            finally:
                tt.set_fill_schedule("auto")
            gbs = 12 * n / (ms * 1e-3) / 1e9
            r[sched] = {"ms": round(ms, 4), "gcells_per_s": round(n / (ms * 1e-3) / 1e9, 3),
                        "roofline": {"bound": "hbm", "achieved": round(gbs, 1),
                                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_cell": 12}}
        r["auto_runs"] = "fifo"
        r["queues_bit_identical"] = bool(torch.equal(outs["fifo"].view(torch.int32),
                                                     outs["priority"].view(torch.int32)))
        res[name] = r
    return res


STENCIL_BYTES = {"identifyflats": 8, "min_filter_3x3": 8, "max_filter_square_5": 12,
                 "gwdt_computecosts": 24}


def stencil_run(n: int, steps: int, cpu: bool):
    """identifyflats, min_filter (3x3), max_filter_square (width 5) and
    gwdt_computecosts (with the flats of the filled DEM) on the filled
    fill-workload DEM (n x n): time per call from HIP events, the roofline
    against the bytes above, and the reference's CPU routine on a
    1024 x 1024 crop."""
    import torch

    import libtopotoolbox_amd as tt

    arrays = make_inputs("fill", n)
    dev = torch.device("cuda:0")
    original = torch.from_numpy(arrays["dem"]).to(dev)
    dem = tt.fillsinks(original, torch.from_numpy(arrays["bc"]).to(dev))
    flats = tt.identifyflats(dem, count=False)
    del arrays
    se = torch.ones((3, 3), dtype=torch.uint8, device=dev)
    out = torch.empty_like(dem)
    tmp = torch.empty_like(dem)
    calls = {
        "identifyflats": lambda: tt.identifyflats(dem, count=False),
        "min_filter_3x3": lambda: tt.min_filter(dem, se, out=out),
        "max_filter_square_5": lambda: tt.max_filter_square(dem, 5, out=out, tmp=tmp),
        "gwdt_computecosts": lambda: tt.gwdt_computecosts(flats, original, dem),
    }
    res = {}
    crop = None
    if cpu:
        import oracle

        crop = np.ascontiguousarray(dem[:1024, :1024].cpu().numpy())
        crop_orig = np.ascontiguousarray(original[:1024, :1024].cpu().numpy())
        crop_flats = np.ascontiguousarray(flats[:1024, :1024].cpu().numpy())
    for name, fn in calls.items():
        fn()
        _, per = time_events(fn, steps)
        ms = float(np.mean(per))
        gbs = STENCIL_BYTES[name] * n * n / (ms * 1e-3) / 1e9
        r = {"ms": round(ms, 4), "gcells_per_s": round(n * n / (ms * 1e-3) / 1e9, 2),
             "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                          "bytes_per_cell": STENCIL_BYTES[name]}}
        if crop is not None and oracle.ref() is not None:
            ref = {"identifyflats": lambda: oracle.ref_identifyflats(crop),
                   "min_filter_3x3": lambda: oracle.ref_value_filter(
                       crop, np.ones((3, 3), np.uint8), "min"),
                   "max_filter_square_5": lambda: oracle.ref_value_filter_square(
                       crop, 5, "max"),
                   "gwdt_computecosts": lambda: oracle.ref_gwdt_computecosts(
                       crop_flats, crop_orig, crop)}[name]
            t0 = time.perf_counter()
            for _ in range(5):
                ref()
            dt = (time.perf_counter() - t0) / 5
            r["cpu_baseline"] = {"value": round(crop.size / dt / 1e9, 4), "unit": "Gcells/s",
                                 "cores": 1, "kind": "reference",
                                 "sample": "1024x1024 crop of the filled DEM, 5 calls"}
        res[name] = r
    del dem, out, tmp, flats, original
    torch.cuda.empty_cache()
    return res


def fmm3d_run(n: int, steps: int, cpu: bool, nlayers: int = 4):
    """excesstopography_fmm3d on the excess-topography workload's DEM
    (n x n, cellsize 30 m) with `nlayers` horizontal layers whose tops are
    equally spaced over the DEM's range, and threshold slopes alternating
    1.0 / 0.5, as in the reference's own test (test/excesstopography.cpp:
    84-145). Reports the time per call from HIP events and a roofline on
    dem + layer tops + output. The compiled reference is timed on a
    2048 x 2048 crop. Parity for fmm3d is partial (the layered rule is
    order-dependent), so no parity spot is reported."""
    import torch

    import libtopotoolbox_amd as tt
    from libtopotoolbox_amd import synthetic as S

    dem_np = S.dem_pits(n, n, 1)
    finite = dem_np[np.isfinite(dem_np)]
    lo, hi = float(finite.min()), float(finite.max())
    tops = np.float32(lo) + np.float32(hi - lo) * (np.arange(nlayers, dtype=np.float32) + 1) \
        / np.float32(nlayers)
    slopes_np = np.where(np.arange(nlayers) % 2 == 0, 1.0, 0.5).astype(np.float32)
    dev = torch.device("cuda:0")
    dem = torch.from_numpy(dem_np).to(dev)
    lith = torch.from_numpy(tops).to(dev).expand(n, n, nlayers).contiguous()
    slopes = torch.from_numpy(slopes_np).to(dev)
    out = torch.empty_like(dem)
    call = lambda: tt.excesstopography_fmm3d(dem, lith, slopes, 30.0, out=out)
    call()
    _, per = time_events(call, steps)
    ms = float(np.mean(per))
    bpc = 8 + 4 * nlayers
    gbs = bpc * n * n / (ms * 1e-3) / 1e9
    res = {"ms": round(ms, 4), "gcells_per_s": round(n * n / (ms * 1e-3) / 1e9, 3),
           "layers": nlayers, "layer_tops_m": [round(float(t), 2) for t in tops],
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "bytes_per_cell": bpc}}
    if cpu:
        import oracle

        if oracle.ref() is not None:
            c = 2048
            crop = np.ascontiguousarray(dem_np[:c, :c])
            lith_c = np.ascontiguousarray(np.broadcast_to(tops, (c, c, nlayers)))
            t0 = time.perf_counter()
            oracle.ref_fmm3d(crop, lith_c, slopes_np, 30.0)
            dt = time.perf_counter() - t0
            res["cpu_baseline"] = {"value": round(c * c / dt / 1e9, 5), "unit": "Gcells/s",
                                   "cores": 1, "kind": "reference", "seconds": round(dt, 2),
                                   "sample": "2048x2048 crop of the DEM, same layers, one call"}
    del dem, lith, slopes, out
    torch.cuda.empty_cache()
    return res


def roofline(workload: str, cells: int, launch_ms: float, traffic):
    algo = BYTES_PER_CELL[workload] * cells
    achieved = algo / (launch_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic}


def measured_traffic(workload: str):
    """HBM bytes per launch from the committed rocprofv3 counter summary
    (profiles/traffic.json, written by `bench.py --traffic`), only when it
    was measured from the kernel sources present now (its source stamp
    equals the library's); otherwise None: a figure from other kernels
    would be stale."""
    from libtopotoolbox_amd.build import source_hash

    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("source_hash") != source_hash():
        return None
    return t.get(workload, {}).get("bytes_per_launch")


def host_info() -> dict:
    """CPU model and core counts of the host the baseline runs on."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu": model, "nproc": os.cpu_count(), "usable_cores": usable}


class pinned_core:
    """Pins this process to one core (the first it may use) for the block:
    the reference path is serial, and one core makes the figure repeatable."""

    def __enter__(self):
        try:
            self.prev = os.sched_getaffinity(0)
            self.core = min(self.prev)
            os.sched_setaffinity(0, {self.core})
        except (AttributeError, OSError):
            self.prev, self.core = None, None
        return self

    def __exit__(self, *exc):
        if self.prev is not None:
            os.sched_setaffinity(0, self.prev)


def cpu_baseline(workload: str, arrays: dict):
    """The reference path on one pinned host core over the whole workload
    (the same 8192 x 8192 inputs the GPU ran), with the iteration count of
    the restatement (validated against the reference: same sweeps)."""
    import oracle

    kind = "reference" if oracle.ref() is not None else "port"
    dem = np.ascontiguousarray(arrays["dem"])
    cells = dem.size
    with pinned_core() as pin:
        if workload == "fill":
            bc = np.ascontiguousarray(arrays["bc"])
            t0 = time.perf_counter()
            if kind == "reference":
                oracle.ref_fillsinks(dem, bc)
            else:
                oracle.fillsinks(dem, bc)
            dt = time.perf_counter() - t0
            _, iters, conv = oracle.fillsinks(dem, bc)
            sample = (f"fillsinks on the whole {dem.shape[1]}x{dem.shape[0]} workload, one call; "
                      f"reconstruct iterations {iters} (cap 1000, converged {conv})")
        else:
            slopes = np.ascontiguousarray(arrays["slopes"])
            t0 = time.perf_counter()
            if kind == "reference":
                oracle.ref_fsm2d(dem, slopes, 30.0)
            else:
                oracle.fsm2d_f32(dem, slopes, 30.0)
            dt = time.perf_counter() - t0
            _, iters = oracle.fsm2d_f32(dem, slopes, 30.0)
            sample = (f"excesstopography_fsm2d on the whole {dem.shape[1]}x{dem.shape[0]} "
                      f"workload, one call; {iters} iterations of the 4 sweeps")
        core = pin.core
    return {"value": round(cells / dt / 1e9, 5), "unit": "Gcells/s", "cores": 1, "kind": kind,
            "sample": sample, "seconds": round(dt, 2), "pinned_core": core, **host_info()}


def batched_cpu_baseline(op: str, dems, aux, workers: int):
    """The reference's forward path over all tiles of the batch, one tile
    per task on `workers` host threads (ctypes calls release the GIL). The
    reference has no adjoint, so this is forward only."""
    from concurrent.futures import ThreadPoolExecutor

    import oracle

    if oracle.ref() is None:
        return None
    if op == "fillsinks":
        fn = lambda k: oracle.ref_fillsinks(dems[k], aux[k])  # noqa: E731
    else:
        fn = lambda k: oracle.ref_fsm2d(dems[k], aux[k], 30.0)  # noqa: E731
    cells = sum(d.size for d in dems)
    t0 = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(fn, range(len(dems))))
    dt = time.perf_counter() - t0
    return {"value": round(cells / dt / 1e9, 5), "unit": "Gcells/s", "cores": workers,
            "kind": "reference", "seconds": round(dt, 2),
            "sample": f"forward only (the reference has no adjoint) over all {len(dems)} "
                      f"tiles, one tile per task on {workers} threads", **host_info()}


# kernels of one call of each workload (name fragments): the persistent
# relaxation and, for the fill, the pre-pass that seeds its queue
TRAFFIC_KERNELS = {"fill": ("fill_prio_persistent", "fill_persistent", "prio_seed_kernel",
                            "prio_queue_kernel", "zero_words"),
                   "fsm": ("fsm_persistent",), "copy": ("copyBuffer",)}


def _pmc_pass(counter: str, tag: str, bench_args: list) -> dict:
    """One rocprofv3 counter pass over a child `python bench.py ...` (the
    parent never touches the GPU); returns {kernel name: [KiB per
    dispatch]} from the CSV it writes under gpurun_out/."""
    import csv
    import glob
    import subprocess

    out = os.path.join(ROOT, "gpurun_out", f"pmc_{tag}_{counter}")
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "p", "--",
           sys.executable, os.path.join(ROOT, "bench.py"), *bench_args]
    print("[traffic] " + " ".join(cmd), file=sys.stderr, flush=True)
    env = dict(os.environ, TMPDIR="/tmp")
    res = subprocess.run(cmd, env=env, timeout=240, capture_output=True, text=True)
    if res.returncode != 0 and not glob.glob(os.path.join(out, "**", "*counter_collection.csv"),
                                             recursive=True):
        raise RuntimeError(f"rocprofv3 pass failed ({res.returncode}): {res.stderr[-2000:]}")
    vals = {}
    for path in glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                vals.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
    return vals


def measure_traffic(n: int):
    """HBM bytes per launch of the fill and excess-topography kernels at
    n x n: FETCH_SIZE and WRITE_SIZE in passes of their own, scaled by what
    the same counters read for a plain 256 MiB device copy."""
    import subprocess

    from libtopotoolbox_amd.build import source_hash

    def pick(vals, key):
        # mean KiB per dispatch of each kernel of the call, by kernel
        got = {}
        for k, v in vals.items():
            for frag in TRAFFIC_KERNELS[key]:
                if frag in k:
                    got.setdefault(frag, []).extend(v)
                    break
        if not got:
            raise RuntimeError(f"no {TRAFFIC_KERNELS[key]} dispatch in the counter output")
        return {frag: float(np.mean(v)) for frag, v in got.items()}

    copy_kib = 256 * 1024
    cal = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        vals = _pmc_pass(counter, "copy", ["--copy-calibration", "256"])
        cal[counter] = pick(vals, "copy")["copyBuffer"]
    scale = {c: copy_kib / v for c, v in cal.items()}
    result = {"calibration": {"copy_kib": copy_kib, "fetch_kib": round(cal["FETCH_SIZE"], 1),
                              "write_kib": round(cal["WRITE_SIZE"], 1),
                              "fetch_scale": round(scale["FETCH_SIZE"], 4),
                              "write_scale": round(scale["WRITE_SIZE"], 4)},
              "source_hash": source_hash(), "size": n}
    try:
        result["commit"] = subprocess.run(["git", "rev-parse", "--short=12", "HEAD"], cwd=ROOT,
                                          capture_output=True, text=True).stdout.strip() or None
    except OSError:
        result["commit"] = None
    for wl in ("fill", "fsm"):
        args = ["--workload", wl, "--size", str(n), "--steps", "3", "--warmup", "0", "--no-cpu",
                "--no-extra", "--no-adjoint", "--no-batched", "--no-stencils"]
        raw = {c: pick(_pmc_pass(c, wl, args), wl) for c in ("FETCH_SIZE", "WRITE_SIZE")}
        rd = sum(raw["FETCH_SIZE"].values()) * 1024 * scale["FETCH_SIZE"]
        wr = sum(raw["WRITE_SIZE"].values()) * 1024 * scale["WRITE_SIZE"]
        algo = BYTES_PER_CELL[wl] * n * n
        result[wl] = {"kernels": sorted(raw["FETCH_SIZE"]),
                      "fetch_kib_raw": {k: round(x, 1) for k, x in raw["FETCH_SIZE"].items()},
                      "write_kib_raw": {k: round(x, 1) for k, x in raw["WRITE_SIZE"].items()},
                      "read_bytes": int(rd), "write_bytes": int(wr),
                      "bytes_per_launch": int(rd + wr), "algorithmic_bytes": algo,
                      "traffic_over_algorithmic": round((rd + wr) / algo, 2)}
    print(json.dumps(result), flush=True)
    write_traffic_report(result)


def write_traffic_report(result: dict):
    """profiles/traffic.json and profiles/hbm_traffic_8192.md from one
    measure_traffic result (also from its printed JSON line, with
    `--traffic-report FILE`, where the GPU run's files do not come back)."""
    n = result["size"]
    cal = {"FETCH_SIZE": result["calibration"]["fetch_kib"],
           "WRITE_SIZE": result["calibration"]["write_kib"]}
    scale = {"FETCH_SIZE": result["calibration"]["fetch_scale"],
             "WRITE_SIZE": result["calibration"]["write_scale"]}
    copy_kib = result["calibration"]["copy_kib"]
    with open(os.path.join(ROOT, "profiles", "traffic.json"), "w") as f:
        json.dump(result, f, indent=1)
    lines = [f"# HBM traffic per call: fillsinks and excesstopography_fsm2d kernels, {n}x{n}",
             "",
             "Measured by `python bench.py --traffic` on one MI355X (kernel sources "
             f"{result['source_hash']}"
             + (f", commit {result['commit']}" if result.get("commit") else "") +
             "). It runs each counter in a rocprofv3 pass of its own over a child process:",
             "", "```"]
    lines += ["rocprofv3 --pmc {FETCH_SIZE|WRITE_SIZE} --output-format csv -d gpurun_out/pmc_copy_<C> "
              "-o p -- python bench.py --copy-calibration 256",
              "rocprofv3 --pmc {FETCH_SIZE|WRITE_SIZE} --output-format csv -d gpurun_out/pmc_<wl>_<C> "
              f"-o p -- python bench.py --workload <fill|fsm> --size {n} --steps 3 --warmup 0 "
              "--no-cpu --no-extra --no-adjoint --no-batched --no-stencils", "```", "",
              "## Calibration", "",
              f"A plain 256 MiB device copy reads FETCH_SIZE = {cal['FETCH_SIZE']:.1f} KiB and "
              f"WRITE_SIZE = {cal['WRITE_SIZE']:.1f} KiB per dispatch against {copy_kib} KiB moved "
              f"each way: scale factors {scale['FETCH_SIZE']:.4f} and {scale['WRITE_SIZE']:.4f}.",
              "", "## Per call (mean KiB per dispatch of each kernel, summed over the kernels "
              "of one call, scaled)", "",
              "| workload | kernel | FETCH_SIZE raw KiB | WRITE_SIZE raw KiB |",
              "|---|---|---:|---:|"]
    for wl in ("fill", "fsm"):
        r = result[wl]
        for k in r["kernels"]:
            lines.append(f"| {wl} | `{k}` | {r['fetch_kib_raw'][k]:.0f} | "
                         f"{r['write_kib_raw'].get(k, 0.0):.0f} |")
    lines += ["", "| workload | HBM read | HBM write | total | algorithmic | ratio |",
              "|---|---:|---:|---:|---:|---:|"]
    for wl in ("fill", "fsm"):
        r = result[wl]
        lines.append(f"| {wl} | {r['read_bytes'] / 1e9:.2f} GB | "
                     f"{r['write_bytes'] / 1e9:.2f} GB | {r['bytes_per_launch'] / 1e9:.2f} GB | "
                     f"{r['algorithmic_bytes'] / 1e9:.3f} GB ({BYTES_PER_CELL[wl]} B/cell) | "
                     f"{r['traffic_over_algorithmic']}x |")
    lines += ["", "`bench.py` reports `bytes_per_launch` as `roofline.traffic` only while the "
              "library's source stamp equals the one above."]
    with open(os.path.join(ROOT, "profiles", "hbm_traffic_8192.md"), "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["fill", "fsm"], default="fill")
    ap.add_argument("--size", type=int, default=N)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-stencils", action="store_true")
    ap.add_argument("--stencils-only", action="store_true",
                    help="only time the stencils next to the path (identifyflats, value "
                         "filters, gwdt_computecosts) and print their JSON")
    ap.add_argument("--batch-tiles", type=int, default=512)
    ap.add_argument("--traffic", action="store_true",
                    help="measure HBM bytes per launch of the fill and excess-topography "
                         "kernels with rocprofv3 counter passes (child processes) and write "
                         "profiles/traffic.json and profiles/hbm_traffic_8192.md")
    ap.add_argument("--traffic-report", metavar="FILE",
                    help="rewrite profiles/traffic.json and profiles/hbm_traffic_8192.md from "
                         "the JSON line a `--traffic` run printed (no GPU)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB, 5 times (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.traffic:
        measure_traffic(args.size)
        return
    if args.traffic_report:
        with open(args.traffic_report) as f:
            rows = [ln for ln in f if ln.startswith("{") and '"calibration"' in ln]
        if not rows:
            sys.exit(f"no measure_traffic JSON line in {args.traffic_report}")
        write_traffic_report(json.loads(rows[-1]))
        return
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)
    n = args.size
    cells = n * n
    other = "fsm" if args.workload == "fill" else "fill"
    if args.stencils_only:
        print(json.dumps({"stencils": stencil_run(n, max(5, args.steps), not args.no_cpu)}),
              flush=True)
        return

    t0 = time.perf_counter()
    arrays = make_inputs(args.workload, n)
    print(f"[bench] inputs for {args.workload} {n}x{n} in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    r = gpu_run(args.workload, arrays, args.steps, args.warmup)
    ms = r["wall_s"] * 1e3 / args.steps
    launch_ms = float(np.mean(r["launch_ms"]))
    print(f"[bench] {args.workload}: {ms:.3f} ms/step, stats {r['stats']}", file=sys.stderr,
          flush=True)
    line = {
        "metric": METRIC,
        "value": round(cells / (ms * 1e-3) / 1e9, 4),
        "unit": "Gcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("fillsinks" if args.workload == "fill"
                                else "excesstopography_fsm2d") + f" {n}x{n} float32",
                   "generator": ("fractal H0.7 0-3000 m + 200 paraboloid depressions + "
                                 "2% NaN blob (bc=1, dem=-INF)" if args.workload == "fill"
                                 else "fractal H0.8 0-1000 m + 5% pits; 64 Voronoi slopes "
                                 "U(0.2,0.8); cellsize 30 m"),
                   "parallelism": "single GPU", "tile": [r["stats"]["tile_x"],
                                                         r["stats"]["tile_y"]]},
        "roofline": roofline(args.workload, cells, launch_ms,
                             measured_traffic(args.workload)),
        "kernel_stats": {k: r["stats"][k] for k in ("requeues", "visits", "local_iters",
                                                    "max_local", "grid")},
    }
    del r
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.workload, arrays)
        print(f"[bench] cpu baseline: {line['cpu_baseline']}", file=sys.stderr, flush=True)
    extra = {}
    # the host-pointer drop-in (reference signature): PCIe copies included
    from libtopotoolbox_amd import ops

    host_in = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
    fn = (ops.fillsinks_host if args.workload == "fill" else
          (lambda dem, slopes: ops.excesstopography_fsm2d_host(dem, slopes, 30.0)))
    fn(*host_in.values())
    t0 = time.perf_counter()
    fn(*host_in.values())
    hd_ms = (time.perf_counter() - t0) * 1e3
    extra["host_dropin"] = {"ms": round(hd_ms, 3),
                            "gcells_per_s": round(cells / (hd_ms * 1e-3) / 1e9, 4),
                            "note": "reference-signature call on host buffers, H2D + kernels "
                                    "+ D2H included"}
    del host_in
    if not args.no_adjoint:
        extra["fwd_adjoint"] = {args.workload: fwd_adjoint_run(args.workload, arrays,
                                                               max(3, args.steps // 2),
                                                               launch_ms)}
        print(f"[bench] fwd+adjoint {args.workload}: {extra['fwd_adjoint'][args.workload]}",
              file=sys.stderr, flush=True)
    if args.workload == "fill" and not args.no_extra:
        extra["reconstruct"] = reconstruct_run(arrays, max(3, args.steps // 2))
        print(f"[bench] reconstruct: {json.dumps(extra['reconstruct'])}", file=sys.stderr,
              flush=True)
    del arrays
    if not args.no_extra:
        a2 = make_inputs(other, n)
        r2 = gpu_run(other, a2, max(3, args.steps // 2), 1)
        ms2 = r2["wall_s"] * 1e3 / max(3, args.steps // 2)
        ex = {"workload": other, "ms_per_step": round(ms2, 4),
              "gcells_per_s": round(cells / (ms2 * 1e-3) / 1e9, 4),
              "roofline": roofline(other, cells, float(np.mean(r2["launch_ms"])),
                                   measured_traffic(other)),
              "kernel_stats": {k: r2["stats"][k] for k in ("requeues", "visits", "local_iters",
                                                           "max_local", "grid")}}
        if not args.no_cpu:
            ex["cpu_baseline"] = cpu_baseline(other, a2)
        extra[other] = ex
        if not args.no_adjoint:
            extra.setdefault("fwd_adjoint", {})[other] = fwd_adjoint_run(
                other, a2, max(3, args.steps // 2), float(np.mean(r2["launch_ms"])))
        del a2, r2
    if not args.no_stencils:
        extra["stencils"] = stencil_run(n, max(5, args.steps), not args.no_cpu)
        print(f"[bench] stencils: {json.dumps(extra['stencils'])}", file=sys.stderr, flush=True)
        extra["fmm3d"] = fmm3d_run(n, max(3, args.steps // 2), not args.no_cpu)
        print(f"[bench] fmm3d: {json.dumps(extra['fmm3d'])}", file=sys.stderr, flush=True)
    if not args.no_batched:
        import torch

        torch.cuda.empty_cache()
        extra["batched"] = {}
        for op in ("fillsinks", "excesstopography"):
            extra["batched"][op] = batched_run(op, args.batch_tiles, max(3, args.steps // 4),
                                               cpu=not args.no_cpu)
            print(f"[bench] batched {op}: {json.dumps(extra['batched'][op])}", file=sys.stderr,
                  flush=True)
    if extra:
        line["extra"] = extra
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
