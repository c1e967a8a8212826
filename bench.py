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
passes, the batched launch and the stencils next to the path
(identifyflats, min / max filters).
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
# forward + adjoint: fill 9 + spill 4 + (spill, g_out, g_dem) 12; fsm 12 +
# (u, dem, slopes, g_out) 16 + (g_dem, g_slopes) 8
BYTES_FWD_ADJ = {"fill": 25, "fsm": 36}


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


def batched_run(op: str, count: int, steps: int):
    """One graph-captured persistent launch over all tiles (forward +
    adjoint, autotuned tile shape and sweep cap) against the same tiles run
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
        "autotune": {"tile": tune["tile"], "max_local": tune["max_local"],
                     "table": tune["table"]},
        "graph_ms": round(ms, 3), "gcells_per_s": round(cells / (ms * 1e-3) / 1e9, 4),
        "forward_only_ms": round(fwd_ms, 3),
        "one_at_a_time_ms": round(seq_ms, 3), "one_at_a_time_forward_ms": round(seq_fwd_ms, 3),
        "one_at_a_time_gcells_per_s": round(cells / (seq_ms * 1e-3) / 1e9, 4),
        "speedup_vs_one_at_a_time": round(seq_ms / ms, 3),
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


def roofline(workload: str, cells: int, launch_ms: float, traffic):
    algo = BYTES_PER_CELL[workload] * cells
    achieved = algo / (launch_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic}


def measured_traffic(workload: str):
    """HBM bytes per launch from the committed rocprofv3 counter summary
    (profiles/traffic.json), or None when none matches this workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(workload, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(workload: str, arrays: dict, crop: int):
    """Reference path on one host core over the top-left crop x crop cells."""
    import oracle

    sub = {k: np.ascontiguousarray(v[:crop, :crop]) for k, v in arrays.items()}
    kind = "reference" if oracle.ref() is not None else "port"
    cells = crop * crop
    if workload == "fill":
        from libtopotoolbox_amd.synthetic import edge_bc

        bc = sub["bc"] | edge_bc(crop, crop)
        t0 = time.perf_counter()
        if kind == "reference":
            oracle.ref_fillsinks(sub["dem"], bc)
        else:
            oracle.fillsinks(sub["dem"], bc)
        dt = time.perf_counter() - t0
        sample = f"fillsinks on the top-left {crop}x{crop} crop (edge bc), one call"
    else:
        t0 = time.perf_counter()
        if kind == "reference":
            oracle.ref_fsm2d(sub["dem"], sub["slopes"], 30.0)
        else:
            oracle.fsm2d_f32(sub["dem"], sub["slopes"], 30.0)
        dt = time.perf_counter() - t0
        sample = f"excesstopography_fsm2d on the top-left {crop}x{crop} crop, one call"
    return {"value": round(cells / dt / 1e9, 5), "unit": "Gcells/s", "cores": 1, "kind": kind,
            "sample": sample, "seconds": round(dt, 2),
            "host": platform.processor() or platform.machine()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["fill", "fsm"], default="fill")
    ap.add_argument("--size", type=int, default=N)
    ap.add_argument("--cpu-crop", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-stencils", action="store_true")
    ap.add_argument("--stencils-only", action="store_true",
                    help="only time the stencils next to the path (identifyflats, value "
                         "filters, gwdt_computecosts) and print their JSON")
    ap.add_argument("--batch-tiles", type=int, default=512)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB, 5 times (the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
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
        line["cpu_baseline"] = cpu_baseline(args.workload, arrays, min(args.cpu_crop, n))
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
            ex["cpu_baseline"] = cpu_baseline(other, a2, min(2048, n))
        extra[other] = ex
        if not args.no_adjoint:
            extra.setdefault("fwd_adjoint", {})[other] = fwd_adjoint_run(
                other, a2, max(3, args.steps // 2), float(np.mean(r2["launch_ms"])))
        del a2, r2
    if not args.no_stencils:
        extra["stencils"] = stencil_run(n, max(5, args.steps), not args.no_cpu)
        print(f"[bench] stencils: {json.dumps(extra['stencils'])}", file=sys.stderr, flush=True)
    if not args.no_batched:
        import torch

        torch.cuda.empty_cache()
        extra["batched"] = {}
        for op in ("fillsinks", "excesstopography"):
            extra["batched"][op] = batched_run(op, args.batch_tiles, max(3, args.steps // 4))
            print(f"[bench] batched {op}: {json.dumps(extra['batched'][op])}", file=sys.stderr,
                  flush=True)
    if extra:
        line["extra"] = extra
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
