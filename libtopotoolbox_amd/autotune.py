"""Autotuner for the batched launch: tile shape, and sweep cap per item size
class.

The knobs change the schedule, never the result:

* tile shape (64x64 quadrants per tile: (2, 1), (1, 2), (1, 1)), one per
  launch: larger tiles carry information further per visit but leave fewer
  tiles to spread over the CUs when the items are small;
* the sweep cap of one tile visit, per item (``Batch.set_sweep_caps``): a
  capped visit writes its tile back and queues itself again, so neighbours
  see progress sooner. That pays on large items, whose drainage paths cross
  many tiles; an item of one or a few tiles gains nothing from it and only
  pays the extra visits. The items are therefore grouped in size classes
  (``CLASS_EDGES``, by cell count) and each class gets its own cap.

``tune`` first times every (tile, global cap) pair, then, on the best tile,
walks the classes from the largest down and gives each the cap that times
fastest with the others fixed (coordinate descent). Times are medians of
``reps`` launches measured with HIP events after a warm-up. Finally the
best global setting and the per-class setting are timed again, interleaved,
and the per-class caps are kept only if they are faster. Results are cached
in a JSON file keyed by op, the item shapes, forward-only and the device
name, so a serving loop tunes once.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import torch

TILES = [(2, 1), (1, 2), (1, 1)]
SWEEP_CAPS = [0, 16, 4, 2, 1]  # 0 = no practical cap (65536)
CLASS_EDGES = (256 * 256, 1024 * 1024)  # cells: small < 65536 <= medium < 1048576 <= large
CLASS_NAMES = ("small", "medium", "large")


def cache_path() -> str:
    return os.environ.get("TT_AMD_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "libtopotoolbox_amd",
                                       "autotune.json"))


def key(batch, forward_only: bool) -> str:
    h = hashlib.sha1(repr(batch.shapes).encode()).hexdigest()[:16]
    dev = torch.cuda.get_device_name(batch.device)
    return f"{batch.op}|{h}|n{len(batch.shapes)}|fwd{int(forward_only)}|{dev}|v2"


def size_classes(shapes) -> np.ndarray:
    """Class index (0 small, 1 medium, 2 large) of every item."""
    cells = np.array([ny * nx for ny, nx in shapes], np.int64)
    return np.searchsorted(np.asarray(CLASS_EDGES), cells, side="right")


def item_caps(classes: np.ndarray, class_caps) -> np.ndarray:
    return np.asarray(class_caps, np.uint32)[classes]


def _load_cache() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(k: str, v: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        c = _load_cache()
        c[k] = v
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only costs a re-tune


def time_launch(batch, forward_only: bool, max_local: int, reps: int) -> float:
    """Median milliseconds of one launch (batch already loaded)."""
    batch.run(forward_only, max_local)  # warm-up
    stream = torch.cuda.current_stream(batch.device)
    times = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        batch.run(forward_only, max_local)
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def apply(batch, result: dict) -> None:
    """Plan ``batch`` for a tuning result (tile shape and per-item caps);
    run it with ``max_local=result["max_local"]``."""
    batch.set_tile(tuple(result["tile"]))
    caps = result.get("class_caps")
    batch.set_sweep_caps(None if caps is None else
                         item_caps(size_classes(batch.shapes), caps))


def tune(batch, forward_only: bool = False, reps: int = 3, tiles=None, caps=None,
         use_cache: bool = True) -> dict:
    """Tune ``batch`` (loaded), leave it planned for the result and return
    {"tile", "max_local", "class_caps" (None: one global cap), "ms",
    "global": best global setting, "gain_vs_global", "table"}."""
    k = key(batch, forward_only)
    if use_cache:
        hit = _load_cache().get(k)
        if hit is not None:
            apply(batch, hit)
            return dict(hit, cached=True)
    caps = list(caps or SWEEP_CAPS)
    classes = size_classes(batch.shapes)
    present = [c for c in (2, 1, 0) if (classes == c).any()]
    table = []

    # 1. tile shape x one cap for every item
    batch.set_sweep_caps(None)
    for tile in (tiles or TILES):
        batch.set_tile(tile)
        for ml in caps:
            ms = time_launch(batch, forward_only, ml, reps)
            table.append({"tile": list(tile), "max_local": ml, "ms": round(ms, 4)})
    best = min(table, key=lambda r: r["ms"])
    tile = tuple(best["tile"])
    batch.set_tile(tile)

    # 2. per-class caps on the best tile (coordinate descent, largest first)
    class_caps = [best["max_local"]] * 3
    cur = best["ms"]
    for c in present:
        for ml in caps:
            if ml == class_caps[c]:
                continue
            trial = list(class_caps)
            trial[c] = ml
            batch.set_sweep_caps(item_caps(classes, trial))
            ms = time_launch(batch, forward_only, 0, reps)
            table.append({"tile": list(tile), "class_caps": dict(zip(CLASS_NAMES, trial)),
                          "ms": round(ms, 4)})
            if ms < cur:
                cur, class_caps = ms, trial

    # 3. the best global setting against the per-class one, interleaved
    g_ms, p_ms = [], []
    for _ in range(3):
        batch.set_sweep_caps(None)
        g_ms.append(time_launch(batch, forward_only, best["max_local"], reps))
        batch.set_sweep_caps(item_caps(classes, class_caps))
        p_ms.append(time_launch(batch, forward_only, 0, reps))
    g, p = float(np.median(g_ms)), float(np.median(p_ms))
    res = {"tile": list(tile), "max_local": 0, "class_caps": class_caps, "ms": round(p, 4)}
    if p >= g:
        batch.set_sweep_caps(None)
        res = {"tile": list(tile), "max_local": best["max_local"], "class_caps": None,
               "ms": round(g, 4)}
    res.update({"global": {"tile": list(tile), "max_local": best["max_local"], "ms": round(g, 4)},
                "per_class_ms": round(p, 4), "gain_vs_global": round(g / min(p, g), 4),
                "classes": {n: int((classes == c).sum()) for c, n in enumerate(CLASS_NAMES)},
                "class_edges_cells": list(CLASS_EDGES), "table": table})
    if use_cache:
        _store(k, res)
    return dict(res, cached=False)
