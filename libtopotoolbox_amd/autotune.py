"""Autotuner for the batched launch: tile shape x per-visit sweep cap.

The two knobs change the schedule, never the result:

* tile shape (64x64 quadrants per tile: (2, 1), (1, 2), (1, 1)): larger
  tiles carry information further per visit but leave fewer tiles to
  spread over the CUs when the items are small;
* ``max_local``, the cap on sweeps of one tile visit: a capped visit writes
  its tile back and queues itself again, so neighbours see progress sooner
  (useful on long drainage paths), at the cost of more visits.

``tune`` times every candidate on the caller's loaded ``Batch`` with HIP
events (median of ``reps`` launches after a warm-up) and returns the
fastest. Results are cached in a JSON file keyed by op, the item shapes,
forward-only and the device name, so a serving loop tunes once.
"""
from __future__ import annotations

import hashlib
import json
import os

import torch

TILES = [(2, 1), (1, 2), (1, 1)]
MAX_LOCAL = [0, 64, 16, 4]  # 0 = no practical cap (65536)


def cache_path() -> str:
    return os.environ.get("TT_AMD_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "libtopotoolbox_amd",
                                       "autotune.json"))


def key(batch, forward_only: bool) -> str:
    h = hashlib.sha1(repr(batch.shapes).encode()).hexdigest()[:16]
    dev = torch.cuda.get_device_name(batch.device)
    return f"{batch.op}|{h}|n{len(batch.shapes)}|fwd{int(forward_only)}|{dev}"


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


def tune(batch, forward_only: bool = False, reps: int = 3, tiles=None, max_locals=None,
         use_cache: bool = True) -> dict:
    """Pick the fastest (tile, max_local) for this batch, leave the batch
    planned for it and return {"tile", "max_local", "ms", "table"}."""
    k = key(batch, forward_only)
    if use_cache:
        hit = _load_cache().get(k)
        if hit is not None:
            batch.set_tile(tuple(hit["tile"]))
            return dict(hit, cached=True)
    table = []
    for tile in (tiles or TILES):
        batch.set_tile(tile)
        for ml in (max_locals or MAX_LOCAL):
            ms = time_launch(batch, forward_only, ml, reps)
            table.append({"tile": list(tile), "max_local": ml, "ms": round(ms, 4)})
    best = min(table, key=lambda r: r["ms"])
    batch.set_tile(tuple(best["tile"]))
    res = {"tile": best["tile"], "max_local": best["max_local"], "ms": best["ms"],
           "table": table}
    if use_cache:
        _store(k, res)
    return dict(res, cached=False)
