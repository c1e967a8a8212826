"""Batched forward (+ adjoint) over many grids of uneven sizes.

A ``Batch`` packs grids ("items") of arbitrary shapes into flat device
buffers, plans them with ``tt_batch_plan`` and runs them with ONE persistent
launch (``tt_batch_run_dev``): the tiles of all items share one work queue
and every item walks through its forward and adjoint phases inside the same
launch. The launch allocates nothing and does not synchronise, so
``Batch.capture`` records it in a CUDA/HIP graph that ``replay`` reruns.

Results per item equal the single-grid entry points (fillsinks, spill
index and fillsinks adjoint bit for bit; excess topography and its adjoint
to FP32 rounding), so the per-item semantics are those of the reference
functions in include/topotoolbox.h (fillsinks :256-293,
excesstopography_fsm2d :536-609). The reference itself has no batched
entry point.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

OPS = {"fillsinks": _lib.BATCH_FILLSINKS, "excesstopography": _lib.BATCH_EXCESSTOPOGRAPHY}


def _stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


class Batch:
    """Packed buffers, plan and workspace of one batch.

    ``shapes``: list of (ny, nx). ``op``: "fillsinks" or "excesstopography".
    ``tile``: tile shape in 64x64 quadrants, (2, 1), (1, 2) or (1, 1).
    ``cellsize``: scalar or one per item (excess topography).
    Item offsets are aligned to ``align`` cells so every item starts on a
    256-byte boundary (128-bit loads).
    """

    def __init__(self, op: str, shapes, *, tile=(2, 1), cellsize=1.0, device="cuda",
                 align: int = 64):
        if op not in OPS:
            raise ValueError(f"op must be one of {sorted(OPS)}")
        if len(shapes) == 0:
            raise ValueError("a batch needs at least one item")
        self.op = op
        self.shapes = [(int(ny), int(nx)) for ny, nx in shapes]
        self.device = torch.device(device)
        n = len(self.shapes)
        cs = np.broadcast_to(np.asarray(cellsize, dtype=np.float32), (n,))
        offs, pos = [], 0
        for ny, nx in self.shapes:
            if ny <= 0 or nx <= 0:
                raise ValueError(f"item shape ({ny}, {nx}) must be positive")
            offs.append(pos)
            pos += -(-(ny * nx) // align) * align
        self.offsets = offs
        self.cells = pos
        self.items = (_lib.BatchItem * n)(
            *[_lib.BatchItem(nx, ny, o, float(c)) for (ny, nx), o, c in zip(self.shapes, offs, cs)])
        f32 = dict(dtype=torch.float32, device=self.device)
        self.dem = torch.zeros(self.cells, **f32)
        self.output = torch.zeros(self.cells, **f32)
        self.g_out = torch.zeros(self.cells, **f32)
        self.g_dem = torch.zeros(self.cells, **f32)
        if op == "fillsinks":
            self.bc = torch.zeros(self.cells, dtype=torch.uint8, device=self.device)
            self.spill = torch.zeros(self.cells, dtype=torch.int32, device=self.device)
            self.slopes = self.g_slopes = None
        else:
            self.slopes = torch.zeros(self.cells, **f32)
            self.g_slopes = torch.zeros(self.cells, **f32)
            self.bc = self.spill = None
        self._buffers = _lib.BatchBuffers(
            self.output.data_ptr(),
            self.spill.data_ptr() if self.spill is not None else None,
            self.dem.data_ptr(),
            self.bc.data_ptr() if self.bc is not None else None,
            self.slopes.data_ptr() if self.slopes is not None else None,
            self.g_out.data_ptr(), self.g_dem.data_ptr(),
            self.g_slopes.data_ptr() if self.g_slopes is not None else None)
        self.tile = None
        self.sweep_caps = None
        self.set_tile(tile)

    # ------------------------------------------------------------ planning
    def set_tile(self, tile) -> None:
        """(Re)plan for another tile shape; the packed buffers stay. The plan
        and workspace are new buffers: capture again after this."""
        qx, qy = int(tile[0]), int(tile[1])
        L = _lib.lib()
        pb, wb = ctypes.c_size_t(0), ctypes.c_size_t(0)
        n = len(self.shapes)
        _lib.check(L.tt_batch_plan(OPS[self.op], self.items, n, qx, qy, None, ctypes.byref(pb),
                                   ctypes.byref(wb)), "tt_batch_plan")
        host = ctypes.create_string_buffer(pb.value)
        _lib.check(L.tt_batch_plan(OPS[self.op], self.items, n, qx, qy, host, ctypes.byref(pb),
                                   ctypes.byref(wb)), "tt_batch_plan")
        self._plan_host = host
        self.plan_dev = None
        self.workspace = torch.empty(wb.value, dtype=torch.uint8, device=self.device)
        self.tile = (qx, qy)
        self.set_sweep_caps(self.sweep_caps)

    def set_sweep_caps(self, caps) -> None:
        """Per-item sweep caps (one per item, 0 = the cap given to ``run``;
        None = all 0). They change the schedule, never the result."""
        n = len(self.shapes)
        arr = np.zeros(n, np.uint32) if caps is None else np.asarray(caps, dtype=np.uint32)
        if arr.shape != (n,):
            raise ValueError(f"sweep caps: expected {n} values, got shape {arr.shape}")
        _lib.check(_lib.lib().tt_batch_plan_set_sweep_caps(
            self._plan_host, n, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
            "tt_batch_plan_set_sweep_caps")
        self.sweep_caps = None if caps is None else arr.copy()
        plan = torch.frombuffer(bytearray(self._plan_host.raw), dtype=torch.uint8)
        if self.plan_dev is None:
            self.plan_dev = plan.to(self.device)
        else:  # in place: a captured graph keeps reading this buffer
            self.plan_dev.copy_(plan)

    # ------------------------------------------------------------ data
    def views(self, flat: torch.Tensor) -> list:
        """2-D (ny, nx) views of every item in a packed buffer."""
        return [flat[o:o + ny * nx].view(ny, nx) for (ny, nx), o in zip(self.shapes,
                                                                        self.offsets)]

    def _fill(self, flat: torch.Tensor, arrays, name: str) -> None:
        if len(arrays) != len(self.shapes):
            raise ValueError(f"{name}: expected {len(self.shapes)} arrays, got {len(arrays)}")
        for v, a, s in zip(self.views(flat), arrays, self.shapes):
            t = torch.as_tensor(a)
            if tuple(t.shape) != s:
                raise ValueError(f"{name}: item of shape {tuple(t.shape)}, planned {s}")
            v.copy_(t.to(v.dtype), non_blocking=False)

    def load(self, dem, bc=None, slopes=None, g_out=None) -> None:
        """Copy per-item inputs (numpy arrays or tensors) into the packed
        buffers. bc: fillsinks; slopes: excess topography; g_out: adjoint
        seed (omit for forward-only runs)."""
        self._fill(self.dem, dem, "dem")
        if self.op == "fillsinks":
            if bc is None:
                raise ValueError("fillsinks needs bc")
            self._fill(self.bc, bc, "bc")
        else:
            if slopes is None:
                raise ValueError("excess topography needs slopes")
            self._fill(self.slopes, slopes, "slopes")
        if g_out is not None:
            self._fill(self.g_out, g_out, "g_out")

    # ------------------------------------------------------------ running
    def run(self, forward_only: bool = False, max_local: int = 0, stats: bool = False):
        """One launch on the current stream. With stats=True the call waits
        and returns the queue counters."""
        st = _lib.Stats() if stats else None
        flags = _lib.BATCH_FORWARD_ONLY if forward_only else 0
        with torch.cuda.device(self.device):
            rc = _lib.lib().tt_batch_run_dev(
                self._plan_host, self.plan_dev.data_ptr(), ctypes.byref(self._buffers),
                self.workspace.data_ptr(), self.workspace.numel(), int(max_local), flags,
                _stream_handle(), st)
        _lib.check(rc, "tt_batch_run_dev")
        return st.as_dict() if st is not None else None

    def capture(self, forward_only: bool = False, max_local: int = 0) -> "torch.cuda.CUDAGraph":
        """Record one launch in a graph (after an eager warm-up run that
        also caches the occupancy query). Replays reuse the plan, the
        workspace and the packed buffers: load new inputs with ``load``
        (or write the packed tensors) between replays."""
        self.run(forward_only, max_local)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.run(forward_only, max_local)
        torch.cuda.current_stream(self.device).wait_stream(s)
        return g

    # ------------------------------------------------------------ results
    def outputs(self) -> dict:
        """Per-item 2-D views of the results (after checking, with a device
        synchronisation, that no launch recorded a device error)."""
        with torch.cuda.device(self.output.device):
            _lib.check_device_error("batch")
        r = {"output": self.views(self.output), "g_dem": self.views(self.g_dem)}
        if self.op == "fillsinks":
            r["spill"] = self.views(self.spill)
        else:
            r["g_slopes"] = self.views(self.g_slopes)
        return r
