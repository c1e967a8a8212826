"""tt_batch_plan is host-only: its layout and error handling run here
without a GPU (packed-buffer bookkeeping of libtopotoolbox_amd.batched too)."""
import ctypes
import struct

import numpy as np
import pytest


def plan(op, shapes, tile=(2, 1), offsets=None):
    from libtopotoolbox_amd import _lib

    L = _lib.lib()
    offs = offsets if offsets is not None else np.cumsum([0] + [a * b for a, b in shapes])[:-1]
    items = (_lib.BatchItem * len(shapes))(
        *[_lib.BatchItem(nx, ny, int(o), 1.0) for (ny, nx), o in zip(shapes, offs)])
    pb, wb = ctypes.c_size_t(0), ctypes.c_size_t(0)
    rc = L.tt_batch_plan(op, items, len(shapes), tile[0], tile[1], None, ctypes.byref(pb),
                         ctypes.byref(wb))
    if rc != 0:
        return rc, _lib.last_error(), None, None
    buf = ctypes.create_string_buffer(pb.value)
    rc = L.tt_batch_plan(op, items, len(shapes), tile[0], tile[1], buf, ctypes.byref(pb),
                         ctypes.byref(wb))
    return rc, buf.raw, pb.value, wb.value


@pytest.mark.parametrize("tile", [(2, 1), (1, 2), (1, 1)])
def test_plan_tiles_cover_every_item(tile):
    shapes = [(1, 1), (300, 1), (1, 300), (129, 257), (64, 128), (1000, 33)]
    rc, raw, pb, wb = plan(0, shapes, tile)
    assert rc == 0
    magic, op, nitems, ntiles = struct.unpack_from("<Iiii", raw, 0)
    assert op == 0 and nitems == len(shapes)
    tx, ty = 64 * tile[0], 64 * tile[1]
    per = [-(-nx // tx) * -(-ny // ty) for ny, nx in shapes]
    assert ntiles == sum(per)
    # the tile -> item table closes the plan: item ids in order, per-item counts
    table = np.frombuffer(raw[pb - 4 * ntiles:], dtype=np.uint32)
    np.testing.assert_array_equal(table, np.repeat(np.arange(len(shapes)), per))
    cells = sum(a * b for a, b in shapes)
    assert wb >= 8 * cells + 16 * ntiles
    rc2, _, _, wb_fsm = plan(1, shapes, tile)
    assert rc2 == 0 and wb_fsm >= 16 * cells


def test_plan_rejects_bad_input():
    rc, msg, _, _ = plan(0, [(10, 10), (10, 10)], offsets=[0, 50])
    assert rc != 0 and "overlap" in msg
    rc, msg, _, _ = plan(0, [(10, 10)], tile=(2, 2))
    assert rc != 0 and "tile shape" in msg
    rc, msg, _, _ = plan(7, [(10, 10)])
    assert rc != 0 and "op" in msg
    rc, msg, _, _ = plan(0, [(0, 10)])
    assert rc != 0 and "item 0" in msg


def test_batch_packing_on_host():
    import torch

    from libtopotoolbox_amd.batched import Batch

    shapes = [(3, 5), (7, 2), (1, 1), (65, 70)]
    b = Batch("fillsinks", shapes, device="cpu")
    assert all(o % 64 == 0 for o in b.offsets)
    dems = [np.arange(a * c, dtype=np.float32).reshape(a, c) + k for k, (a, c) in
            enumerate(shapes)]
    bcs = [np.ones(s, np.uint8) for s in shapes]
    b.load(dems, bc=bcs)
    for v, d in zip(b.views(b.dem), dems):
        np.testing.assert_array_equal(v.numpy(), d)
    with pytest.raises(ValueError):
        b.load(dems[::-1], bc=bcs)
    b.set_tile((1, 1))
    assert b.tile == (1, 1) and b.plan_dev.dtype == torch.uint8


def test_plan_sweep_caps_host_only():
    """Per-item sweep caps are written into the host plan (the item records
    change, nothing else); a plan of another item count or no caps is
    refused."""
    from libtopotoolbox_amd import _lib

    L = _lib.lib()
    shapes = [(64, 64), (300, 200), (17, 900)]
    rc, raw, pb, wb = plan(1, shapes)
    assert rc == 0
    buf = ctypes.create_string_buffer(raw, pb)
    caps = (ctypes.c_uint32 * 3)(7, 0, 123456)
    assert L.tt_batch_plan_set_sweep_caps(buf, 3, caps) == 0
    new = buf.raw[:pb]
    diff = [i for i in range(pb) if new[i] != raw[i]]
    assert diff, "caps left the plan unchanged"
    for v in (7, 123456):  # each non-zero cap is stored as a little-endian word
        assert struct.pack("<I", v) in new
    assert L.tt_batch_plan_set_sweep_caps(buf, 3, (ctypes.c_uint32 * 3)(0, 0, 0)) == 0
    assert buf.raw[:pb] == raw
    assert L.tt_batch_plan_set_sweep_caps(buf, 4, caps) != 0
    assert "not a plan of 4 items" in _lib.last_error()
    assert L.tt_batch_plan_set_sweep_caps(buf, 3, None) != 0
