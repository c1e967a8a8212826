"""The device-error channel of asynchronous calls.

A host-set wait bound of one tick (10 ns) makes the persistent kernel's
bounded queue waits give up as soon as a workgroup has to wait for a tile:
the launch then ends early, with every workgroup exiting normally and the
timeout bit set (the GPU is never left spinning). A call without `stats`
has already returned 0 by then; the bit must reach the per-GPU sticky word
that tt_device_error reads, and a call with `stats` must return
1000 + bits itself.
"""
import numpy as np
import pytest
import torch

from conftest import assert_bitwise

pytestmark = pytest.mark.gpu


def _pits(oracle_lib, shape=(1024, 1024), seed=11):
    rng = np.random.default_rng(seed)
    dem = oracle_lib.grid_random(seed, shape)
    pits = rng.uniform(size=shape) < 0.05
    dem[pits] -= rng.uniform(0, 20, size=int(pits.sum())).astype(np.float32)
    bc = np.zeros(shape, np.uint8)
    bc[0, :] = bc[-1, :] = 1
    bc[:, 0] = bc[:, -1] = 1
    return dem, bc


@pytest.mark.parametrize("schedule", ["fifo", "priority"])
def test_expired_wait_reaches_sticky_word(cuda, oracle_lib, schedule):
    """Both work queues: the FIFO pop's slot wait and the lowest-level-first
    pop's wait for a queued tile give up at the bound."""
    import libtopotoolbox_amd as tt
    from libtopotoolbox_amd import _lib

    L = _lib.lib()
    dem, bc = _pits(oracle_lib)
    d = torch.from_numpy(dem).to(cuda)
    b = torch.from_numpy(bc).to(cuda)
    tt.set_tile_shape(2, 2)
    tt.set_fill_schedule(schedule)
    try:
        _expired_wait(tt, _lib, L, d, b, dem, bc, oracle_lib)
    finally:
        tt.set_fill_schedule("auto")


def _expired_wait(tt, _lib, L, d, b, dem, bc, oracle_lib):
    assert _lib.device_error(clear=True) in (0, 1)  # start clean
    print("forced timeout: fillsinks 1024^2, wait bound 1 tick", flush=True)
    L.tt_set_wait_limit(1)
    try:
        tt.fillsinks(d, b)  # no stats: returns before the kernel ends
        bits = _lib.device_error(clear=True)
        stats = _lib.Stats()
        rc = L.tt_fillsinks_dev(torch.empty_like(d).data_ptr(), d.data_ptr(), b.data_ptr(),
                                dem.shape[1], dem.shape[0],
                                torch.cuda.current_stream().cuda_stream, stats)
        msg = _lib.last_error()
        bits_stats = _lib.device_error(clear=True)
    finally:
        L.tt_set_wait_limit(0)
    assert bits & 1, "the expired wait did not reach the sticky word"
    assert rc == 1001 and stats.error & 1, (rc, stats.error, msg)
    assert "time limit" in msg
    assert bits_stats & 1
    # with the default bound the same calls are clean and exact
    got = tt.fillsinks(d, b).cpu().numpy()
    assert _lib.device_error(clear=True) == 0
    assert_bitwise(got, oracle_lib.fillsinks_pf(dem, bc), "after the error path")
    with pytest.raises(_lib.TTError):
        L.tt_set_wait_limit(1)
        try:
            tt.fillsinks(d, b)
            _lib.check_device_error("fillsinks")
        finally:
            L.tt_set_wait_limit(0)
    assert _lib.device_error(clear=True) == 0
