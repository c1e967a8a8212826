"""Iteration cap and resume on the GPU (the reference's reconstruct stops
silently after 1000 iterations, src/reconstruct.c:273; the marker is its
whole state, so a reconstruction can continue from a partial one).

fillsinks_capped / reconstruct_capped stop after about max_visits tile
visits. A capped output must be a valid partial result, pinned by the
mathematics of the reconstruction rather than by the reference's own
capped marker (which another update order produces):
  * it lies between the full result and the initial marker;
  * reconstruct from it reaches the full result bit for bit (the full
    result is checked against the oracle's fillsinks);
  * an uncapped-enough run reports converged and equals the full result;
  * the number of visits stays within the cap plus one visit per
    workgroup (the pops that find their slot filled after the stop).
"""
import numpy as np
import pytest
import torch

from conftest import assert_bitwise

pytestmark = pytest.mark.gpu


def to_gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def cases():
    from libtopotoolbox_amd import synthetic as S

    out = []
    dem = S.dem_depressions(2048, 2048, 31)
    d, bc, _ = S.nan_to_bc(dem, S.edge_bc(2048, 2048))
    out.append(("depressions 2048^2, NaN as bc", d, bc))
    raw = S.dem_depressions(1500, 1100, 32)  # NaN blob left as NaN
    out.append(("depressions 1100x1500, raw NaN", raw, S.edge_bc(1500, 1100)))
    return out


def same(got, want, what):
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want), err_msg=f"{what}: NaN cells")
    ok = ~np.isnan(want)
    assert_bitwise(got[ok], want[ok], what)


@pytest.mark.parametrize("cap_per_tile", [1, 3])
def test_fillsinks_cap_and_resume(cuda, oracle_lib, cap_per_tile):
    from libtopotoolbox_amd import ops

    for name, dem, bc in cases():
        full, _, c = oracle_lib.fillsinks(dem, bc)
        assert c
        ntiles = -(-dem.shape[0] // 128) * -(-dem.shape[1] // 128)
        cap = cap_per_tile * ntiles
        (part, conv), st = ops.fillsinks_capped(to_gpu(dem), to_gpu(bc), cap, stats=True)
        assert st["error"] == 0
        assert not conv, f"{name}: converged within {cap} visits"
        assert ntiles <= st["visits"] <= cap + st["grid"], (name, st["visits"], cap)
        p = part.cpu().numpy()
        ok = ~np.isnan(dem)
        np.testing.assert_array_equal(np.isnan(p), ~ok, err_msg=name)
        assert (p[ok] >= full[ok]).all(), f"{name}: partial fill below the full fill"
        assert (p[ok] >= dem[ok]).all(), f"{name}: partial fill below the DEM"
        assert (p[ok] != full[ok]).any(), f"{name}: the capped run is already complete"
        # resume: reconstruct(-partial, -dem), negated back
        m = ops.reconstruct(-to_gpu(p), -to_gpu(dem))
        same((-m).cpu().numpy(), full, f"{name}: resumed fill")


def test_fillsinks_cap_not_reached(cuda, oracle_lib):
    from libtopotoolbox_amd import ops

    name, dem, bc = cases()[0]
    full, _, _ = oracle_lib.fillsinks(dem, bc)
    out, conv = ops.fillsinks_capped(to_gpu(dem), to_gpu(bc), 2**31)
    assert conv
    same(out.cpu().numpy(), full, name)


def test_reconstruct_cap_and_resume(cuda, oracle_lib):
    from libtopotoolbox_amd import ops

    name, dem, bc = cases()[0]
    full, _, _ = oracle_lib.fillsinks(dem, bc)
    marker0 = np.where(bc == 1, -dem, -np.inf).astype(np.float32)
    mask = (-dem).astype(np.float32)
    ntiles = (2048 // 128) ** 2
    (part, conv), st = ops.reconstruct_capped(to_gpu(marker0), to_gpu(mask), ntiles, stats=True)
    assert st["error"] == 0 and not conv
    p = part.cpu().numpy()
    assert (p >= marker0).all() and (p <= -full).all(), "partial marker out of its bounds"
    done, conv2 = ops.reconstruct_capped(part, to_gpu(mask), 2**31)
    assert conv2
    same(-done.cpu().numpy(), full, "resumed reconstruct")
    # resuming twice more changes nothing
    again = ops.reconstruct(done, to_gpu(mask))
    assert_bitwise(again.cpu().numpy(), done.cpu().numpy(), "reconstruct of a converged marker")
