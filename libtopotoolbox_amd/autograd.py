"""Differentiable fillsinks and excess topography (torch.autograd.Function
over the C ABI's adjoint entry points).

* ``FillSinks``: output[p] = dem[spill[p]], where the spill cell is carried
  through the reconstruction as the index half of a lexicographic
  (level, index) key; the backward scatters g_out onto spill cells with a
  deterministic fixed-point reduction. The gradient is the one-hot
  subgradient selected by that order wherever the fill is not
  differentiable (ties). ``bc`` is not differentiable.
* ``ExcessTopography``: the backward solves the adjoint of the converged
  Godunov system by reverse relaxation sweeps, giving gradients with
  respect to the DEM and to the threshold slopes.

Each backward ends by reading the GPU's device-error word (a device
synchronisation): a forward or backward launch whose bounded queue wait
expired raises ``TTError`` there instead of returning incomplete
gradients.
"""
from __future__ import annotations

import torch

from . import _lib, ops


class FillSinks(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dem: torch.Tensor, bc: torch.Tensor) -> torch.Tensor:
        out, spill = ops.fillsinks_spill(dem.contiguous(), bc.contiguous())
        ctx.save_for_backward(spill)
        ctx.mark_non_differentiable(spill)
        return out

    @staticmethod
    def backward(ctx, g_out: torch.Tensor):
        (spill,) = ctx.saved_tensors
        g = ops.fillsinks_backward(spill, g_out.contiguous().to(torch.float32))
        with torch.cuda.device(g.device):
            _lib.check_device_error("fillsinks forward/backward")
        return g, None


class ExcessTopography(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dem: torch.Tensor, threshold_slopes: torch.Tensor,
                cellsize: float) -> torch.Tensor:
        dem = dem.contiguous()
        threshold_slopes = threshold_slopes.contiguous()
        u = ops.excesstopography_fsm2d(dem, threshold_slopes, cellsize)
        ctx.save_for_backward(u, dem, threshold_slopes)
        ctx.cellsize = float(cellsize)
        return u

    @staticmethod
    def backward(ctx, g_out: torch.Tensor):
        u, dem, s = ctx.saved_tensors
        g_dem, g_s = ops.excesstopography_fsm2d_backward(
            u, dem, s, ctx.cellsize, g_out.contiguous().to(torch.float32))
        with torch.cuda.device(g_dem.device):
            _lib.check_device_error("excesstopography_fsm2d forward/backward")
        return g_dem, g_s, None


def fillsinks(dem: torch.Tensor, bc: torch.Tensor) -> torch.Tensor:
    """Differentiable fillsinks (gradient with respect to ``dem``)."""
    return FillSinks.apply(dem, bc)


def excesstopography_fsm2d(dem: torch.Tensor, threshold_slopes: torch.Tensor,
                           cellsize: float) -> torch.Tensor:
    """Differentiable excesstopography_fsm2d (gradients with respect to
    ``dem`` and ``threshold_slopes``)."""
    return ExcessTopography.apply(dem, threshold_slopes, cellsize)
