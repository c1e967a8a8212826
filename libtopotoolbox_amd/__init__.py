"""libtopotoolbox_amd: MI355X-native sink filling and excess topography.

Drop-in for libtopotoolbox's fillsinks / reconstruct / excesstopography_fsm2d
hot path and the steps next to it (identifyflats, gwdt_computecosts, min/max
filters). The compute runs in HIP kernels for gfx950 behind a C ABI
(include/topotoolbox_amd.h); this package is the host-side mirror of that
interface for torch tensors (ops), the differentiable wrappers (autograd)
and the batched tile scheduler (batched).
"""
from . import _lib
from .ops import (
    excesstopography_fmm2d_host,
    excesstopography_fmm3d,
    excesstopography_fmm3d_host,
    excesstopography_fsm2d,
    excesstopography_fsm2d_host,
    fillsinks,
    fillsinks_capped,
    fillsinks_host,
    fillsinks_hybrid_host,
    gwdt_computecosts,
    gwdt_computecosts_host,
    identifyflats,
    identifyflats_host,
    max_filter,
    max_filter_host,
    max_filter_square,
    max_filter_square_host,
    min_filter,
    min_filter_host,
    min_filter_square,
    min_filter_square_host,
    reconstruct,
    reconstruct_capped,
    reconstruct_host,
    reconstruct_hybrid_host,
    set_tile_shape,
    set_fill_schedule,
)

__all__ = [
    "excesstopography_fmm2d_host",
    "excesstopography_fmm3d",
    "excesstopography_fmm3d_host",
    "excesstopography_fsm2d",
    "excesstopography_fsm2d_host",
    "fillsinks",
    "fillsinks_capped",
    "fillsinks_host",
    "fillsinks_hybrid_host",
    "gwdt_computecosts",
    "gwdt_computecosts_host",
    "identifyflats",
    "identifyflats_host",
    "max_filter",
    "max_filter_host",
    "max_filter_square",
    "max_filter_square_host",
    "min_filter",
    "min_filter_host",
    "min_filter_square",
    "min_filter_square_host",
    "reconstruct",
    "reconstruct_capped",
    "reconstruct_host",
    "reconstruct_hybrid_host",
    "set_tile_shape",
    "set_fill_schedule",
]


def has_topotoolbox() -> int:
    return _lib.lib().has_topotoolbox()
