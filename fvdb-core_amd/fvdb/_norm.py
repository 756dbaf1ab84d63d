"""Group normalization over the voxels of each grid, on the gfx950 kernels of
``fvdbk_group_norm_forward`` / ``_backward`` (``include/fvdbk.h``).

Semantics of the reference's ``fvdb.nn.GroupNorm`` (``fvdb/nn/modules.py:407-432``): for each
grid, ``torch.nn.GroupNorm`` on the grid's features viewed as [1, C, n] -- statistics per
(grid, group) over the grid's voxels and the group's channels, biased variance, per-channel
affine.  Here the batch is normalized in two native passes (statistics, then normalize +
affine, optionally + ReLU) instead of a loop over grids with transposed copies, with
deterministic reductions.
"""
from __future__ import annotations

from typing import Optional

import torch

from ._lib import ext


class GroupNormFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, weight, bias, grid, num_groups: int, eps: float, relu: bool):
        impl = grid._impl
        jidx = impl.jidx if grid.grid_count > 1 else torch.empty(0, dtype=torch.int32, device=x.device)
        y, mean, rstd = ext.group_norm_forward(x, impl.joffsets, impl.counts, jidx, int(num_groups), weight, bias,
                                               float(eps), bool(relu))
        ctx.save_for_backward(x, weight, mean, rstd, y if relu else torch.empty(0, device=x.device))
        ctx.grid, ctx.groups = grid, int(num_groups)
        ctx.has_w, ctx.has_b = weight is not None, bias is not None
        return y

    @staticmethod
    def backward(ctx, dy):
        x, weight, mean, rstd, y = ctx.saved_tensors
        impl = ctx.grid._impl
        jidx = impl.jidx if ctx.grid.grid_count > 1 else torch.empty(0, dtype=torch.int32, device=x.device)
        dx, dg, db = ext.group_norm_backward(x, dy.contiguous(), y, impl.joffsets, impl.counts, jidx, ctx.groups,
                                             mean, rstd, weight if ctx.has_w else None)
        gw = dg.to(weight.dtype) if ctx.has_w and ctx.needs_input_grad[1] else None
        gb = db.to(weight.dtype if ctx.has_w else x.dtype) if ctx.has_b and ctx.needs_input_grad[2] else None
        return dx, gw, gb, None, None, None, None


def group_norm(x: torch.Tensor, grid, num_groups: int, weight: Optional[torch.Tensor] = None,
               bias: Optional[torch.Tensor] = None, eps: float = 1e-5, relu: bool = False) -> torch.Tensor:
    """GroupNorm of `x` [N, C] (rows = the voxels of `grid`), per grid; ReLU fused when asked."""
    if x.dim() != 2 or x.shape[0] != grid.total_voxels:
        raise ValueError("features must have shape [total_voxels, C]")
    if x.shape[1] % num_groups != 0:
        raise ValueError("num_channels must be divisible by num_groups")
    return GroupNormFunction.apply(x, weight, bias, grid, num_groups, eps, relu)
