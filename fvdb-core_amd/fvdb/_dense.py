"""Sparse rows <-> dense tensors through ``fvdbk_dense_rows`` (``include/fvdbk.h``), with autograd.

Semantics of ``src/fvdb/detail/autograd/ReadIntoDense.cpp`` / ``ReadFromDense.cpp`` and the
voxel callbacks of ``src/fvdb/detail/ops/ReadIntoDense.cu:17-91`` / ``ReadFromDense.cu:15-89``:
voxel ijk of grid b pairs with dense cell ijk - origin[b] when that cell is inside the frame.
Writing fills a zero tensor; reading gives zeros for voxels outside.  Each is the other's
gradient.  Layouts: "xyzc" ([B, X, Y, Z, *]) and "czyx" ([B, C, Z, Y, X]); the kernel takes
the tensor's element strides, so neither layout is copied or permuted in memory.
"""
from __future__ import annotations

from typing import Tuple

import torch

from ._lib import ext


def _frame(dense: torch.Tensor, layout: str) -> torch.Tensor:
    """[B, X, Y, Z, C] view of a dense tensor in `layout` (trailing dims of xyzc flattened)."""
    if layout == "xyzc":
        return dense.reshape(tuple(dense.shape[:4]) + (-1,))
    return dense.permute(0, 4, 3, 2, 1)


def _jidx(grid) -> torch.Tensor:
    impl = grid._impl
    return impl.jidx if grid.grid_count > 1 else torch.empty(0, dtype=torch.int32, device=impl.ijk.device)


class DenseWriteFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rows: torch.Tensor, grid, origins: torch.Tensor, size: Tuple[int, int, int], layout: str):
        n = rows.shape[0]
        r2 = rows.reshape(n, -1).contiguous()
        c = r2.shape[1]
        b = grid.grid_count
        if layout == "xyzc":
            out = torch.zeros((b,) + tuple(size) + tuple(rows.shape[1:]), dtype=rows.dtype, device=rows.device)
        else:
            out = torch.zeros((b, c) + tuple(size)[::-1], dtype=rows.dtype, device=rows.device)
        ext.dense_rows(True, grid._impl.ijk, _jidx(grid), origins, _frame(out, layout), r2)
        ctx.grid, ctx.layout, ctx.rows_shape = grid, layout, rows.shape
        ctx.save_for_backward(origins)
        return out

    @staticmethod
    def backward(ctx, grad):
        (origins,) = ctx.saved_tensors
        n = ctx.rows_shape[0]
        f = _frame(grad, ctx.layout)
        g = torch.empty((n, f.shape[4]), dtype=grad.dtype, device=grad.device)
        ext.dense_rows(False, ctx.grid._impl.ijk, _jidx(ctx.grid), origins, f, g)
        return g.reshape(ctx.rows_shape), None, None, None, None


class DenseReadFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dense: torch.Tensor, grid, origins: torch.Tensor, layout: str):
        f = _frame(dense, layout)
        n = grid.total_voxels
        out = torch.empty((n, f.shape[4]), dtype=dense.dtype, device=dense.device)
        ext.dense_rows(False, grid._impl.ijk, _jidx(grid), origins, f, out)
        ctx.grid, ctx.layout, ctx.dense_shape = grid, layout, dense.shape
        ctx.save_for_backward(origins)
        return out.reshape((n,) + tuple(dense.shape[4:])) if layout == "xyzc" else out

    @staticmethod
    def backward(ctx, grad):
        (origins,) = ctx.saved_tensors
        gd = torch.zeros(ctx.dense_shape, dtype=grad.dtype, device=grad.device)
        n = grad.shape[0]
        ext.dense_rows(True, ctx.grid._impl.ijk, _jidx(ctx.grid), origins, _frame(gd, ctx.layout),
                       grad.reshape(n, -1).contiguous())
        return gd, None, None, None


def write_dense(rows: torch.Tensor, grid, origins: torch.Tensor, size, layout: str) -> torch.Tensor:
    return DenseWriteFunction.apply(rows, grid, origins.to(torch.int32).contiguous(), tuple(int(v) for v in size),
                                    layout)


def read_dense(dense: torch.Tensor, grid, origins: torch.Tensor, layout: str) -> torch.Tensor:
    return DenseReadFunction.apply(dense, grid, origins.to(torch.int32).contiguous(), layout)
