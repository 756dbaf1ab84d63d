"""Pooling, nearest upsampling and injection between grids: the operators beside the
sparse convolution in a U-Net.

* max / average pooling: a window map (``fvdbk_kmap_window``: coarse voxel c reads the
  fine voxels ``c*stride + [0, pool_factor)``) reduced by ``fvdbk_pool_max`` /
  ``fvdbk_pool_avg``.  Forward semantics of ``src/fvdb/detail/ops/DownsampleGridMaxPool.cu:17-63``
  (max over existing voxels, -inf when none) and ``DownsampleGridAvgPool.cu:17-62``
  (window sum times ``1 / prod(pool_factor)``, including missing voxels in the divisor).
  Backward: max pooling routes each coarse gradient to the first voxel holding the
  maximum (``:65-114``); average pooling gives every existing voxel ``grad * factor``
  (``DownsampleGridAvgPool.cu:64-107``).  Where windows overlap (stride < pool factor)
  the reference assigns these gradients concurrently (the last writer wins); here they
  are summed.
* nearest upsampling (``UpsampleGridNearest.cu:17-96``): a fine voxel takes the value of
  the coarse voxel ``floor(ijk / factor)`` (zero when that voxel does not exist); the
  backward sums fine gradients into their coarse voxel.  The coarse coordinate is an
  integer floor division; the reference divides in the data's floating-point type, which
  agrees wherever that type represents the coordinates exactly.
* injection (``autograd/Inject.cpp``, ``ops/Inject.cu``): voxels of the destination grid
  that exist in the source grid take the source value; the backward assigns the
  destination gradient of those voxels to the source voxels and zeroes it in the
  destination gradient.
"""
from __future__ import annotations

import torch

from ._lib import ext


def _pad128(n: int) -> int:
    return (n + 127) // 128 * 128


def window_map(coarse, fine, pool_factor, stride) -> torch.Tensor:
    """[ceil(N_coarse/128)*128, prod(pool_factor)] int32 indices of the fine voxels in each window."""
    return ext.window_map(coarse._impl.tensors(), fine._impl.tensors(), coarse.grid_count, list(pool_factor),
                          list(stride), _pad128(coarse.total_voxels))


def _flat(data: torch.Tensor) -> torch.Tensor:
    return data.reshape(data.shape[0], -1).contiguous()


class MaxPoolFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, data, wmap, n_coarse):
        x = _flat(data)
        # the argmax (the voxel each coarse gradient goes to) only when a backward will run
        out, arg = ext.pool_max(x, wmap, n_coarse, bool(ctx.needs_input_grad[0]))
        ctx.save_for_backward(arg)
        ctx.in_shape = data.shape
        return out.reshape((n_coarse,) + tuple(data.shape[1:]))

    @staticmethod
    def backward(ctx, grad):
        (arg,) = ctx.saved_tensors
        g = _flat(grad)
        C = g.shape[1]
        gi = torch.zeros(ctx.in_shape[0] * C, dtype=g.dtype, device=g.device)
        valid = arg >= 0
        flat = (arg.long() * C + torch.arange(C, device=g.device))[valid]
        gi.index_add_(0, flat, g[valid])
        return gi.reshape(ctx.in_shape), None, None


class AvgPoolFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, data, wmap, n_coarse, scale):
        x = _flat(data)
        out = ext.pool_avg(x, wmap, n_coarse, float(scale))
        ctx.save_for_backward(wmap)
        ctx.n_coarse, ctx.scale, ctx.in_shape = n_coarse, float(scale), data.shape
        return out.reshape((n_coarse,) + tuple(data.shape[1:]))

    @staticmethod
    def backward(ctx, grad):
        (wmap,) = ctx.saved_tensors
        g = _flat(grad)
        # grad * factor in the data type, as the reference multiplies two values of it
        gs = g * torch.tensor(ctx.scale, dtype=g.dtype, device=g.device)
        m = wmap[: ctx.n_coarse]
        rows, taps = (m >= 0).nonzero(as_tuple=True)
        gi = torch.zeros((ctx.in_shape[0], g.shape[1]), dtype=g.dtype, device=g.device)
        gi.index_add_(0, m[rows, taps].long(), gs[rows])
        return gi.reshape(ctx.in_shape), None, None, None


class GatherRowsFunction(torch.autograd.Function):
    """out[i] = data[idx[i]] where idx >= 0, else zero; the backward sums into data."""

    @staticmethod
    def forward(ctx, data, idx):
        out = ext.gather_rows(data, idx)  # one pass: rows of missing sources are zero
        ctx.save_for_backward(idx)
        ctx.n_src = data.shape[0]
        return out

    @staticmethod
    def backward(ctx, grad):
        (idx,) = ctx.saved_tensors
        valid = idx >= 0
        gd = torch.zeros((ctx.n_src,) + tuple(grad.shape[1:]), dtype=grad.dtype, device=grad.device)
        gd.index_add_(0, idx[valid].long(), grad[valid])
        return gd, None


class InjectFunction(torch.autograd.Function):
    """dst[i] = src[idx[i]] where idx >= 0 (idx: source index of each destination voxel)."""

    @staticmethod
    def forward(ctx, src, dst, idx):
        if dst.dtype == src.dtype:
            out = ext.gather_rows(src, idx, dst)  # one pass: destination rows without a source keep dst
        else:
            valid = idx >= 0
            out = dst.clone()
            out[valid] = src[idx[valid].long()]
        ctx.save_for_backward(idx)
        ctx.n_src = src.shape[0]
        return out

    @staticmethod
    def backward(ctx, grad):
        (idx,) = ctx.saved_tensors
        valid = idx >= 0
        gs = torch.zeros((ctx.n_src,) + tuple(grad.shape[1:]), dtype=grad.dtype, device=grad.device)
        gs[idx[valid].long()] = grad[valid]
        gd = grad.clone()
        gd[valid] = 0
        return gs, gd, None
