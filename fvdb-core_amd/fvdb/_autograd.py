"""Autograd for the sparse convolution: forward, dgrad and wgrad on gfx950 kernels.

Replaces ``src/fvdb/detail/autograd/SparseConvolutionKernelMap.cpp:19-191`` and
``SparseConvolutionImplicitGEMM.cpp:17-226``.  Weight layouts follow the reference:
a forward weight ``[O, I, k0, k1, k2]`` is flattened with ``permute(4,3,2,1,0)`` to
``[K, I, O]`` (``SparseConvolutionKernelMap.cpp:75``), a transposed weight
``[I, O, k0, k1, k2]`` with ``permute(4,3,2,0,1)`` (``:92``), and the gradient is
mapped back with the inverse permutations (``:170-187``).

* forward: implicit GEMM over the kernel map (rows = output voxels);
* dgrad:   the same kernel over the opposite map with the un-transposed weights;
* wgrad:   split-K MFMA reduction over the neighbour-map pairs (each offset's 64-pair
  rounds interleaved over its partials) with an fp32 Neumaier-compensated sum of the
  partials.
* row order: rows grouped by tap mask (the pack's ``_order_for``) so that a tile skips
  the taps none of its rows uses; the result does not depend on the order.
* epilogue: an optional bias and ReLU applied to the fp32 accumulator before the
  output is rounded (the module's separate bias add, ``fvdb/nn/modules.py:400-403``,
  and a following ReLU, in one pass); the ReLU's gradient mask is read from the output.
"""
from __future__ import annotations

from typing import Optional

import torch

from ._lib import ext, require_device


def weight_to_kio(weight: torch.Tensor, transposed: bool) -> torch.Tensor:
    w = weight.permute(4, 3, 2, 1, 0) if not transposed else weight.permute(4, 3, 2, 0, 1)
    return w.reshape(-1, w.shape[3], w.shape[4]).contiguous()


def weight_to_koi(weight: torch.Tensor, transposed: bool) -> torch.Tensor:
    """The forward contraction's operand [K, O, I] (``weight_to_kio`` with the last two
    dimensions swapped) in one permute."""
    w = weight.permute(4, 3, 2, 0, 1) if not transposed else weight.permute(4, 3, 2, 1, 0)
    return w.reshape(-1, w.shape[3], w.shape[4]).contiguous()


def kio_to_weight(gk: torch.Tensor, kernel_size, transposed: bool) -> torch.Tensor:
    g = gk.reshape(int(kernel_size[2]), int(kernel_size[1]), int(kernel_size[0]), gk.shape[1], gk.shape[2])
    return g.permute(4, 3, 2, 1, 0) if not transposed else g.permute(3, 4, 2, 1, 0)


def _wgrad_tile(c: int) -> int:
    """Channels per workgroup tile along one side of the wgrad kernel (32, 64 or 128)."""
    return 32 if c <= 32 else (64 if c <= 64 else 128)


def wgrad_chunk(n_pairs: int, dtype: torch.dtype = torch.float32, cin: int = 128, cout: int = 128,
                n_cu: int = 0) -> int:
    """Pairs per split-K chunk, a multiple of 64 and at least 256: offset k gets
    ceil(pairs of k / chunk) partials, which take its 64-pair rounds interleaved.

    fp32 (MFMA-bound) runs best with ~2048 partials per launch (planes + spheres, 6.7M pairs,
    128 -> 128 on MI355X: 2.40 ms against 2.52 ms at 473 and 2.60 ms at 4040).  16-bit inputs
    are bound by their gathers, and run best when the launch is one wave of workgroups that all
    stay resident: 128 x 128 channel tiles hold two workgroups per CU, and a launch just above
    that pays a second, mostly idle wave (same layer: 473 partials 0.51 ms, 525 partials
    0.67 ms).  Narrower tiles measured best at 3 (64-channel) and 4 (32-channel) workgroups
    per CU.  Workgroups per partial = channel tiles of the layer."""
    if dtype in (torch.bfloat16, torch.float16):
        if n_cu <= 0:
            n_cu = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count \
                if torch.cuda.is_available() else 256
        tm, tn = _wgrad_tile(cin), _wgrad_tile(cout)
        tiles = -(-cin // tm) * -(-cout // tn)
        per_cu = 1.75 if (tm, tn) == (128, 128) else (3.0 if max(tm, tn) >= 64 else 4.0)
        target = max(1, int(per_cu * n_cu / tiles))
    else:
        target = 2048
    c = -(-max(n_pairs, 1) // target)
    c = -(-c // 64) * 64
    return int(max(256, c))


def check_conv_args(feats: torch.Tensor, weight: torch.Tensor, n_in: int, transposed: bool,
                    feats2: Optional[torch.Tensor] = None) -> None:
    """Input checks of ``autograd/SparseConvolutionKernelMap.cpp:34-93``; with `feats2` the
    input is the channel concatenation [feats | feats2] (same shape and dtype)."""
    if feats2 is not None and (feats2.shape != feats.shape or feats2.dtype != feats.dtype or
                               feats2.device != feats.device):
        raise ValueError(f"concatenated features must match the features: {tuple(feats2.shape)} {feats2.dtype} "
                         f"vs {tuple(feats.shape)} {feats.dtype}")
    if not feats.is_floating_point():
        raise TypeError("features must have a floating point type")
    if feats.dim() != 2:
        raise ValueError(f"Expected features to have 2 dimensions (shape (n, nF)) but got {feats.dim()} dimensions")
    if not weight.is_floating_point():
        raise TypeError("kernels must have a floating point type")
    for i in range(weight.dim()):
        if weight.shape[i] == 0:
            raise ValueError(f"kernels tensor has zero dimension (dim = {i})")
    if feats.shape[0] != n_in:
        raise ValueError("The number of input features must match the number of voxels")
    if weight.dim() != 5:
        raise ValueError(f"Expected kernels to have 5 dimensions but got {weight.dim()} dimensions")
    cin = weight.shape[0] if transposed else weight.shape[1]
    nf = feats.shape[1] * (2 if feats2 is not None else 1)
    if cin != nf:
        raise ValueError(f"Expected input channels of kernels ({cin}) to equal input channels of features: {nf}")
    if weight.dtype != feats.dtype:
        raise TypeError(f"features ({feats.dtype}) and kernels ({weight.dtype}) must have the same dtype")
    require_device(feats, "features")


class SparseConvFunction(torch.autograd.Function):
    """out = act(conv(feats, weight) + bias) with the bias and ReLU fused into the
    contraction's epilogue (one rounding to the feature dtype); bias may be None."""

    @staticmethod
    def forward(ctx, feats, weight, bias, pack, transposed: bool, relu: bool = False, feats2=None):
        n_src, n_tgt = pack.source_grid.total_voxels, pack.target_grid.total_voxels
        cin, dt = feats.shape[1] * (2 if feats2 is not None else 1), feats.dtype
        koi = weight_to_koi(weight, transposed)
        if not transposed:
            fmap, order, n_out = pack._map_fwd(), pack._order_for(True, cin, dt, koi.shape[1]), n_tgt
        else:
            fmap, order, n_out = pack._map_bwd(), pack._order_for(False, cin, dt, koi.shape[1]), n_src
        if feats2 is None:
            out = ext.implicit_gemm(feats.contiguous(), koi, fmap, n_out, order, bias, bool(relu))
        else:
            # input channels [feats | feats2], read from both without forming the concatenation
            out = ext.implicit_gemm_cat2(feats, feats2, koi, fmap, n_out, order, bias, bool(relu))
        # the [K, I, O] operand of dgrad is built from the weight in backward, when needed
        ctx.save_for_backward(feats, weight, out if relu else None, feats2)
        ctx.pack = pack
        ctx.transposed = transposed
        ctx.relu = bool(relu)
        ctx.bias_dtype = bias.dtype if bias is not None else None
        return out

    @staticmethod
    def backward(ctx, grad_out):
        feats, weight, out, feats2 = ctx.saved_tensors
        pack, transposed = ctx.pack, ctx.transposed
        kio = weight_to_kio(weight, transposed)  # [K, I, O]
        grad_out = grad_out.contiguous()
        if grad_out.dtype != feats.dtype:
            grad_out = grad_out.to(feats.dtype)
        if ctx.relu:
            grad_out = torch.where(out > 0, grad_out, torch.zeros_like(grad_out))
        gb = None
        if ctx.needs_input_grad[2]:
            gb = grad_out.sum(0, dtype=torch.float64 if grad_out.dtype == torch.float64 else torch.float32)
            gb = gb.to(ctx.bias_dtype)
        gin = gw = gin2 = None
        if ctx.needs_input_grad[0] or (feats2 is not None and ctx.needs_input_grad[6]):
            cg, dt = grad_out.shape[1], grad_out.dtype
            co = kio.shape[1]  # the dgrad contraction's output channels (the layer's inputs)
            if not transposed:
                dmap, order, n_in = pack._map_bwd(), pack._order_for(False, cg, dt, co), pack.source_grid.total_voxels
            else:
                dmap, order, n_in = pack._map_fwd(), pack._order_for(True, cg, dt, co), pack.target_grid.total_voxels
            gin = ext.implicit_gemm(grad_out, kio, dmap, n_in, order)
            if feats2 is not None:  # the gradient of each half of the concatenation
                c = feats.shape[1]
                gin, gin2 = gin[:, :c], gin[:, c:]
        if ctx.needs_input_grad[1]:
            x = feats if feats2 is None else torch.cat([feats, feats2], 1)
            pairs, _ = pack._pairs()
            chunk = wgrad_chunk(pairs.shape[0], x.dtype, x.shape[1], grad_out.shape[1])
            koff, cstart, total = pack._wgrad_tables(chunk)
            gk = ext.wgrad_tabled(x.contiguous(), grad_out, pairs, koff, cstart, total, transposed, chunk)
            gw = kio_to_weight(gk.to(kio.dtype), pack.kernel_size, transposed)
        return gin, gw, gb, None, None, None, gin2


def sparse_conv(feats: torch.Tensor, weight: torch.Tensor, pack, transposed: bool, bias=None,
                relu: bool = False, feats2: Optional[torch.Tensor] = None) -> torch.Tensor:
    """The convolution of `feats`, or with `feats2` of the channel concatenation
    [feats | feats2] without forming it."""
    n_in = pack.target_grid.total_voxels if transposed else pack.source_grid.total_voxels
    check_conv_args(feats, weight, n_in, transposed, feats2)
    cout = weight.shape[1] if transposed else weight.shape[0]
    if bias is not None and (bias.dim() != 1 or bias.shape[0] != cout):
        raise ValueError(f"bias must have shape [{cout}], got {tuple(bias.shape)}")
    return SparseConvFunction.apply(feats, weight, bias, pack, transposed, relu, feats2)
