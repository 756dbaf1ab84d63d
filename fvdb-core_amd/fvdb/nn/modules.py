"""``fvdb.nn`` modules on the convolution path.

``SparseConv3d`` mirrors ``fvdb/nn/modules.py:156-404`` of the reference: weight
storage (``[Do, Di, k0, k1, k2]`` viewed over ``(k2, k1, k0, Di, Do)`` storage,
``:234-242``), ``reset_parameters`` (``U(+-1/sqrt(fan_in))``, ``:261-265``), the 1x1x1
stride-1 matmul fast path and bias (``:384-404``), the kernel-map cache of
``VDBTensor`` (reuse for stride 1 on the same grid, ``:325-341``), and the backend
names of ``_dispatch_conv`` (``:267-348``).  Every backend name runs on the gfx950
kernels: "default"/"igemm_mode*"/"cutlass"/"lggs"/"halo" use the implicit-GEMM
pack, "legacy"/"me" the gather-scatter pack (same kernels, the reference's GS
metadata), "dense" is the reference's dense torch path.  ``allow_tf32`` is accepted;
gfx950 has no TF32 MFMA, so FP32 convolutions always run in full FP32.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence, Union

import torch
import torch.nn as nn
from torch.profiler import record_function

from ..grid_batch import GridBatch
from ..jagged_tensor import JaggedTensor, jcat_tensors
from ..sparse_conv_pack_info import ConvPackBackend, SparseConvPackInfo
from .vdbtensor import VDBTensor


def fvnn_module(module):
    old_forward = module.forward

    def _forward(self, *args, **kwargs):
        with record_function(repr(self)):
            return old_forward(self, *args, **kwargs)

    module.forward = _forward
    return module


GridOrVDBTensor = Union[GridBatch, VDBTensor]


def _ref_grid(ref):
    if isinstance(ref, VDBTensor):
        return ref.grid, ref.kmap
    if isinstance(ref, GridBatch):
        return ref, None
    return None, None


@fvnn_module
class MaxPool(nn.Module):
    """3D max pooling over voxel windows (``fvdb/nn/modules.py:33-69`` of the reference).

    Coarse voxels whose window holds no fine voxel get zero (the reference's ``-inf -> 0``).
    """

    def __init__(self, kernel_size, stride=None):
        super().__init__()
        self.kernel_size = kernel_size
        self.stride = stride or self.kernel_size

    def forward(self, input: VDBTensor, ref_coarse_data: Optional[GridOrVDBTensor] = None) -> VDBTensor:
        coarse_grid, coarse_kmap = _ref_grid(ref_coarse_data)
        new_feature, new_grid = input.grid.max_pool(self.kernel_size, input.data, stride=self.stride,
                                                    coarse_grid=coarse_grid)
        jd = new_feature.jdata
        jd = torch.where(torch.isinf(jd), torch.zeros_like(jd), jd)
        return VDBTensor(new_grid, new_grid.jagged_like(jd), kmap=coarse_kmap)

    def extra_repr(self) -> str:
        return f"kernel_size={self.kernel_size}, stride={self.stride}"


@fvnn_module
class AvgPool(nn.Module):
    """3D average pooling over voxel windows (``fvdb/nn/modules.py:70-101`` of the reference)."""

    def __init__(self, kernel_size, stride=None):
        super().__init__()
        self.kernel_size = kernel_size
        self.stride = stride or self.kernel_size

    def forward(self, input: VDBTensor, ref_coarse_data: Optional[GridOrVDBTensor] = None) -> VDBTensor:
        coarse_grid, coarse_kmap = _ref_grid(ref_coarse_data)
        new_feature, new_grid = input.grid.avg_pool(self.kernel_size, input.data, stride=self.stride,
                                                    coarse_grid=coarse_grid)
        return VDBTensor(new_grid, new_feature, kmap=coarse_kmap)

    def extra_repr(self) -> str:
        return f"kernel_size={self.kernel_size}, stride={self.stride}"


@fvnn_module
class UpsamplingNearest(nn.Module):
    """Nearest upsampling by a scale factor (``fvdb/nn/modules.py:102-129`` of the reference)."""

    def __init__(self, scale_factor):
        super().__init__()
        self.scale_factor = scale_factor

    def forward(self, input: VDBTensor, mask: Optional[JaggedTensor] = None,
                ref_fine_data: Optional[GridOrVDBTensor] = None) -> VDBTensor:
        fine_grid, fine_kmap = _ref_grid(ref_fine_data)
        new_feature, new_grid = input.grid.subdivide(self.scale_factor, input.data, mask, fine_grid=fine_grid)
        return VDBTensor(new_grid, new_feature, kmap=fine_kmap)

    def extra_repr(self) -> str:
        return f"scale_factor={self.scale_factor}"


@fvnn_module
class FillFromGrid(nn.Module):
    """Copy a VDBTensor's features onto another grid (``fvdb/nn/modules.py:131-154``)."""

    def __init__(self, default_value: float = 0.0) -> None:
        super().__init__()
        self.default_value = default_value

    def forward(self, input: VDBTensor, other_data: Optional[GridOrVDBTensor] = None) -> VDBTensor:
        other_grid, other_kmap = _ref_grid(other_data)
        if other_grid is None:
            return input
        new_feature = other_grid.inject_from(input.grid, input.data, default_value=self.default_value)
        return VDBTensor(other_grid, new_feature, kmap=other_kmap)


@fvnn_module
class SparseConv3d(nn.Module):
    CUTLASS_SUPPORTED_CHANNELS = [
        (32, 64), (64, 128), (128, 256), (32, 32), (64, 64), (128, 128), (256, 256), (128, 64), (64, 32),
        (256, 128), (384, 256), (192, 128), (256, 512), (512, 256), (512, 512),
    ]
    backend: str = "default"
    allow_tf32: bool = True

    def __init__(self, in_channels: int, out_channels: int, kernel_size: Union[int, Sequence] = 3,
                 stride: Union[int, Sequence] = 1, bias: bool = True, transposed: bool = False) -> None:
        super().__init__()
        self.in_channels = in_channels
        self.out_channels = out_channels
        if isinstance(kernel_size, int):
            kernel_size = (kernel_size,) * 3
        assert len(kernel_size) == 3
        if isinstance(stride, int):
            stride = (stride,) * 3
        assert len(stride) == 3
        self.kernel_size = tuple(kernel_size)
        self.stride = tuple(stride)
        self.transposed = transposed
        if self.transposed:
            out_channels, in_channels = in_channels, out_channels
        self.kernel_volume = math.prod(self.kernel_size)
        if self.kernel_volume > 1:
            weight_shape = [out_channels, in_channels] + list(self.kernel_size)
            weight = torch.zeros(*weight_shape[::-1]).permute(4, 3, 2, 1, 0)
            self.weight = nn.Parameter(weight)
        else:
            self.weight = nn.Parameter(torch.zeros(out_channels, in_channels))
        if bias:
            self.bias = nn.Parameter(torch.empty(self.out_channels))
        else:
            self.register_parameter("bias", None)
        self.reset_parameters()

    def extra_repr(self) -> str:
        s = "{in_channels}, {out_channels}, kernel_size={kernel_size}"
        if self.stride != (1, 1, 1):
            s += ", stride={stride}"
        if self.bias is None:
            s += ", bias=False"
        if self.transposed:
            s += ", transposed=True"
        return s.format(**self.__dict__)

    def reset_parameters(self) -> None:
        std = 1 / math.sqrt((self.out_channels if self.transposed else self.in_channels) * self.kernel_volume)
        self.weight.data.uniform_(-std, std)
        if self.bias is not None:
            self.bias.data.uniform_(-std, std)

    def _build_kmap_and_convert_backend(self, kmap: SparseConvPackInfo, backend: str) -> ConvPackBackend:
        if backend in ["legacy", "me"]:
            kmap.build_gather_scatter(backend == "me")
            return ConvPackBackend.GATHER_SCATTER
        if backend == "cutlass":
            kmap.build_cutlass(benchmark=False)
            return ConvPackBackend.CUTLASS
        if backend in ("igemm_mode0", "igemm_mode1", "igemm_mode2"):
            kmap.build_implicit_gemm(sorted=backend != "igemm_mode0", split_mask_num=3 if backend == "igemm_mode2"
                                     else 1, training=self.training, split_mask_num_bwd=3, use_tf32=self.allow_tf32)
            return ConvPackBackend.IGEMM
        if backend == "lggs":
            kmap.build_lggs()
            return ConvPackBackend.LGGS
        raise NotImplementedError(f"Backend {backend} is not supported")

    def _dispatch_conv(self, in_feature: JaggedTensor, in_grid: GridBatch, in_kmap, out_grid, relu: bool = False,
                       concat: Optional[JaggedTensor] = None):
        """(out_grid, out_feature, out_kmap, fused): fused tells whether bias and ReLU
        were applied in the convolution's epilogue.  `concat`: features whose channels follow
        in_feature's (the implicit-GEMM backends read both without concatenating them)."""
        if concat is not None and (self.transposed or self.backend in ("halo", "dense")):
            in_feature, concat = jcat_tensors([in_feature, concat], dim=1), None
        backend = self.backend
        if backend == "cutlass" and (self.in_channels, self.out_channels) not in self.CUTLASS_SUPPORTED_CHANNELS:
            backend = "default"
        if backend == "lggs" and (self.in_channels, self.out_channels) != (128, 128):
            backend = "default"
        if backend == "default":
            # the reference picks "legacy" for fp64 and "igemm_mode1" otherwise; here both run
            # the gfx950 implicit GEMM, without sort metadata (rows are already in tree order)
            backend = "legacy" if in_feature.dtype == torch.float64 else "igemm_mode0"
        if backend == "halo" and self.stride == (1, 1, 1) and self.kernel_size == (3, 3, 3):
            assert out_grid is None or in_grid.is_same(out_grid)
            return in_grid, in_grid.sparse_conv_halo(in_feature, self.weight, 8), None, False
        if backend == "dense" and self.stride == (1, 1, 1):
            assert out_grid is None or in_grid.is_same(out_grid)
            min_coord = in_grid.ijk.jdata.min(dim=0).values
            dense = in_grid.write_to_dense_czyx(in_feature, min_coord=min_coord)
            dense = torch.nn.functional.conv3d(dense, self.weight, padding=1, stride=1)
            return in_grid, in_grid.read_from_dense_czyx(dense.contiguous(), dense_origins=min_coord), None, False
        if backend in ("halo", "dense"):
            backend = "igemm_mode0"
        can_cache = self.stride == (1, 1, 1) and (out_grid is None or out_grid.is_same(in_grid))
        if in_kmap is not None and in_kmap.kernel_size == self.kernel_size and can_cache:
            kmap, out_grid = in_kmap, in_grid
        else:
            if self.transposed:
                assert out_grid is not None
                kmap, _ = out_grid.sparse_conv_kernel_map(self.kernel_size, self.stride, in_grid)
            else:
                kmap, out_grid = in_grid.sparse_conv_kernel_map(self.kernel_size, self.stride, out_grid)
        out_kmap = kmap if can_cache else None
        be = self._build_kmap_and_convert_backend(kmap, backend)
        if not self.transposed:
            out_feature = kmap.sparse_conv_3d(in_feature, self.weight, be, bias=self.bias, relu=relu, concat=concat)
        else:
            out_feature = kmap.sparse_transpose_conv_3d(in_feature, self.weight, be, bias=self.bias, relu=relu)
        return out_grid, out_feature, out_kmap, True

    def forward(self, input: VDBTensor, out_grid: Optional[GridBatch] = None, *, fuse_relu: bool = False,
                concat: Optional[VDBTensor] = None) -> VDBTensor:
        """``fuse_relu`` applies a ReLU in the convolution's epilogue (same values as a
        following ``fvdb.nn.ReLU``, without a separate pass over the output).  ``concat`` (a
        VDBTensor on the input's grid): the layer convolves ``fvdb.jcat([input, concat],
        dim=1)``, the U-Net decoder's skip connection, read from both tensors without
        forming the concatenation (same result)."""
        in_feature, in_grid, in_kmap = input.data, input.grid, input.kmap
        cat = None
        if concat is not None:
            if not (concat.grid is in_grid or concat.grid.is_same(in_grid)):
                raise ValueError("concat must be on the input's grid")
            cat = concat.data
            if in_kmap is None:
                in_kmap = concat.kmap
        fused = False
        if cat is not None and self.kernel_size == (1, 1, 1) and self.stride == (1, 1, 1):
            in_feature, cat = jcat_tensors([in_feature, cat], dim=1), None
        if self.kernel_size == (1, 1, 1) and self.stride == (1, 1, 1):
            out_feature = in_feature.jagged_like(in_feature.jdata.matmul(self.weight.transpose(0, 1)))
            out_grid, out_kmap = in_grid, in_kmap
        else:
            out_grid, out_feature, out_kmap, fused = self._dispatch_conv(in_feature, in_grid, in_kmap, out_grid,
                                                                         fuse_relu, cat)
        if not fused:
            if self.bias is not None:
                out_feature.jdata = out_feature.jdata + self.bias
            if fuse_relu:
                out_feature.jdata = torch.relu(out_feature.jdata)
        if out_grid is None:
            raise ValueError("Failed to compute output grid. This is a bug in the implementation.")
        return VDBTensor(out_grid, out_feature, out_kmap)


class ElementwiseMixin:
    def forward(self, input: VDBTensor) -> VDBTensor:
        res = super().forward(input.data.jdata)  # type: ignore[misc]
        return VDBTensor(input.grid, input.grid.jagged_like(res), input.kmap)


@fvnn_module
class ReLU(ElementwiseMixin, nn.ReLU):
    pass


@fvnn_module
class LeakyReLU(ElementwiseMixin, nn.LeakyReLU):
    pass


@fvnn_module
class SiLU(ElementwiseMixin, nn.SiLU):
    pass


@fvnn_module
class GELU(ElementwiseMixin, nn.GELU):
    pass


@fvnn_module
class Sigmoid(ElementwiseMixin, nn.Sigmoid):
    pass


@fvnn_module
class Tanh(ElementwiseMixin, nn.Tanh):
    pass


@fvnn_module
class Linear(ElementwiseMixin, nn.Linear):
    pass


@fvnn_module
class Dropout(ElementwiseMixin, nn.Dropout):
    pass


@fvnn_module
class BatchNorm(nn.BatchNorm1d):
    def forward(self, input: VDBTensor) -> VDBTensor:
        assert input.data.jdata.size(1) == self.num_features
        res = super().forward(input.data.jdata)
        return VDBTensor(input.grid, input.grid.jagged_like(res), input.kmap)


@fvnn_module
class GroupNorm(nn.GroupNorm):
    """``torch.nn.GroupNorm`` per grid of a VDBTensor (``fvdb/nn/modules.py:407-432``): statistics
    per (grid, group) over the grid's voxels.  Runs on the native group-norm kernels (two
    passes over the batch, deterministic reductions); ``fuse_relu=True`` applies a ReLU in the
    normalize pass."""

    def forward(self, input: VDBTensor, *, fuse_relu: bool = False) -> VDBTensor:
        from .._norm import group_norm

        C = input.data.jdata.size(1)
        assert C == self.num_channels, "Input feature should have the same number of channels as GroupNorm"
        out = group_norm(input.data.jdata, input.grid, self.num_groups, self.weight, self.bias, self.eps,
                         relu=fuse_relu)
        return VDBTensor(input.grid, input.grid.jagged_like(out), input.kmap)


def fuse_conv_bn(conv: "SparseConv3d", bn: "BatchNorm") -> "SparseConv3d":
    """A SparseConv3d equal to ``bn(conv(x))`` with the batch norm in inference mode folded into
    the convolution: W' = W * g / sqrt(var + eps) per output channel, b' = (b - mean) * g /
    sqrt(var + eps) + beta.  The folded bias runs in the convolution's epilogue (with a fused
    ReLU when the caller passes ``fuse_relu=True``), so the normalization costs no pass of its
    own."""
    if bn.training or bn.running_mean is None:
        raise ValueError("fuse_conv_bn folds a batch norm in inference mode (bn.eval(), running statistics)")
    if bn.num_features != conv.out_channels:
        raise ValueError("the batch norm's features must be the convolution's output channels")
    fused = SparseConv3d(conv.in_channels, conv.out_channels, conv.kernel_size, conv.stride, bias=True,
                         transposed=conv.transposed).to(device=conv.weight.device, dtype=conv.weight.dtype)
    fused.backend = conv.backend
    with torch.no_grad():
        scale = bn.weight.double() / torch.sqrt(bn.running_var.double() + bn.eps) if bn.affine else \
            1.0 / torch.sqrt(bn.running_var.double() + bn.eps)
        shift = (bn.bias.double() if bn.affine else 0.0) - bn.running_mean.double() * scale
        w = conv.weight.double()
        # output channels are dim 0, or dim 1 for a transposed conv's [I, O, k...] weight
        shape = [1] * w.dim()
        shape[1 if conv.transposed else 0] = -1
        fused.weight.copy_((w * scale.reshape(shape)).to(conv.weight.dtype))
        b = conv.bias.double() if conv.bias is not None else torch.zeros_like(scale)
        fused.bias.copy_((b * scale + shift).to(conv.weight.dtype))
    return fused
