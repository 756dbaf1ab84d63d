"""VDBTensor: a grid batch with its features and its cached stride-1 kernel map.

Mirrors ``fvdb/nn/vdbtensor.py:22-391`` of the reference: the dataclass checks of
``__post_init__`` (``:38-55``), the kernel-map cache (only a stride-1 map whose source
and target are this grid is kept), element-wise arithmetic on the features, and
``vdbtensor_from_dense``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Optional

import torch

from ..grid_batch import GridBatch
from ..jagged_tensor import JaggedTensor
from ..sparse_conv_pack_info import SparseConvPackInfo


@dataclass
class VDBTensor:
    grid: GridBatch
    data: JaggedTensor
    kmap: Optional[SparseConvPackInfo] = None

    def __post_init__(self):
        if not isinstance(self.grid, GridBatch):
            raise TypeError("grid should be of type GridBatch")
        if isinstance(self.data, torch.Tensor):
            self.data = self.grid.jagged_like(self.data)
        if not isinstance(self.data, JaggedTensor):
            raise TypeError("data must be a JaggedTensor or a torch.Tensor")
        if self.grid.grid_count != len(self.data):
            raise ValueError("grid and feature should have the same batch size")
        if self.grid.total_voxels != self.data.jdata.size(0):
            raise ValueError("grid and feature should have the same total voxel count")
        if self.kmap is not None:
            if not isinstance(self.kmap, SparseConvPackInfo):
                raise TypeError("kmap should be of type SparseConvPackInfo")
            if not (self.is_same(self.kmap.source_grid) and self.is_same(self.kmap.target_grid)
                    and self.kmap.stride == (1, 1, 1)):
                raise ValueError("kmap should operate on the same grid as this tensor")

    def __getitem__(self, idx):
        return VDBTensor(self.grid[idx], self.data[idx])

    def __len__(self):
        return self.grid.grid_count

    def to_dense(self) -> torch.Tensor:
        return self.grid.write_to_dense_xyzc(self.data)

    def clear_cache(self):
        self.kmap = None

    def is_same(self, other: Any) -> bool:
        if isinstance(other, VDBTensor):
            return self.grid.address == other.grid.address and self.grid_count == other.grid_count
        return self.grid.has_same_address_and_grid_count(other)

    def _binop(self, other, op):
        o = other.data if isinstance(other, VDBTensor) else other
        return VDBTensor(self.grid, op(self.data, o), self.kmap)

    def __add__(self, o): return self._binop(o, lambda a, b: a + b)
    def __sub__(self, o): return self._binop(o, lambda a, b: a - b)
    def __mul__(self, o): return self._binop(o, lambda a, b: a * b)
    def __truediv__(self, o): return self._binop(o, lambda a, b: a / b)
    def __pow__(self, o): return self._binop(o, lambda a, b: a ** b)
    def __neg__(self): return VDBTensor(self.grid, -self.data, self.kmap)

    def detach(self):
        return VDBTensor(self.grid, self.data.detach(), self.kmap)

    def type(self, dtype: torch.dtype):
        return VDBTensor(self.grid, self.data.type(dtype), self.kmap)

    def to(self, device_or_dtype: Any):
        if isinstance(device_or_dtype, torch.dtype):
            return VDBTensor(self.grid, self.data.to(device_or_dtype), self.kmap)
        return VDBTensor(self.grid.to(device_or_dtype), self.data.to(device_or_dtype), self.kmap)

    def cuda(self):
        return self.to("cuda")

    def clone(self):
        return VDBTensor(self.grid, self.data.clone(), self.kmap)

    def requires_grad_(self, flag: bool = True):
        self.data.requires_grad_(flag)
        return self

    @property
    def num_tensors(self):
        return self.data.num_tensors

    @property
    def is_cuda(self):
        return self.data.is_cuda

    @property
    def device(self):
        return self.data.device

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def jdata(self):
        return self.data.jdata

    @property
    def jidx(self):
        return self.data.jidx

    @property
    def joffsets(self):
        return self.data.joffsets

    @property
    def grid_count(self) -> int:
        return self.grid.grid_count

    @property
    def ijk(self) -> JaggedTensor:
        return self.grid.ijk

    @property
    def num_voxels(self):
        return self.grid.num_voxels

    @property
    def total_voxels(self) -> int:
        return self.grid.total_voxels

    @property
    def cum_voxels(self):
        return self.grid.cum_voxels

    @property
    def voxel_sizes(self):
        return self.grid.voxel_sizes

    @property
    def origins(self):
        return self.grid.origins


def vdbtensor_from_dense(dense_data: torch.Tensor, ijk_min=0, voxel_sizes=1, origins=0) -> VDBTensor:
    grid = GridBatch.from_dense(dense_data.size(0), tuple(dense_data.shape[1:4]), ijk_min=ijk_min,
                                voxel_sizes=voxel_sizes, origins=origins, device=dense_data.device)
    data = grid.read_from_dense_xyzc(dense_data.contiguous(), dense_origins=ijk_min)
    return VDBTensor(grid, data)
