"""Sparse U-Net built from ``fvdb.nn`` modules (the network of the U-Net workload).

Structure (levels l = 0..L-1, channels c_l, default 32/64/128/256):

* encoder level 0: SparseConv3d(in -> c0, 3), SparseConv3d(c0 -> c0, 3)
* encoder level l>0: SparseConv3d(c_{l-1} -> c_l, 2, stride 2) down, then two 3x3x3 convs
* decoder level l<L-1: SparseConv3d(c_{l+1} -> c_l, 2, stride 2, transposed) onto the level-l
  grid, ``fvdb.jcat([up, skip_l], dim=1)`` (2 c_l channels), then two 3x3x3 convs
  (2 c_l -> c_l -> c_l); the first reads the concatenation's two halves from ``up`` and
  ``skip_l`` (``SparseConv3d(..., concat=skip_l)``) instead of forming it
* optional 1x1x1 head

A ReLU follows every convolution except the head; it and the bias run in the
convolution's epilogue (``SparseConv3d(..., fuse_relu=True)``).  This is the usage pattern of the
reference's U-Net tutorial (``docs/tutorials/simple_unet.md:108-245``: strided down
convs, ``transposed=True`` up convs with ``out_grid=``, ``jcat(dim=1)`` skips), without
batch norm so that each voxel's output depends only on its own neighbourhood.
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import torch
import torch.nn as nn

from .modules import ReLU, SparseConv3d
from .vdbtensor import VDBTensor


class SparseUNet(nn.Module):
    def __init__(self, in_channels: int = 32, channels: Sequence[int] = (32, 64, 128, 256),
                 out_channels: Optional[int] = None, bias: bool = True):
        super().__init__()
        self.channels = list(channels)
        L = len(self.channels)
        self.enc = nn.ModuleList()
        self.down = nn.ModuleList()
        for l, c in enumerate(self.channels):
            cin = in_channels if l == 0 else c
            if l > 0:
                self.down.append(SparseConv3d(self.channels[l - 1], c, 2, stride=2, bias=bias))
            self.enc.append(nn.ModuleList([SparseConv3d(cin, c, 3, bias=bias), SparseConv3d(c, c, 3, bias=bias)]))
        self.up = nn.ModuleList()
        self.dec = nn.ModuleList()
        for l in range(L - 2, -1, -1):
            c = self.channels[l]
            self.up.append(SparseConv3d(self.channels[l + 1], c, 2, stride=2, bias=bias, transposed=True))
            self.dec.append(nn.ModuleList([SparseConv3d(2 * c, c, 3, bias=bias), SparseConv3d(c, c, 3, bias=bias)]))
        self.head = SparseConv3d(self.channels[0], out_channels, 1, bias=bias) if out_channels else None
        self.relu = ReLU()

    def forward(self, x: VDBTensor) -> VDBTensor:
        skips: List[VDBTensor] = []
        for l, convs in enumerate(self.enc):
            if l > 0:
                x = self.down[l - 1](x, fuse_relu=True)
            for conv in convs:
                x = conv(x, fuse_relu=True)
            skips.append(x)
        L = len(self.channels)
        for i, l in enumerate(range(L - 2, -1, -1)):
            skip = skips[l]
            x = self.up[i](x, out_grid=skip.grid, fuse_relu=True)
            x = VDBTensor(skip.grid, x.data, skip.kmap)  # same grid: reuse its stride-1 kernel map
            # the first decoder convolution takes fvdb.jcat([x, skip], dim=1), read from both
            x = self.dec[i][0](x, fuse_relu=True, concat=skip)
            for conv in self.dec[i][1:]:
                x = conv(x, fuse_relu=True)
        if self.head is not None:
            x = self.head(x)
        return x
