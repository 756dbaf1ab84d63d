"""SparseConvPackInfo: kernel maps of one (source grid, target grid, kernel, stride).

Mirrors the reference's ``fvdb/sparse_conv_pack_info.py:33-374`` and
``src/fvdb/SparseConvPackInfo.cpp:26-330``:

* constructor: target = source for stride 1, ``conv_grid`` otherwise, or the given
  target (``SparseConvPackInfo.cpp:26-59``); a 1x1x1 stride-1 conv is rejected.
* ``build_gather_scatter``: kernel map -> ``neighborhood_map`` [P,2] (source index,
  target row) ordered by (k, row) and ``neighborhood_sizes`` [K] (``:61-88``).
* ``build_implicit_gemm``: ``out_in_map`` [ceil(N_out/128)*128, K] and, for training,
  ``out_in_map_bwd`` (``:90-148``).  With ``sorted=True`` the reference's sort
  metadata (bitmask per split, descending sort -> ``reorder_loc``, reordered map,
  OR-reduced masks) is produced with the same definitions; the order of equal masks is
  unspecified there (``torch::sort`` is not stable) and stable here (row order), and with
  one split ``reorder_loc`` is the order the sorted contraction runs.
* ``build_cutlass`` / ``build_lggs``: accepted; both backends run on the same gfx950
  implicit-GEMM kernel (their buffers are reported as ``None``).
* ``sparse_conv_3d`` / ``sparse_transpose_conv_3d`` with the reference's checks
  (``:248-330``).

Every map here is computed on the GPU by ``fvdbk_kmap_build``; backward maps and
neighbour pairs needed by a backward pass are built on first use.
"""
from __future__ import annotations

import enum
from typing import Optional, Tuple

import torch

from ._autograd import sparse_conv
from ._lib import ext
from .jagged_tensor import JaggedTensor
from .types import is_Vec3iOrScalar, to_vec3i


class ConvPackBackend(enum.Enum):
    GATHER_SCATTER = 0
    IGEMM = 1
    CUTLASS = 2
    LGGS = 3


GATHER_SCATTER = ConvPackBackend.GATHER_SCATTER
IGEMM = ConvPackBackend.IGEMM
CUTLASS = ConvPackBackend.CUTLASS
LGGS = ConvPackBackend.LGGS


def _pad128(n: int) -> int:
    return (n + 127) // 128 * 128


AUTO_ORDER_WINDOW = 8192


def wants_row_order(kernel_volume: int, cin: int, dtype: torch.dtype, cout: int = 0, rows: int = 0) -> bool:
    """Whether a contraction runs faster with its rows ordered by tap mask per window
    (measured on MI355X): fp32 (MFMA-bound, so each skipped tap is saved time), 16-bit
    inputs of 128+ channels, and 16-bit 3x3x3 layers of 64 inputs and outputs in
    multiples of 128 channels.  16-bit 3x3x3 layers of 64 or 128 inputs then run the
    direct-to-LDS kernel with compacted steps (planes + spheres, 1.1M rows: 128 -> 128
    0.73 ms windowed vs 0.76 ms in tree order, 128 -> 64 0.37 vs 0.40 ms, 64 -> 128 0.47 vs
    1.02 ms).  Other narrow 16-bit layers stay in tree order, whose gathers hit cache
    (3x3x3 64 -> 64 on the sphere batch: 0.73 ms tree vs 0.81 ms windowed); that includes
    maps with one tap per row (a 2x2x2 stride-2 kernel's fine-to-coarse map), where tree
    order computes 8 taps per row.  Only the kernel volumes with tap-skipping kernels (27,
    8) qualify; fp64 runs the scalar kernel.  ``rows`` is accepted for callers that pass
    it and does not change the choice."""
    del rows
    if kernel_volume not in (8, 27) or dtype == torch.float64:
        return False
    if dtype == torch.float32:
        return True
    if kernel_volume == 27 and cin == 64 and cout > 0 and cout % 128 == 0:
        return True
    return cin >= 128


class SparseConvPackInfo:
    def __init__(self, kernel_size=None, stride=None, source_grid=None, target_grid=None, *, impl=None):
        from .grid_batch import GridBatch

        if impl is not None:
            raise ValueError("constructing from an implementation object is internal to the reference binding")
        if kernel_size is None or stride is None or source_grid is None:
            raise ValueError("kernel_size, stride, and source_grid must be provided when impl is None")
        if not isinstance(source_grid, GridBatch):
            raise TypeError("source_grid must be a GridBatch")
        if target_grid is not None and not isinstance(target_grid, GridBatch):
            raise TypeError("target_grid must be a GridBatch")
        if not is_Vec3iOrScalar(kernel_size) or not is_Vec3iOrScalar(stride):
            raise TypeError(f"kernel_size and stride must be of type Vec3iOrScalar, but got {type(kernel_size)} "
                            f"and {type(stride)}")
        ks = to_vec3i(kernel_size, "kernel_size")
        st = to_vec3i(stride, "stride")
        if any(k <= 0 for k in ks):
            raise RuntimeError(f"Expect kernel size to be larger than {{0,0,0}}, but got {ks}.")
        if any(s <= 0 for s in st):
            raise RuntimeError(f"Expect stride to be larger than 0, but got {st}.")
        if target_grid is None:
            target_grid = source_grid if st == (1, 1, 1) else source_grid.conv_grid(ks, st)
        if source_grid.device != target_grid.device:
            raise RuntimeError("Source and target grids must both be on the same device")
        if ks == (1, 1, 1) and st == (1, 1, 1):
            raise RuntimeError("1x1 conv does not need kernel map to be built!")
        if source_grid.grid_count != target_grid.grid_count:
            raise ValueError("source and target grids must hold the same number of grids")
        self._ks: Tuple[int, int, int] = ks
        self._st: Tuple[int, int, int] = st
        self._src = source_grid
        self._tgt = target_grid
        self._kmap: Optional[torch.Tensor] = None   # [pad128(N_tgt), K]
        self._tmap: Optional[torch.Tensor] = None   # [pad128(N_src), K]
        self._pairs_t: Optional[torch.Tensor] = None
        self._sizes_t: Optional[torch.Tensor] = None
        self._sizes_host: Optional[torch.Tensor] = None
        self._wg_tables = {}
        self._gs_built = False
        self._use_me = False
        self._igemm = None  # dict of build flags
        self._sorted = False
        self._ord_fwd: Optional[torch.Tensor] = None
        self._ord_bwd: Optional[torch.Tensor] = None
        self._ord_auto = {}
        self._sorted_meta = {}
        self._use_tf32 = False
        self._cutlass = None
        self._lggs = False

    # ----------------------------------------------------------------------------
    # maps (GPU)
    # ----------------------------------------------------------------------------
    @property
    def kernel_volume(self) -> int:
        return self._ks[0] * self._ks[1] * self._ks[2]

    def _map_fwd(self) -> torch.Tensor:
        """Kernel map, rows = target voxels (``ConvolutionKernelMap.cu:104-138``)."""
        if self._kmap is None:
            self._kmap = ext.kernel_map(self._tgt._impl.tensors(), self._src._impl.tensors(), self._src.grid_count,
                                        list(self._ks), list(self._st), False, _pad128(self._tgt.total_voxels))
        return self._kmap

    def _map_bwd(self) -> torch.Tensor:
        """Transposed map, rows = source voxels (``IGEMMBitOperations.cu:92-105``)."""
        if self._tmap is None:
            self._tmap = ext.kernel_map(self._src._impl.tensors(), self._tgt._impl.tensors(), self._src.grid_count,
                                        list(self._ks), list(self._st), True, _pad128(self._src.total_voxels))
        return self._tmap

    def _order_fwd(self) -> Optional[torch.Tensor]:
        """Target rows ordered by tap mask (for the sorted contraction), or None.  For a
        build with one split this is ``reorder_loc[0][:N_out]``: the order the property
        reports is the order the contraction runs."""
        if not self._sorted:
            return None
        if self._ord_fwd is None:
            self._ord_fwd = ext.row_order(self._map_fwd(), self._tgt.total_voxels)
        return self._ord_fwd

    def _order_bwd(self) -> Optional[torch.Tensor]:
        """Source rows of the transposed map ordered by tap mask, or None (for a training
        build with one backward split, ``reorder_loc_bwd[0][:N_in]``)."""
        if not self._sorted:
            return None
        if self._ord_bwd is None:
            self._ord_bwd = ext.row_order(self._map_bwd(), self._src.total_voxels)
        return self._ord_bwd

    # Rows of a tile that share taps let the contraction skip the taps none of them uses.
    # Without an explicit sorted build, the rows are ordered by tap mask inside windows
    # of AUTO_ORDER_WINDOW consecutive rows (neighbours stay close in memory) where
    # ``wants_row_order`` says so.  Narrow 16-bit layers stay in tree order, whose
    # gathers hit cache.
    def _order_for(self, fwd: bool, cin: int, dtype: torch.dtype, cout: int = 0) -> Optional[torch.Tensor]:
        """Row order for a contraction over the forward (rows = target voxels) or
        transposed (rows = source voxels) map with ``cin`` input and ``cout`` output
        channels, or None."""
        if self._sorted:
            return self._order_fwd() if fwd else self._order_bwd()
        n = self._tgt.total_voxels if fwd else self._src.total_voxels
        if not wants_row_order(self.kernel_volume, cin, dtype, cout, n):
            return None
        if n == 0:
            return None
        if self._ord_auto.get(fwd) is None:
            m = self._map_fwd() if fwd else self._map_bwd()
            self._ord_auto[fwd] = ext.row_order(m, n, AUTO_ORDER_WINDOW)
        return self._ord_auto[fwd]

    def _wgrad_tables(self, chunk: int):
        """Device split-K tables of the weight gradient for `chunk` pairs per chunk: pair offsets
        per tap (int64 [K+1]), first chunk per tap (int32 [K+1]) and the chunk count; built once
        per chunk size from the host pair counts."""
        if chunk not in self._wg_tables:
            _, sizes = self._pairs()
            s = sizes.tolist()
            koff = [0]
            cst = [0]
            for n in s:
                koff.append(koff[-1] + n)
                cst.append(cst[-1] + -(-n // chunk))
            dev = self._src.device
            self._wg_tables[chunk] = (torch.tensor(koff, dtype=torch.int64).to(dev),
                                      torch.tensor(cst, dtype=torch.int32).to(dev), cst[-1])
        return self._wg_tables[chunk]

    def _pairs(self):
        if self._pairs_t is None:
            pairs, sizes = ext.neighbor_map(self._map_fwd(), self._tgt.total_voxels)
            self._pairs_t, self._sizes_t = pairs, sizes
            self._sizes_host = sizes.to("cpu", torch.int64)
        return self._pairs_t, self._sizes_host

    # ----------------------------------------------------------------------------
    # builders
    # ----------------------------------------------------------------------------
    def build_gather_scatter(self, use_me: bool = False) -> None:
        if self._gs_built:
            if self._use_me != use_me:
                raise RuntimeError("Gather scatter is already built with different use_me value")
            return
        self._pairs()
        self._gs_built = True
        self._use_me = bool(use_me)

    def build_implicit_gemm(self, sorted: bool = False, split_mask_num: int = 1, training: bool = False,
                            split_mask_num_bwd: int = 1, use_tf32: bool = False) -> None:
        if self._igemm is not None:
            if self._igemm["sorted"] and self._igemm["split"] != split_mask_num:
                raise RuntimeError("Implicit GEMM is already built with different splitMaskNum value")
            if self._igemm["training"] != training:
                raise RuntimeError("Implicit GEMM is already built with different training flag")
            return
        self._map_fwd()
        if training:
            self._map_bwd()
        self._igemm = {"sorted": bool(sorted), "split": int(split_mask_num), "training": bool(training),
                       "split_bwd": int(split_mask_num_bwd)}
        self._use_tf32 = bool(use_tf32)
        self._sorted = bool(sorted)
        if sorted:
            self._sorted_meta = self._build_sorted_meta(self._kmap, split_mask_num, self._tgt.total_voxels, 128,
                                                        self._order_fwd())
        if training and sorted:
            bwd = self._build_sorted_meta(self._tmap, split_mask_num_bwd, self._src.total_voxels, 128,
                                          self._order_bwd())
            self._sorted_meta.update({
                "reorder_loc_bwd": bwd["reorder_loc"],
                "reorder_out_in_map_bwd": bwd["reorder_out_in_map"],
                "sorted_mask_bwd_d": bwd["reduced_sorted_mask"],
                "sorted_mask_bwd_w": self._reduce_mask(bwd["sorted_mask"], 64),
            })

    @staticmethod
    def _reduce_mask(sorted_mask: torch.Tensor, tile: int) -> torch.Tensor:
        S, R = sorted_mask.shape
        n = (R + tile - 1) // tile
        pad = n * tile - R
        m = torch.nn.functional.pad(sorted_mask, (0, pad)) if pad else sorted_mask
        out = torch.zeros((S, n), dtype=sorted_mask.dtype, device=sorted_mask.device)
        m = m.reshape(S, n, tile)
        for i in range(tile):  # OR over the rows of each tile
            out |= m[:, :, i]
        return out

    def _build_sorted_meta(self, kmap: torch.Tensor, splits: int, n_rows: int, tile: int,
                           rows_order: torch.Tensor) -> dict:
        """Sort metadata of ``IGEMMBitOperations.cu:14-170`` (bitmask, reorder, reduce).

        The reference sorts each split's bitmask over all padded rows with an unstable
        descending sort, so the order of equal masks is unspecified there.  Here the sort
        is stable (equal masks in row order).  With one split the bitmask is the tap mask
        of ``fvdbk_kmap_row_keys`` (27 taps fit its 31 bits), and ``reorder_loc[0]`` is
        `rows_order` -- the stable order the sorted contraction runs -- followed by the
        padding rows (mask 0, which a stable sort keeps after every real row)."""
        K = kmap.shape[1]
        per = (K + splits - 1) // splits
        valid = kmap >= 0
        valid[n_rows:] = False
        bits = torch.zeros((splits, kmap.shape[0]), dtype=torch.int32, device=kmap.device)
        for s in range(splits):
            for i in range(per):
                k = s * per + i
                if k < K:
                    bits[s] |= valid[:, k].to(torch.int32) << i
        if splits == 1 and K <= 31:
            pad = torch.arange(n_rows, kmap.shape[0], dtype=torch.long, device=kmap.device)
            order = torch.cat([rows_order.long(), pad])[None]
            smask = bits.gather(1, order)
        else:
            smask, order = torch.sort(bits, dim=-1, descending=True, stable=True)
        reorder_loc = order.to(torch.int32)
        reordered = kmap[order[0].long()] if splits == 1 else kmap[order.long()]
        return {"sorted_mask": smask, "reorder_loc": reorder_loc, "reorder_out_in_map": reordered,
                "reduced_sorted_mask": self._reduce_mask(smask, tile)}

    def build_cutlass(self, benchmark: bool = False) -> None:
        if self._cutlass is not None and self._cutlass != benchmark:
            raise RuntimeError("Cutlass is already built with different benchmark flag")
        self._map_fwd()
        self._cutlass = bool(benchmark)

    def build_lggs(self) -> None:
        if self._ks != (3, 3, 3):
            raise RuntimeError("LGGS only supports 3x3x3 kernel size")
        self._map_fwd()
        self._lggs = True

    # ----------------------------------------------------------------------------
    # convolution
    # ----------------------------------------------------------------------------
    def _check_backend(self, backend: ConvPackBackend, transposed: bool) -> None:
        if not isinstance(backend, ConvPackBackend):
            raise TypeError(f"backend must be a ConvPackBackend, got {type(backend)}")
        if backend == ConvPackBackend.GATHER_SCATTER and not self._gs_built:
            raise RuntimeError("Neighbor map must be built for sparse convolution")
        if backend == ConvPackBackend.IGEMM and self._igemm is None:
            raise RuntimeError("Implicit GEMM map must be built for sparse convolution")
        if backend == ConvPackBackend.CUTLASS:
            if transposed:
                raise RuntimeError("Cutlass does not support transpose convolution yet")
            if self._cutlass is None:
                raise RuntimeError("Cutlass buffer is not built")
        if backend == ConvPackBackend.LGGS:
            if transposed:
                raise RuntimeError("Unknown backend")
            if not self._lggs:
                raise RuntimeError("LGGS buffer is not built")

    def sparse_conv_3d(self, input, weights: torch.Tensor,
                       backend: ConvPackBackend = ConvPackBackend.GATHER_SCATTER, *, bias=None,
                       relu: bool = False, concat=None) -> JaggedTensor:
        """``concat`` (same layout as ``input``): convolve the channel concatenation
        [input | concat], as ``fvdb.jcat([input, concat], dim=1)`` would form it, without
        forming it (16-bit layers read both arrays in the contraction; bit-identical)."""
        j = input if isinstance(input, JaggedTensor) else JaggedTensor(input)
        if j.num_tensors != self._src.grid_count:
            raise ValueError("Input batch size must match target grid batch size")
        if j.element_count != self._src.total_voxels:
            raise ValueError("Input element count must match target grid total voxels")
        self._check_backend(backend, False)
        f2 = None
        if concat is not None:
            f2 = concat.jdata if isinstance(concat, JaggedTensor) else concat
        out = sparse_conv(j.jdata, weights, self, False, bias, relu, f2)
        return self._tgt.jagged_like(out)

    def sparse_transpose_conv_3d(self, input, weights: torch.Tensor,
                                 backend: ConvPackBackend = ConvPackBackend.GATHER_SCATTER, *, bias=None,
                                 relu: bool = False) -> JaggedTensor:
        j = input if isinstance(input, JaggedTensor) else JaggedTensor(input)
        if j.num_tensors != self._tgt.grid_count:
            raise ValueError("Input batch size must match target grid batch size")
        if j.element_count != self._tgt.total_voxels:
            raise ValueError("Input element count must match target grid total voxels")
        self._check_backend(backend, True)
        out = sparse_conv(j.jdata, weights, self, True, bias, relu)
        return self._src.jagged_like(out)

    # ----------------------------------------------------------------------------
    # device moves
    # ----------------------------------------------------------------------------
    def to(self, to_device) -> "SparseConvPackInfo":
        dev = torch.device(to_device) if not isinstance(to_device, torch.device) else to_device
        if dev.type != "cuda":
            raise ValueError("fvdb-core_amd kernel maps live on HIP devices only")
        return self

    def cuda(self) -> "SparseConvPackInfo":
        return self.to("cuda")

    def cpu(self) -> "SparseConvPackInfo":
        return self.to("cpu")

    # ----------------------------------------------------------------------------
    # properties (``sparse_conv_pack_info.py:272-368``)
    # ----------------------------------------------------------------------------
    @property
    def source_grid(self):
        return self._src

    @property
    def target_grid(self):
        return self._tgt

    @property
    def kernel_size(self) -> tuple:
        return self._ks

    @property
    def stride(self) -> tuple:
        return self._st

    @property
    def neighborhood_map(self):
        return self._pairs_t if self._gs_built else None

    @property
    def neighborhood_sizes(self):
        return self._sizes_t if self._gs_built else None

    @property
    def use_me(self) -> bool:
        return self._use_me

    @property
    def use_tf32(self) -> bool:
        return self._use_tf32

    @property
    def out_in_map(self):
        return self._kmap if self._igemm is not None else None

    @property
    def out_in_map_bwd(self):
        return self._tmap if (self._igemm is not None and self._igemm["training"]) else None

    @property
    def reorder_loc(self):
        return self._sorted_meta.get("reorder_loc")

    @property
    def sorted_mask(self):
        return self._sorted_meta.get("sorted_mask")

    @property
    def reduced_sorted_mask(self):
        return self._sorted_meta.get("reduced_sorted_mask")

    @property
    def reorder_out_in_map(self):
        return self._sorted_meta.get("reorder_out_in_map")

    @property
    def reorder_loc_bwd(self):
        return self._sorted_meta.get("reorder_loc_bwd")

    @property
    def reorder_out_in_map_bwd(self):
        return self._sorted_meta.get("reorder_out_in_map_bwd")

    @property
    def sorted_mask_bwd_w(self):
        return self._sorted_meta.get("sorted_mask_bwd_w")

    @property
    def sorted_mask_bwd_d(self):
        return self._sorted_meta.get("sorted_mask_bwd_d")

    @property
    def halo_index_buffer(self):
        return None

    @property
    def output_index_buffer(self):
        return None

    @property
    def block_kernel_ranges(self):
        return None

    @property
    def block_kernel_in_idx(self):
        return None

    @property
    def block_kernel_rel_out_idx(self):
        return None

    def __repr__(self) -> str:
        return (f"SparseConvPackInfo(kernel_size={self._ks}, stride={self._st}, source={self._src.total_voxels}, "
                f"target={self._tgt.total_voxels})")
