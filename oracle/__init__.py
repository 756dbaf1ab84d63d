"""CPU oracle for fVDB's sparse-convolution path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py``, always as the checker (or the timed CPU baseline),
never as the thing measured on the GPU and never by the ``fvdb`` package itself.

Two halves:

* integer work (voxel order, target grid, kernel map, neighbour map) restated in
  plain C in ``oracle_kmap.c`` and loaded here through ctypes;
* floating-point work: the reference's CPU gather-GEMM-scatter
  (``src/fvdb/detail/ops/convolution/backend/SparseConvolutionKernelMap.cu:417-615``)
  restated with torch CPU ops, run in float64 for accuracy checks or float32 for the
  CPU baseline, together with the weight permutations of
  ``src/fvdb/detail/autograd/SparseConvolutionKernelMap.cpp:19-191``.

Parity pins (each checked in ``tests/test_oracle.py``):

* voxel order: closed form on dense boxes, bijection and per-grid contiguity, and
  the golden fixtures under ``tests/golden``.  The root-tile order of grids that span
  several 4096^3 tiles (negative coordinates) follows the GPU builder's unsigned
  root key; no reference test pins it, so that part is "parity unpinned".
* kernel map: brute force with a Python dict on tiny grids, closed form on dense
  boxes (a shift, -1 outside), invariants (identity centre column at stride 1).
* target grid: brute-force set construction on tiny grids, dense-box shapes.
* gather-scatter forward/backward: dense ``torch.nn.functional.conv3d`` /
  ``conv_transpose3d`` in float64 on dense boxes (the reference's own test strategy,
  ``tests/unit/test_conv.py:361-461``), and a brute-force triple loop on tiny inputs.
* pooling window map, max / average pooling, nearest upsampling, injection:
  ``torch.nn.functional.max_pool3d`` / ``avg_pool3d`` / ``interpolate(mode="nearest")``
  on dense tensors that hold -inf / 0 / nothing at missing voxels, on dense boxes and
  random sparse sets, and set intersection for injection.
* row order by tap mask (global, and per window of consecutive rows): a brute-force
  ``sorted`` with an explicit (window, -mask, row) key on small random maps.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from typing import List, Sequence, Tuple

import numpy as np
import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    """Compile ``liboracle.so`` with the committed Makefile and return its path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        i32p = ctypes.POINTER(ctypes.c_int32)
        lib.oracle_cmp_order.argtypes = [i32p, i32p]
        lib.oracle_cmp_order.restype = ctypes.c_int
        lib.oracle_sort_unique.argtypes = [i32p, ctypes.c_int64]
        lib.oracle_sort_unique.restype = ctypes.c_int64
        lib.oracle_conv_grid_capacity.argtypes = [ctypes.c_int64, i32p, i32p]
        lib.oracle_conv_grid_capacity.restype = ctypes.c_int64
        lib.oracle_conv_grid.argtypes = [i32p, ctypes.c_int64, i32p, i32p, i32p]
        lib.oracle_conv_grid.restype = ctypes.c_int64
        lib.oracle_kernel_map.argtypes = [i32p, ctypes.c_int64, ctypes.c_int64, i32p, ctypes.c_int64,
                                          i32p, i32p, i32p]
        lib.oracle_kernel_map.restype = ctypes.c_int
        lib.oracle_neighbor_map.argtypes = [i32p, ctypes.c_int64, ctypes.c_int32, i32p, i32p]
        lib.oracle_neighbor_map.restype = ctypes.c_int64
        lib.oracle_transpose_map.argtypes = [i32p, ctypes.c_int64, ctypes.c_int32, i32p, ctypes.c_int64]
        lib.oracle_transpose_map.restype = None
        _LIB = lib
    return _LIB


def _p(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _vec3(v) -> np.ndarray:
    if isinstance(v, int):
        v = (v, v, v)
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(3))
    return a


# --------------------------------------------------------------------------------------
# integer half
# --------------------------------------------------------------------------------------


def sort_unique(ijk: np.ndarray) -> np.ndarray:
    """Unique voxels of one grid in NanoVDB OnIndex tree order (``[n, 3]`` int32)."""
    a = np.ascontiguousarray(np.asarray(ijk, dtype=np.int32).reshape(-1, 3)).copy()
    m = _lib().oracle_sort_unique(_p(a), a.shape[0])
    return a[:m].copy()


def cmp_order(a, b) -> int:
    x = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    y = np.ascontiguousarray(np.asarray(b, dtype=np.int32))
    return int(_lib().oracle_cmp_order(_p(x), _p(y)))


def conv_grid(ijk: np.ndarray, kernel_size, stride) -> np.ndarray:
    """Target voxels of one grid for a strided conv (``BuildGridForConv.cu:190-255``)."""
    src = np.ascontiguousarray(np.asarray(ijk, dtype=np.int32).reshape(-1, 3))
    ks, st = _vec3(kernel_size), _vec3(stride)
    cap = _lib().oracle_conv_grid_capacity(src.shape[0], _p(ks), _p(st))
    out = np.empty((max(cap, 1), 3), dtype=np.int32)
    m = _lib().oracle_conv_grid(_p(src), src.shape[0], _p(ks), _p(st), _p(out))
    return out[:m].copy()


def kernel_map(src_grids: Sequence[np.ndarray], tgt_grids: Sequence[np.ndarray], kernel_size, stride) -> np.ndarray:
    """Batched kernel map ``[sum(n_tgt), K]`` int32 (``ConvolutionKernelMap.cu:61-102``).

    ``src_grids[b]`` and ``tgt_grids[b]`` are grid b's voxels in tree order.
    """
    ks, st = _vec3(kernel_size), _vec3(stride)
    K = int(ks.prod())
    n_tgt = sum(int(t.shape[0]) for t in tgt_grids)
    out = np.empty((n_tgt, K), dtype=np.int32)
    src_base = 0
    tgt_base = 0
    for s, t in zip(src_grids, tgt_grids):
        s = np.ascontiguousarray(s, dtype=np.int32)
        t = np.ascontiguousarray(t, dtype=np.int32)
        part = np.empty((t.shape[0], K), dtype=np.int32)
        if t.shape[0] > 0:
            rc = _lib().oracle_kernel_map(_p(s), s.shape[0], src_base, _p(t), t.shape[0], _p(ks), _p(st), _p(part))
            if rc != 0:
                raise MemoryError("oracle kernel map allocation failed")
        out[tgt_base:tgt_base + t.shape[0]] = part
        src_base += s.shape[0]
        tgt_base += t.shape[0]
    return out


def neighbor_map(kmap: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """(nbmap [P,2], nbsizes [K]) from a kernel map (``SparseConvPackInfo.cpp:61-88``)."""
    m = np.ascontiguousarray(kmap, dtype=np.int32)
    n, K = m.shape
    nb = np.empty((max(n * K, 1), 2), dtype=np.int32)
    sizes = np.empty((K,), dtype=np.int32)
    p = _lib().oracle_neighbor_map(_p(m), n, K, _p(nb), _p(sizes))
    return nb[:p].copy(), sizes


def transpose_map(kmap: np.ndarray, n_src: int) -> np.ndarray:
    m = np.ascontiguousarray(kmap, dtype=np.int32)
    n, K = m.shape
    out = np.empty((n_src, K), dtype=np.int32)
    _lib().oracle_transpose_map(_p(m), n, K, _p(out), n_src)
    return out


def kernel_start(kernel_size) -> List[int]:
    ks = _vec3(kernel_size)
    return [int(math.floor(-int(k) / 2.0 + 1.0)) for k in ks]


# --------------------------------------------------------------------------------------
# floating-point half: gather-GEMM-scatter restated with torch CPU ops
# --------------------------------------------------------------------------------------


def weight_to_kio(weight: torch.Tensor, transposed: bool) -> torch.Tensor:
    """[O,I,k0,k1,k2] (or [I,O,k0,k1,k2] when transposed) -> [K, I, O].

    ``autograd/SparseConvolutionKernelMap.cpp:75`` (permute 4,3,2,1,0) and ``:92``
    (permute 4,3,2,0,1).
    """
    if not transposed:
        w = weight.permute(4, 3, 2, 1, 0)
    else:
        w = weight.permute(4, 3, 2, 0, 1)
    return w.reshape(-1, w.shape[3], w.shape[4]).contiguous()


def kio_grad_to_weight(gw: torch.Tensor, kernel_size, transposed: bool) -> torch.Tensor:
    """Inverse layout change for the weight gradient (``autograd/...KernelMap.cpp:170-187``)."""
    ks = _vec3(kernel_size)
    g = gw.reshape(int(ks[2]), int(ks[1]), int(ks[0]), gw.shape[1], gw.shape[2])
    return g.permute(4, 3, 2, 1, 0) if not transposed else g.permute(3, 4, 2, 1, 0)


def gs_forward(in_feat: torch.Tensor, kernel_kio: torch.Tensor, nbmap: torch.Tensor, nbsizes: torch.Tensor,
               n_out: int, transpose: bool, middle_acceleration: bool) -> torch.Tensor:
    """``dispatchSparseConvolutionKernelMap<kCPU>`` (backend/SparseConvolutionKernelMap.cu:417-509).

    Centre tap first as one dense product when the kernel volume is odd, the row
    counts match and the stride is 1; then offsets in ascending order, each as
    gather -> mm -> scatter-add.  Runs in the dtype of ``in_feat``.
    """
    K = kernel_kio.shape[0]
    cout = kernel_kio.shape[2]
    out = torch.zeros((n_out, cout), dtype=in_feat.dtype)
    sizes = [int(x) for x in nbsizes.tolist()]
    flag = False
    if K % 2 == 1 and n_out == in_feat.shape[0] and middle_acceleration:
        flag = True
        torch.mm(in_feat, kernel_kio[K // 2], out=out)
    g_col, s_col = (1, 0) if transpose else (0, 1)
    cur = 0
    nb = nbmap.long()
    for i in range(K):
        n = sizes[i]
        if flag and i == K // 2:
            cur += n
            continue
        if n == 0:
            continue
        pairs = nb[cur:cur + n]
        buf = in_feat.index_select(0, pairs[:, g_col])
        ob = torch.mm(buf, kernel_kio[i])
        out.index_add_(0, pairs[:, s_col], ob)
        cur += n
    return out


def gs_backward(in_feat: torch.Tensor, grad_out: torch.Tensor, kernel_kio: torch.Tensor, nbmap: torch.Tensor,
                nbsizes: torch.Tensor, transpose: bool) -> Tuple[torch.Tensor, torch.Tensor]:
    """``dispatchSparseConvolutionKernelMapGrad<kCPU>`` (backend/SparseConvolutionKernelMap.cu:511-615).

    Returns (grad_in [N_in, I], grad_kernel [K, I, O]).
    """
    K = kernel_kio.shape[0]
    grad_in = torch.zeros_like(in_feat)
    grad_k = torch.zeros_like(kernel_kio)
    sizes = [int(x) for x in nbsizes.tolist()]
    # gather grad_out through column !transpose, input through column transpose
    go_col, in_col = (0, 1) if transpose else (1, 0)
    nb = nbmap.long()
    cur = 0
    for i in range(K):
        n = sizes[i]
        if n == 0:
            continue
        pairs = nb[cur:cur + n]
        gob = grad_out.index_select(0, pairs[:, go_col])
        inb = in_feat.index_select(0, pairs[:, in_col])
        gib = torch.mm(gob, kernel_kio[i].t())
        grad_k[i] = torch.mm(inb.t(), gob)
        grad_in.index_add_(0, pairs[:, in_col], gib)
        cur += n
    return grad_in, grad_k


def sparse_conv(in_feat: torch.Tensor, weight: torch.Tensor, kmap: np.ndarray, n_out: int, stride,
                transposed: bool = False, dtype=torch.float64) -> torch.Tensor:
    """Full reference forward: weight permute + neighbour map + gather-scatter, on CPU.

    ``kmap`` is the non-transposed kernel map (rows = target voxels).  For a
    transposed conv the output has one row per source voxel (``n_out`` = N_src).
    """
    nb, sizes = neighbor_map(kmap)
    w = weight_to_kio(weight.detach().to("cpu", dtype), transposed)
    x = in_feat.detach().to("cpu", dtype)
    middle = tuple(_vec3(stride).tolist()) == (1, 1, 1)
    return gs_forward(x, w, torch.from_numpy(nb), torch.from_numpy(sizes), n_out, transposed, middle)


def sparse_conv_backward(in_feat: torch.Tensor, grad_out: torch.Tensor, weight: torch.Tensor, kmap: np.ndarray,
                         kernel_size, transposed: bool = False, dtype=torch.float64):
    """Reference backward on CPU: returns (grad_in, grad_weight in the module's layout)."""
    nb, sizes = neighbor_map(kmap)
    w = weight_to_kio(weight.detach().to("cpu", dtype), transposed)
    gi, gk = gs_backward(in_feat.detach().to("cpu", dtype), grad_out.detach().to("cpu", dtype), w,
                         torch.from_numpy(nb), torch.from_numpy(sizes), transposed)
    return gi, kio_grad_to_weight(gk, kernel_size, transposed)


# ---------------------------------------------------------------------------------
# pooling, nearest upsampling, injection (the operators beside the convolution)
# ---------------------------------------------------------------------------------

def _index_of(grids: Sequence[np.ndarray]):
    """{(b, i, j, k): cumulative index} over grids given in tree order."""
    d, base = {}, 0
    for b, g in enumerate(grids):
        for n, c in enumerate(np.asarray(g).reshape(-1, 3).tolist()):
            d[(b, c[0], c[1], c[2])] = base + n
        base += len(g)
    return d


def window_map(coarse_grids: Sequence[np.ndarray], fine_grids: Sequence[np.ndarray], pool_factor, stride) -> np.ndarray:
    """map[c][k] = fine index of ``c*stride + (i, j, l)``, k = (i*f1 + j)*f2 + l, or -1.

    Follows the window loops of ``src/fvdb/detail/ops/DownsampleGridMaxPool.cu:40-60``
    (x outermost); pool_factor and stride in (x, y, z) order."""
    f = [int(v) for v in _vec3(pool_factor)]
    s = [int(v) for v in _vec3(stride)]
    fine = _index_of(fine_grids)
    rows = []
    for b, g in enumerate(coarse_grids):
        for c in np.asarray(g).reshape(-1, 3).tolist():
            row = []
            for i in range(f[0]):
                for j in range(f[1]):
                    for l in range(f[2]):
                        row.append(fine.get((b, c[0] * s[0] + i, c[1] * s[1] + j, c[2] * s[2] + l), -1))
            rows.append(row)
    K = f[0] * f[1] * f[2]
    return np.asarray(rows, dtype=np.int32).reshape(-1, K)


def max_pool(fine_data: torch.Tensor, wmap: np.ndarray) -> Tuple[torch.Tensor, torch.Tensor]:
    """(max over each window's voxels, -inf when none; argmax = first voxel holding it)
    (``DownsampleGridMaxPool.cu:17-63,65-114``)."""
    x = fine_data.reshape(fine_data.shape[0], -1)
    n, C = wmap.shape[0], x.shape[1]
    out = torch.full((n, C), -math.inf, dtype=x.dtype)
    arg = torch.full((n, C), -1, dtype=torch.int64)
    for r in range(n):
        for s_ in wmap[r].tolist():
            if s_ < 0:
                continue
            v = x[s_]
            better = v > out[r]
            out[r] = torch.where(better, v, out[r])
            arg[r] = torch.where(better, torch.full_like(arg[r], s_), arg[r])
    return out, arg


def avg_pool(fine_data: torch.Tensor, wmap: np.ndarray, scale: float) -> torch.Tensor:
    """T(sum) * T(scale), the sum accumulated in fp32 (fp64 for fp64) over taps in order
    (``DownsampleGridAvgPool.cu:17-62``)."""
    x = fine_data.reshape(fine_data.shape[0], -1)
    acc = torch.float64 if x.dtype == torch.float64 else torch.float32
    n, C = wmap.shape[0], x.shape[1]
    out = torch.empty((n, C), dtype=x.dtype)
    sc = torch.tensor(scale, dtype=x.dtype).to(acc)
    for r in range(n):
        s_ = torch.zeros(C, dtype=acc)
        for i in wmap[r].tolist():
            if i >= 0:
                s_ = s_ + x[i].to(acc)
        out[r] = (s_.to(x.dtype).to(acc) * sc).to(x.dtype)
    return out


def parent_index(fine_grids: Sequence[np.ndarray], coarse_grids: Sequence[np.ndarray], factor) -> np.ndarray:
    """Index of the coarse voxel floor(ijk / factor) of each fine voxel, or -1
    (``UpsampleGridNearest.cu:34-50``, integer floor division)."""
    f = _vec3(factor)
    coarse = _index_of(coarse_grids)
    out = []
    for b, g in enumerate(fine_grids):
        for c in np.asarray(g).reshape(-1, 3).tolist():
            p = np.floor_divide(np.asarray(c), f).tolist()
            out.append(coarse.get((b, p[0], p[1], p[2]), -1))
    return np.asarray(out, dtype=np.int64)


def inject_index(dst_grids: Sequence[np.ndarray], src_grids: Sequence[np.ndarray]) -> np.ndarray:
    """Source index of each destination voxel present in the source grid, or -1 (``ops/Inject.cu``)."""
    src = _index_of(src_grids)
    out = []
    for b, g in enumerate(dst_grids):
        for c in np.asarray(g).reshape(-1, 3).tolist():
            out.append(src.get((b, c[0], c[1], c[2]), -1))
    return np.asarray(out, dtype=np.int64)


def row_order(kmap: np.ndarray, window: int = 0) -> np.ndarray:
    """Rows ordered by tap mask, descending, ties in row order; with ``window`` > 0 each
    window of that many consecutive rows is ordered on its own and stays in place.

    The global form is the reference's ``reorder_loc`` (``SparseConvPackInfo.cpp:117-128``:
    a descending sort of the per-row bitmask of ``IGEMMBitOperations.cu:14-36``), made
    deterministic by keeping ties in row order.  The bitmask covers the first
    min(K, 31) taps (min(K, 27) in the windowed form).
    """
    kmap = np.asarray(kmap)
    n, K = kmap.shape
    bits = min(K, 27 if window > 0 else 31)
    mask = ((kmap[:, :bits] >= 0).astype(np.int64) << np.arange(bits, dtype=np.int64)).sum(1)
    rows = np.arange(n, dtype=np.int64)
    win = rows // window if window > 0 else np.zeros(n, np.int64)
    return np.lexsort((rows, -mask, win)).astype(np.int32)


# ---------------------------------------------------------------------------------
# voxelisation of points, dense read/write, ijk -> index (the steps either side of the
# convolution that produce and consume its rows)
# ---------------------------------------------------------------------------------

def points_to_ijk(points: np.ndarray, voxel_size, origin) -> Tuple[np.ndarray, np.ndarray]:
    """Voxel coordinates of points and the distance of each pre-rounding coordinate to a
    rounding tie.

    ``ijkForPointsCallback`` (``src/fvdb/detail/ops/BuildGridFromPoints.cu:28-44``) applies the
    primal transform of ``VoxelCoordTransform.h:296-307`` -- scale = 1 / voxel_size and
    translate = -origin / voxel_size, both computed in float64 and used in the point
    precision (``:88-107``) -- as ``p * scale + translate`` in the point precision, then
    rounds with NanoVDB's ``Vec3::round`` (x -> floor(x + 0.5)).  The returned distances
    (float64, per coordinate: |frac(v) - 0.5|) let a caller skip points that sit within an
    ulp or two of a tie, where a fused multiply-add and two roundings may round apart."""
    p = np.asarray(points)
    dt = p.dtype.type
    vs = np.asarray(voxel_size, dtype=np.float64).reshape(-1)
    og = np.asarray(origin, dtype=np.float64).reshape(-1)
    vs = np.broadcast_to(vs, (3,)) if vs.size == 1 else vs
    og = np.broadcast_to(og, (3,)) if og.size == 1 else og
    scale = (1.0 / vs).astype(dt)
    trans = (-og / vs).astype(dt)
    v = (p * scale).astype(dt) + trans
    v = v.astype(dt)
    ijk = np.floor(v.astype(np.float64) + 0.5).astype(np.int32)
    frac = v.astype(np.float64) - np.floor(v.astype(np.float64))
    return ijk, np.abs(frac - 0.5)


def ijk_to_index(grids: Sequence[np.ndarray], queries: Sequence[np.ndarray]) -> np.ndarray:
    """Cumulative index of each query voxel (per grid b, in tree order across grids), or -1
    (``src/fvdb/detail/ops/IjkToIndex.cu``: ``-1`` for inactive voxels)."""
    index = _index_of(grids)
    out = []
    for b, q in enumerate(queries):
        for c in np.asarray(q).reshape(-1, 3).tolist():
            out.append(index.get((b, c[0], c[1], c[2]), -1))
    return np.asarray(out, dtype=np.int64)


def write_to_dense(grids: Sequence[np.ndarray], data: np.ndarray, min_coord: np.ndarray, size) -> np.ndarray:
    """Dense [B, X, Y, Z, C] array: row n of grid b lands at ijk - min_coord[b] when inside
    [0, size); zeros elsewhere (``autograd/ReadIntoDense.cpp:20-80`` with
    ``ops/ReadIntoDense.cu``)."""
    B = len(grids)
    C = data.shape[1]
    out = np.zeros((B,) + tuple(int(s) for s in size) + (C,), dtype=data.dtype)
    base = 0
    for b, g in enumerate(grids):
        g = np.asarray(g).reshape(-1, 3)
        for n, c in enumerate(g.tolist()):
            r = [c[d] - int(min_coord[b][d]) for d in range(3)]
            if all(0 <= r[d] < size[d] for d in range(3)):
                out[b, r[0], r[1], r[2]] = data[base + n]
        base += len(g)
    return out


def read_from_dense(grids: Sequence[np.ndarray], dense: np.ndarray, origins: np.ndarray) -> np.ndarray:
    """Rows [N, C] read at ijk - origins[b] from a dense [B, X, Y, Z, C] array, zeros
    outside it (``autograd/ReadFromDense.cpp:18-60`` with ``ops/ReadFromDense.cu``)."""
    rows = []
    size = dense.shape[1:4]
    for b, g in enumerate(grids):
        for c in np.asarray(g).reshape(-1, 3).tolist():
            r = [c[d] - int(origins[b][d]) for d in range(3)]
            if all(0 <= r[d] < size[d] for d in range(3)):
                rows.append(dense[b, r[0], r[1], r[2]])
            else:
                rows.append(np.zeros(dense.shape[4:], dtype=dense.dtype))
    return np.stack(rows) if rows else np.zeros((0,) + dense.shape[4:], dtype=dense.dtype)
