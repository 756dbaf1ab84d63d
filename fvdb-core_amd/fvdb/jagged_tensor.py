"""JaggedTensor: a batch of tensors of different lengths stored as one tensor.

Same layout as the reference (``src/fvdb/JaggedTensor.h:16-18,167-172``):

* ``jdata``    [N, *]   the concatenated elements;
* ``joffsets`` int64 [B+1]  start of each list in ``jdata``;
* ``jidx``     int32 [N]  list index of each element (empty when B == 1, as in the
  reference, see ``JaggedTensor.h:596``);
* ``jlidx``    int32 [B, ldim]  list ids (empty for single-level lists).

Only single-level lists (``ldim == 1``) are used on the convolution path.
"""
from __future__ import annotations

from typing import List, Sequence, Union

import torch


def _offsets_from_sizes(sizes: Sequence[int], device) -> torch.Tensor:
    o = torch.zeros(len(sizes) + 1, dtype=torch.int64)
    if len(sizes) > 0:
        o[1:] = torch.cumsum(torch.as_tensor(list(sizes), dtype=torch.int64), 0)
    return o.to(device)


def _jidx_from_offsets(joffsets: torch.Tensor, n: int) -> torch.Tensor:
    B = joffsets.numel() - 1
    if B <= 1:
        return torch.empty(0, dtype=torch.int32, device=joffsets.device)
    sizes = joffsets[1:] - joffsets[:-1]
    return torch.repeat_interleave(torch.arange(B, dtype=torch.int32, device=joffsets.device), sizes, output_size=n)


class JaggedTensor:
    def __init__(self, tensors: Union[torch.Tensor, Sequence[torch.Tensor], None] = None, *, _jdata=None,
                 _joffsets=None, _jidx=None):
        if _jdata is not None:
            self._jdata = _jdata
            self._joffsets = _joffsets
            self._jidx = _jidx if _jidx is not None else _jidx_from_offsets(_joffsets, _jdata.shape[0])
            return
        if isinstance(tensors, torch.Tensor):
            tensors = [tensors]
        if tensors is None or len(tensors) == 0:
            raise ValueError("JaggedTensor needs at least one tensor")
        if any(isinstance(t, (list, tuple)) for t in tensors):
            raise NotImplementedError("nested (ldim > 1) jagged tensors are not used on the convolution path")
        dev = tensors[0].device
        self._jdata = torch.cat(list(tensors), 0) if len(tensors) > 1 else tensors[0]
        self._joffsets = _offsets_from_sizes([t.shape[0] for t in tensors], dev)
        self._jidx = _jidx_from_offsets(self._joffsets, self._jdata.shape[0])

    # ---- constructors ---------------------------------------------------------------
    @staticmethod
    def from_data_and_offsets(data: torch.Tensor, offsets: torch.Tensor) -> "JaggedTensor":
        offsets = offsets.to(device=data.device, dtype=torch.int64)
        return JaggedTensor(_jdata=data, _joffsets=offsets)

    @staticmethod
    def from_data_and_indices(data: torch.Tensor, indices: torch.Tensor, num_tensors: int) -> "JaggedTensor":
        counts = torch.bincount(indices.long(), minlength=num_tensors)
        offsets = torch.zeros(num_tensors + 1, dtype=torch.int64, device=data.device)
        offsets[1:] = torch.cumsum(counts, 0)
        return JaggedTensor(_jdata=data, _joffsets=offsets)

    @staticmethod
    def from_data_indices_and_list_ids(data, indices, list_ids, num_tensors) -> "JaggedTensor":
        return JaggedTensor.from_data_and_indices(data, indices, num_tensors)

    @staticmethod
    def from_data_offsets_and_list_ids(data, offsets, list_ids) -> "JaggedTensor":
        return JaggedTensor.from_data_and_offsets(data, offsets)

    # ---- layout -----------------------------------------------------------------------
    @property
    def jdata(self) -> torch.Tensor:
        return self._jdata

    @jdata.setter
    def jdata(self, value: torch.Tensor) -> None:
        if value.shape[0] != self._jdata.shape[0]:
            raise ValueError("new jdata must have the same number of elements")
        self._jdata = value

    @property
    def joffsets(self) -> torch.Tensor:
        return self._joffsets

    @property
    def jidx(self) -> torch.Tensor:
        return self._jidx

    @property
    def jlidx(self) -> torch.Tensor:
        return torch.empty((0, 1), dtype=torch.int32, device=self._jdata.device)

    @property
    def num_tensors(self) -> int:
        return self._joffsets.numel() - 1

    def num_outer_lists(self) -> int:
        return self.num_tensors

    @property
    def element_count(self) -> int:
        return int(self._jdata.shape[0])

    @property
    def ldim(self) -> int:
        return 1

    @property
    def edim(self) -> int:
        return self._jdata.dim() - 1

    @property
    def eshape(self) -> List[int]:
        return list(self._jdata.shape[1:])

    @property
    def lshape(self) -> List[int]:
        return (self._joffsets[1:] - self._joffsets[:-1]).tolist()

    @property
    def rshape(self):
        return tuple(self._jdata.shape)

    def rsize(self, d: int) -> int:
        return int(self._jdata.shape[d])

    @property
    def device(self) -> torch.device:
        return self._jdata.device

    @property
    def dtype(self) -> torch.dtype:
        return self._jdata.dtype

    @property
    def is_cuda(self) -> bool:
        return self._jdata.is_cuda

    @property
    def is_cpu(self) -> bool:
        return not self._jdata.is_cuda

    @property
    def requires_grad(self) -> bool:
        return self._jdata.requires_grad

    def is_floating_point(self) -> bool:
        return self._jdata.is_floating_point()

    def __len__(self) -> int:
        return self.num_tensors

    def jagged_like(self, data: torch.Tensor) -> "JaggedTensor":
        if data.shape[0] != self._jdata.shape[0]:
            raise ValueError("data must have the same number of elements as this JaggedTensor")
        return JaggedTensor(_jdata=data, _joffsets=self._joffsets, _jidx=self._jidx)

    def unbind(self) -> List[torch.Tensor]:
        o = self._joffsets.tolist()
        return [self._jdata[o[i]:o[i + 1]] for i in range(len(o) - 1)]

    def __getitem__(self, idx) -> "JaggedTensor":
        parts = self.unbind()
        if isinstance(idx, int):
            if idx < 0:
                idx += len(parts)
            return JaggedTensor([parts[idx]])
        if isinstance(idx, slice):
            sel = parts[idx]
            if len(sel) == 0:
                raise IndexError("empty slice of a JaggedTensor")
            return JaggedTensor(sel)
        raise TypeError(f"unsupported index type {type(idx)}")

    # ---- conversions ------------------------------------------------------------------
    def _map(self, fn) -> "JaggedTensor":
        d = fn(self._jdata)
        return JaggedTensor(_jdata=d, _joffsets=self._joffsets.to(d.device), _jidx=self._jidx.to(d.device))

    def to(self, *args, **kwargs) -> "JaggedTensor":
        return self._map(lambda t: t.to(*args, **kwargs))

    def cpu(self) -> "JaggedTensor":
        return self._map(lambda t: t.cpu())

    def cuda(self) -> "JaggedTensor":
        return self._map(lambda t: t.cuda())

    def type(self, dtype) -> "JaggedTensor":
        return self._map(lambda t: t.type(dtype))

    def type_as(self, other) -> "JaggedTensor":
        o = other.jdata if isinstance(other, JaggedTensor) else other
        return self._map(lambda t: t.type_as(o))

    def float(self):
        return self.type(torch.float32)

    def double(self):
        return self.type(torch.float64)

    def int(self):
        return self.type(torch.int32)

    def long(self):
        return self.type(torch.int64)

    def detach(self):
        return self._map(lambda t: t.detach())

    def clone(self):
        return self._map(lambda t: t.clone())

    def requires_grad_(self, flag: bool = True):
        self._jdata.requires_grad_(flag)
        return self

    # ---- arithmetic -------------------------------------------------------------------
    def _binop(self, other, op) -> "JaggedTensor":
        o = other.jdata if isinstance(other, JaggedTensor) else other
        return self.jagged_like(op(self._jdata, o))

    def __add__(self, o): return self._binop(o, lambda a, b: a + b)
    def __radd__(self, o): return self._binop(o, lambda a, b: b + a)
    def __sub__(self, o): return self._binop(o, lambda a, b: a - b)
    def __rsub__(self, o): return self._binop(o, lambda a, b: b - a)
    def __mul__(self, o): return self._binop(o, lambda a, b: a * b)
    def __rmul__(self, o): return self._binop(o, lambda a, b: b * a)
    def __truediv__(self, o): return self._binop(o, lambda a, b: a / b)
    def __floordiv__(self, o): return self._binop(o, lambda a, b: a // b)
    def __mod__(self, o): return self._binop(o, lambda a, b: a % b)
    def __pow__(self, o): return self._binop(o, lambda a, b: a ** b)
    def __gt__(self, o): return self._binop(o, lambda a, b: a > b)
    def __ge__(self, o): return self._binop(o, lambda a, b: a >= b)
    def __lt__(self, o): return self._binop(o, lambda a, b: a < b)
    def __le__(self, o): return self._binop(o, lambda a, b: a <= b)
    def __eq__(self, o): return self._binop(o, lambda a, b: a == b)  # noqa: E704
    def __ne__(self, o): return self._binop(o, lambda a, b: a != b)
    def __neg__(self): return self.jagged_like(-self._jdata)

    __hash__ = object.__hash__

    def abs(self): return self.jagged_like(self._jdata.abs())
    def sqrt(self): return self.jagged_like(self._jdata.sqrt())
    def floor(self): return self.jagged_like(self._jdata.floor())
    def ceil(self): return self.jagged_like(self._jdata.ceil())
    def round(self, decimals: int = 0): return self.jagged_like(self._jdata.round(decimals=decimals))

    def jsum(self, dim: int = 0, keepdim: bool = False) -> "JaggedTensor":
        B = self.num_tensors
        idx = self._jidx.long() if B > 1 else torch.zeros(self._jdata.shape[0], dtype=torch.long,
                                                           device=self._jdata.device)
        out = torch.zeros((B,) + tuple(self._jdata.shape[1:]), dtype=self._jdata.dtype, device=self._jdata.device)
        out.index_add_(0, idx, self._jdata)
        return JaggedTensor([out[i:i + 1] for i in range(B)])

    def __repr__(self) -> str:
        return f"JaggedTensor(num_tensors={self.num_tensors}, jdata={tuple(self._jdata.shape)}, dtype={self.dtype}, device={self.device})"


def jcat_tensors(items: Sequence[JaggedTensor], dim=None) -> JaggedTensor:
    """Concatenate jagged tensors: along the list axis (dim None) or an element dim."""
    if dim is None:
        parts = []
        for j in items:
            parts.extend(j.unbind())
        return JaggedTensor(parts)
    first = items[0]
    for j in items[1:]:
        if j.num_tensors != first.num_tensors or not torch.equal(j.joffsets, first.joffsets):
            raise ValueError("jcat along an element dimension needs identical list layouts")
    d = dim if dim < 0 else dim
    if d == 0:
        # dim 0 concatenates list by list
        per = [j.unbind() for j in items]
        return JaggedTensor([torch.cat([p[i] for p in per], 0) for i in range(first.num_tensors)])
    return first.jagged_like(torch.cat([j.jdata for j in items], d))
