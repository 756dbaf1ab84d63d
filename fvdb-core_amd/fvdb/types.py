"""Argument normalisation (the subset of ``fvdb/types.py`` the convolution path uses)."""
from __future__ import annotations

from typing import Sequence, Tuple, Union

import numpy as np
import torch

Vec3iOrScalar = Union[int, Sequence[int], torch.Tensor, np.ndarray]
NumericMaxRank1 = Union[int, float, Sequence, torch.Tensor, np.ndarray]


def is_Vec3iOrScalar(x) -> bool:
    if isinstance(x, (int, np.integer)):
        return True
    if isinstance(x, torch.Tensor):
        return x.numel() in (1, 3) and not x.is_floating_point()
    if isinstance(x, (list, tuple, np.ndarray)):
        return len(x) in (1, 3) and all(isinstance(v, (int, np.integer)) for v in np.asarray(x).reshape(-1).tolist())
    return False


def to_vec3i(x, name: str = "value", positive: bool = False, non_negative: bool = False) -> Tuple[int, int, int]:
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().reshape(-1).tolist()
    elif isinstance(x, np.ndarray):
        x = x.reshape(-1).tolist()
    if isinstance(x, (int, np.integer)):
        v = (int(x),) * 3
    elif isinstance(x, (list, tuple)):
        if len(x) == 1:
            v = (int(x[0]),) * 3
        elif len(x) == 3:
            if any(isinstance(e, float) and not float(e).is_integer() for e in x):
                raise TypeError(f"{name} must be integer valued")
            v = tuple(int(e) for e in x)
        else:
            raise ValueError(f"{name} must be a scalar or have 3 entries, got {len(x)}")
    elif isinstance(x, float) and x.is_integer():
        v = (int(x),) * 3
    else:
        raise TypeError(f"{name} must be an int or a sequence of 3 ints, got {type(x)}")
    if positive and any(e <= 0 for e in v):
        raise ValueError(f"{name} must be positive, got {v}")
    if non_negative and any(e < 0 for e in v):
        raise ValueError(f"{name} must be non-negative, got {v}")
    return v  # type: ignore[return-value]


def to_vec3f_batch(x, batch: int, name: str = "value", positive: bool = False) -> torch.Tensor:
    """Broadcast a scalar, a 3-vector or a [B,3] array to a float64 CPU tensor [B,3]."""
    t = torch.as_tensor(x if not isinstance(x, torch.Tensor) else x.detach().cpu(), dtype=torch.float64)
    if t.dim() == 0:
        t = t.reshape(1, 1).expand(batch, 3)
    elif t.dim() == 1:
        if t.numel() == 1:
            t = t.reshape(1, 1).expand(batch, 3)
        elif t.numel() == 3:
            t = t.reshape(1, 3).expand(batch, 3)
        else:
            raise ValueError(f"{name} must broadcast to shape ({batch}, 3)")
    elif t.dim() == 2:
        if t.shape == (1, 3):
            t = t.expand(batch, 3)
        elif t.shape != (batch, 3):
            raise ValueError(f"{name} must broadcast to shape ({batch}, 3), got {tuple(t.shape)}")
    else:
        raise ValueError(f"{name} must broadcast to shape ({batch}, 3)")
    t = t.contiguous().clone()
    if positive and bool((t <= 0).any()):
        raise ValueError(f"{name} must be positive")
    return t


def resolve_device(device, inherit_from=None) -> torch.device:
    if device is None:
        if inherit_from is not None:
            if hasattr(inherit_from, "device"):
                return torch.device(inherit_from.device)
        return torch.device("cpu")
    if isinstance(device, str):
        d = torch.device(device)
        if d.type == "cuda" and d.index is None:
            d = torch.device("cuda", torch.cuda.current_device())
        return d
    return torch.device(device)
