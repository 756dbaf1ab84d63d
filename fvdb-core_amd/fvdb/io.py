"""``save_gridbatch`` / ``load_gridbatch`` with the reference's signatures
(``fvdb/grid_batch.py:2056-2200``), in this package's own container.

The reference writes NanoVDB byte buffers (``.nvdb``) through NanoVDB's I/O.  This build has
no NanoVDB, so it writes a NumPy archive (``numpy.savez``, loaded with ``allow_pickle=False``):
the voxel coordinates of every grid in linear (tree) order, the grid offsets, voxel sizes,
origins and names, and optionally the voxel data.  Loading rebuilds each grid's index on the
device in the same voxel order, so the data rows line up with the grid as saved.  A real
NanoVDB file is recognised by its magic number and rejected with an error.
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from .grid_batch import GridBatch
from .jagged_tensor import JaggedTensor
from .types import resolve_device

MAGIC = "fvdbk-gridbatch-1"
_NANOVDB_MAGIC = (b"NanoVDB0", b"NanoVDB1")
_DTYPES = {"float32": torch.float32, "float64": torch.float64, "float16": torch.float16, "bfloat16": torch.bfloat16,
           "int32": torch.int32, "int64": torch.int64, "uint8": torch.uint8, "bool": torch.bool}


def _names_list(names, name, n: int) -> List[str]:
    if name is not None:
        return [str(name)] * n
    if names is None:
        return [""] * n
    if isinstance(names, str):
        return [names] * n
    names = [str(x) for x in names]
    if len(names) != n:
        raise ValueError(f"names has {len(names)} entries for {n} grids")
    return names


def save_gridbatch(path: str, grid_batch: GridBatch, data: Optional[JaggedTensor] = None,
                   names: Union[Sequence[str], str, None] = None, name: Optional[str] = None,
                   compressed: bool = False, verbose: bool = False) -> None:
    """Save `grid_batch` (and `data`, one row per voxel) to `path`; `compressed` uses zip deflate."""
    if not isinstance(grid_batch, GridBatch):
        raise TypeError("grid_batch must be a GridBatch")
    impl = grid_batch._impl
    rec = {"magic": np.array(MAGIC), "ijk": impl.ijk.cpu().numpy(),
           "joffsets": grid_batch.joffsets.cpu().numpy().astype(np.int64),
           "voxel_sizes": impl.voxel_sizes.double().cpu().numpy(), "origins": impl.origins.double().cpu().numpy(),
           "names": np.array(_names_list(names, name, grid_batch.grid_count), dtype=np.str_)}
    if data is not None:
        d = data.jdata if isinstance(data, JaggedTensor) else data
        if d.shape[0] != grid_batch.total_voxels:
            raise ValueError("data must have one row per voxel of the grid batch")
        dn = str(d.dtype).replace("torch.", "")
        if dn not in _DTYPES:
            raise TypeError(f"unsupported data dtype {d.dtype}")
        raw = d.detach().contiguous().cpu()
        if d.dtype == torch.bfloat16:
            raw = raw.view(torch.int16)  # numpy has no bfloat16: keep the bits
        rec.update({"data": raw.numpy(), "data_dtype": np.array(dn)})
    with open(path, "wb") as f:
        (np.savez_compressed if compressed else np.savez)(f, **rec)
    if verbose:
        print(f"saved {grid_batch.grid_count} grids ({grid_batch.total_voxels} voxels) to {path}")


def load_gridbatch(path: str, *, indices: Optional[List[int]] = None, index: Optional[int] = None,
                   names: Optional[List[str]] = None, name: Optional[str] = None, device="cuda",
                   verbose: bool = False) -> Tuple[GridBatch, JaggedTensor, List[str]]:
    """(grid batch, data, names) of the grids saved in `path`, all or those selected by one of
    `indices`, `index`, `names`, `name`.  Without saved data the data has zero channels."""
    if sum(s is not None for s in (indices, index, names, name)) > 1:
        raise ValueError("Only one of indices, index, names, or name can be specified")
    with open(path, "rb") as f:
        head = f.read(8)
    if head in _NANOVDB_MAGIC:
        raise NotImplementedError(f"{path} is a NanoVDB file: this build reads the archives save_gridbatch writes, "
                                  "not NanoVDB buffers")
    z = np.load(path, allow_pickle=False)
    if "magic" not in z.files or str(z["magic"]) != MAGIC:
        raise ValueError(f"{path} is not a grid batch written by save_gridbatch")
    all_names = [str(s) for s in z["names"]]
    B = len(all_names)
    if index is not None:
        sel = [int(index)]
    elif indices is not None:
        sel = [int(i) for i in indices]
    elif name is not None:
        sel = [i for i, s in enumerate(all_names) if s == name]
    elif names is not None:
        sel = [all_names.index(s) for s in names]
    else:
        sel = list(range(B))
    if not sel or any(i < 0 or i >= B for i in sel):
        raise IndexError(f"grid selection {sel} is out of range for {B} grids")
    dev = resolve_device(device)
    jo = z["joffsets"]
    ijk = torch.from_numpy(z["ijk"])
    parts = [ijk[jo[i]:jo[i + 1]].to(dev) for i in sel]
    vs = torch.from_numpy(z["voxel_sizes"][sel])
    og = torch.from_numpy(z["origins"][sel])
    grid = GridBatch.from_ijk(JaggedTensor(parts), voxel_sizes=vs, origins=og)
    if "data" in z.files:
        raw = torch.from_numpy(z["data"])
        dt = _DTYPES[str(z["data_dtype"])]
        d = raw.view(torch.bfloat16) if dt == torch.bfloat16 else raw
        rows = torch.cat([d[jo[i]:jo[i + 1]] for i in sel]).to(dev)
    else:
        rows = torch.zeros((grid.total_voxels, 0), device=dev)
    if verbose:
        print(f"loaded {len(sel)} grids ({grid.total_voxels} voxels) from {path}")
    return grid, grid.jagged_like(rows), [all_names[i] for i in sel]
