"""fvdb (MI355X build): sparse 3D convolution on gfx950.

Drop-in for the convolution path of fvdb-core: ``GridBatch``, ``JaggedTensor``,
``SparseConvPackInfo``, ``ConvPackBackend``, ``config``, ``jcat`` and ``fvdb.nn``
(``SparseConv3d``, ``VDBTensor``).  The compute runs in hand-written HIP kernels behind
the C ABI of ``include/fvdbk.h``; there is no CPU compute path.
"""
import torch  # noqa: F401  (loads the HIP runtime the native libraries bind to)

from ._lib import ext
from .jagged_tensor import JaggedTensor, jcat_tensors
from .grid_batch import GridBatch
from .sparse_conv_pack_info import CUTLASS, GATHER_SCATTER, IGEMM, LGGS, ConvPackBackend, SparseConvPackInfo
from . import nn
from .io import load_gridbatch, save_gridbatch


class _Config:
    """Global flags of ``src/fvdb/Config.h:9-24`` (accepted; they do not change the kernels)."""

    enable_ultra_sparse_acceleration: bool = False
    pedantic_error_checking: bool = False


config = _Config()


def jcat(things_to_cat, dim=None):
    """Concatenate GridBatch / JaggedTensor / VDBTensor (``fvdb/__init__.py:98-128``)."""
    if len(things_to_cat) == 0:
        raise ValueError("Cannot concatenate empty list")
    first = things_to_cat[0]
    if isinstance(first, GridBatch):
        if dim is not None:
            raise ValueError("GridBatch concatenation does not support dim argument")
        ijks = [g.ijk for g in things_to_cat]
        cat = jcat_tensors(ijks)
        vs = torch.cat([g._impl.voxel_sizes for g in things_to_cat])
        og = torch.cat([g._impl.origins for g in things_to_cat])
        return GridBatch.from_ijk(cat, voxel_sizes=vs, origins=og)
    if isinstance(first, JaggedTensor):
        return jcat_tensors(things_to_cat, dim)
    if isinstance(first, nn.VDBTensor):
        if dim == 0:
            raise ValueError("VDBTensor concatenation does not support dim=0")
        grid = jcat([t.grid for t in things_to_cat]) if dim is None else first.grid
        return nn.VDBTensor(grid, jcat_tensors([t.data for t in things_to_cat], dim))
    raise TypeError("jcat() can only cat GridBatch, JaggedTensor, or VDBTensor")


__version__ = "0.2.1+mi355x"

__all__ = ["GridBatch", "JaggedTensor", "SparseConvPackInfo", "ConvPackBackend", "config", "jcat", "nn",
           "GATHER_SCATTER", "IGEMM", "CUTLASS", "LGGS", "load_gridbatch", "save_gridbatch"]
