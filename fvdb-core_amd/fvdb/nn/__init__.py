"""Neural-network layers on sparse voxel grids (the convolution path of ``fvdb.nn``)."""
from .modules import (BatchNorm, Dropout, ElementwiseMixin, GELU, GroupNorm, LeakyReLU, Linear, ReLU, Sigmoid, SiLU,
                      SparseConv3d, Tanh, fvnn_module)
from .unet import SparseUNet
from .vdbtensor import VDBTensor, vdbtensor_from_dense

__all__ = ["SparseConv3d", "SparseUNet", "VDBTensor", "vdbtensor_from_dense", "ReLU", "LeakyReLU", "SiLU", "GELU", "Sigmoid",
           "Tanh", "Linear", "Dropout", "BatchNorm", "GroupNorm", "ElementwiseMixin", "fvnn_module"]
