"""Neural-network layers on sparse voxel grids (the convolution path of ``fvdb.nn``)."""
from .modules import (AvgPool, BatchNorm, Dropout, ElementwiseMixin, FillFromGrid, GELU, GroupNorm, LeakyReLU, Linear,
                      MaxPool, ReLU, Sigmoid, SiLU, SparseConv3d, Tanh, UpsamplingNearest, fuse_conv_bn, fvnn_module)
from .unet import SparseUNet
from .vdbtensor import VDBTensor, vdbtensor_from_dense

__all__ = ["SparseConv3d", "MaxPool", "AvgPool", "UpsamplingNearest", "FillFromGrid", "SparseUNet", "VDBTensor", "vdbtensor_from_dense", "ReLU", "LeakyReLU", "SiLU", "GELU", "Sigmoid",
           "Tanh", "Linear", "Dropout", "BatchNorm", "GroupNorm", "ElementwiseMixin", "fvnn_module", "fuse_conv_bn"]
