"""CPU checks of the boundary: the C-ABI library loads and exports every symbol its
header declares, the extension loads, and argument errors map to the reference's
Python exception classes without touching a GPU."""
import ctypes
import os
import re

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fvdbk.h")
LIB = os.path.join(ROOT, "fvdb-core_amd", "fvdb", "libfvdbk.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(fvdbk_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ["fvdbk_kmap_build", "fvdbk_conv_implicit_gemm", "fvdbk_conv_wgrad", "fvdbk_last_error",
              "fvdbk_grid_leaf_keys", "fvdbk_nbmap_emit"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libfvdbk.so not built")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    lib.fvdbk_version.restype = ctypes.c_char_p
    assert lib.fvdbk_version().startswith(b"fvdbk")


def test_library_rejects_bad_arguments_without_gpu():
    if not os.path.exists(LIB):
        pytest.skip("libfvdbk.so not built")
    lib = ctypes.CDLL(LIB)
    lib.fvdbk_last_error.restype = ctypes.c_char_p
    # negative channel count -> FVDBK_EINVAL before any launch
    rc = lib.fvdbk_conv_implicit_gemm(0, None, ctypes.c_int64(0), -1, None, None, ctypes.c_int64(0), 27, 8, None,
                                      0, None)
    assert rc == -1
    assert b"positive" in lib.fvdbk_last_error()
    # wgrad workspace size is plain arithmetic
    lib.fvdbk_conv_wgrad_workspace_size.restype = ctypes.c_size_t
    assert lib.fvdbk_conv_wgrad_workspace_size(0, ctypes.c_int64(10), 32, 64) == 10 * 32 * 64 * 4
    assert lib.fvdbk_conv_wgrad_workspace_size(1, ctypes.c_int64(3), 2, 2) == 3 * 2 * 2 * 8


def test_extension_error_classes_on_cpu_tensors():
    import fvdb

    e = fvdb.ext
    with pytest.raises(ValueError):
        e.implicit_gemm(torch.zeros(3, 4), torch.zeros(1, 2, 4), torch.zeros(3, 1, dtype=torch.int32), 3)
    with pytest.raises(ValueError):
        fvdb.GridBatch.from_ijk(torch.zeros(4, 3, dtype=torch.int32))  # CPU tensors: no CPU path
    with pytest.raises(ValueError):
        e.implicit_gemm_cat2(torch.zeros(3, 2), torch.zeros(3, 2), torch.zeros(1, 2, 4),
                             torch.zeros(3, 1, dtype=torch.int32), 3)


def test_concat_argument_checks():
    from fvdb._autograd import check_conv_args

    w = torch.zeros(8, 6, 3, 3, 3)
    # [3 | 3] = 6 input channels passes every shape check; CPU tensors then fail the device check
    with pytest.raises(ValueError, match="HIP device"):
        check_conv_args(torch.zeros(5, 3), w, 5, False, torch.zeros(5, 3))
    with pytest.raises(ValueError, match="concatenated features"):
        check_conv_args(torch.zeros(5, 3), w, 5, False, torch.zeros(5, 2))  # halves of different widths
    with pytest.raises(ValueError, match="concatenated features"):
        check_conv_args(torch.zeros(5, 3), w, 5, False, torch.zeros(5, 3, dtype=torch.float64))
    with pytest.raises(ValueError, match="input channels"):
        check_conv_args(torch.zeros(5, 3), w, 5, False)  # 3 channels against a 6-channel kernel


def test_package_surface():
    import fvdb

    for name in ["GridBatch", "JaggedTensor", "SparseConvPackInfo", "ConvPackBackend", "config", "jcat"]:
        assert hasattr(fvdb, name)
    assert hasattr(fvdb.nn, "SparseConv3d") and hasattr(fvdb.nn, "VDBTensor")
    assert {b.name for b in fvdb.ConvPackBackend} == {"GATHER_SCATTER", "IGEMM", "CUTLASS", "LGGS"}


def test_module_weight_layout_and_init():
    import fvdb

    torch.manual_seed(0)
    m = fvdb.nn.SparseConv3d(4, 8, (3, 2, 1))
    assert tuple(m.weight.shape) == (8, 4, 3, 2, 1)
    # storage is (k2, k1, k0, Di, Do): the permute to [K, I, O] is a view
    assert m.weight.permute(4, 3, 2, 1, 0).is_contiguous()
    bound = 1 / (4 * 6) ** 0.5
    assert float(m.weight.abs().max()) <= bound and float(m.bias.abs().max()) <= bound
    t = fvdb.nn.SparseConv3d(4, 8, 2, stride=2, transposed=True)
    assert tuple(t.weight.shape) == (4, 8, 2, 2, 2)


def test_jagged_tensor_layout():
    import fvdb

    j = fvdb.JaggedTensor([torch.zeros(3, 2), torch.ones(5, 2)])
    assert j.joffsets.tolist() == [0, 3, 8]
    assert j.jidx.tolist() == [0, 0, 0, 1, 1, 1, 1, 1]
    assert j.num_tensors == 2 and j.element_count == 8
    single = fvdb.JaggedTensor(torch.zeros(4, 2))
    assert single.jidx.numel() == 0 and single.joffsets.tolist() == [0, 4]
    cat = fvdb.jcat([j, j], dim=1)
    assert tuple(cat.jdata.shape) == (8, 4)
