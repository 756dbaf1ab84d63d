"""The weight gradient that allocates nothing (``fvdbk_conv_wgrad_noalloc``) against the one
with a caller workspace (``fvdbk_conv_wgrad``), through the C ABI on device pointers.

Both run the same kernels in the same order, so the gradients must be identical bit for
bit -- on two calls in a row, the second with a smaller input (the library's scratch keeps
the first call's partials, which the second must not see) -- and a call whose partials do
not fit the scratch fails with FVDBK_ENOSPC before any launch.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DT = {torch.float32: 0, torch.float64: 1, torch.float16: 2, torch.bfloat16: 3}


def _setup():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb

    lib = ctypes.CDLL(fvdb._lib.lib_path())
    lib.fvdbk_last_error.restype = ctypes.c_char_p
    lib.fvdbk_conv_wgrad_workspace_size.restype = ctypes.c_size_t
    lib.fvdbk_conv_wgrad_scratch_bytes.restype = ctypes.c_size_t
    return fvdb, lib


def _problem(fvdb, n_points, seed, cin, cout, dtype):
    from fvdb.utils.synthetic import sphere_points, voxelize

    grid = voxelize(sphere_points(seed=seed, n_points=n_points, device="cuda"), 1 / 160)
    pack, _ = grid.sparse_conv_kernel_map(3, 1)
    pack.build_gather_scatter()
    pairs, sizes = pack._pairs()
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.randn((grid.total_voxels, cin), generator=g, device="cuda").to(dtype)
    go = torch.randn((grid.total_voxels, cout), generator=g, device="cuda").to(dtype)
    return pairs.contiguous(), sizes.tolist(), x, go


def _tables(sizes, chunk):
    koff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    cstart = np.concatenate([[0], np.cumsum([-(-s // chunk) for s in sizes])]).astype(np.int32)
    return torch.from_numpy(koff).cuda(), torch.from_numpy(cstart).cuda(), int(cstart[-1])


def _wgrad(lib, noalloc, x, go, pairs, koff, cstart, K, chunk, total, dtype):
    acc = torch.float64 if dtype == torch.float64 else torch.float32
    gw = torch.empty((K, x.shape[1], go.shape[1]), dtype=acc, device="cuda")
    args = [DT[dtype], ctypes.c_void_p(x.data_ptr()), x.shape[1], ctypes.c_void_p(go.data_ptr()), go.shape[1],
            ctypes.c_void_p(pairs.data_ptr()), ctypes.c_void_p(koff.data_ptr()), ctypes.c_void_p(cstart.data_ptr()),
            K, chunk, ctypes.c_int64(total), 0]
    torch.cuda.synchronize()
    if noalloc:
        rc = lib.fvdbk_conv_wgrad_noalloc(*args, ctypes.c_void_p(gw.data_ptr()), None)
    else:
        nbytes = lib.fvdbk_conv_wgrad_workspace_size(DT[dtype], ctypes.c_int64(total), x.shape[1], go.shape[1])
        ws = torch.empty(max(int(nbytes), 8), dtype=torch.uint8, device="cuda")
        rc = lib.fvdbk_conv_wgrad(*args, ctypes.c_void_p(ws.data_ptr()), ctypes.c_size_t(nbytes),
                                  ctypes.c_void_p(gw.data_ptr()), None)
    assert rc == 0, lib.fvdbk_last_error()
    torch.cuda.synchronize()
    return gw


@pytest.mark.parametrize("dtype,cin,cout", [(torch.bfloat16, 128, 128), (torch.float32, 64, 32),
                                            (torch.float16, 32, 64), (torch.float64, 8, 16)])
def test_noalloc_wgrad_equals_workspace_wgrad(dtype, cin, cout):
    fvdb, lib = _setup()
    assert lib.fvdbk_conv_wgrad_scratch_bytes() == 64 << 20
    for n_points, seed in ((400_000, 3), (60_000, 4)):  # the second call is smaller
        pairs, sizes, x, go = _problem(fvdb, n_points, seed, cin, cout, dtype)
        chunk = fvdb._autograd.wgrad_chunk(sum(sizes), dtype, cin, cout)
        koff, cstart, total = _tables(sizes, chunk)
        print(f"[no-alloc] {dtype} {cin}x{cout}: {x.shape[0]} voxels, {sum(sizes)} pairs, {total} chunks",
              flush=True)
        ref = _wgrad(lib, False, x, go, pairs, koff, cstart, len(sizes), chunk, total, dtype)
        got = _wgrad(lib, True, x, go, pairs, koff, cstart, len(sizes), chunk, total, dtype)
        assert torch.equal(got, ref)


def test_noalloc_wgrad_rejects_partials_over_the_scratch():
    fvdb, lib = _setup()
    pairs, sizes, x, go = _problem(fvdb, 60_000, 5, 256, 256, torch.bfloat16)
    chunk = 64  # 256 x 256 fp32 partials of 64-pair chunks: far more than the scratch holds
    koff, cstart, total = _tables(sizes, chunk)
    need = lib.fvdbk_conv_wgrad_workspace_size(3, ctypes.c_int64(total), 256, 256)
    assert need > lib.fvdbk_conv_wgrad_scratch_bytes()
    gw = torch.full((len(sizes), 256, 256), 7.0, device="cuda")
    torch.cuda.synchronize()
    rc = lib.fvdbk_conv_wgrad_noalloc(3, ctypes.c_void_p(x.data_ptr()), 256, ctypes.c_void_p(go.data_ptr()), 256,
                                      ctypes.c_void_p(pairs.data_ptr()), ctypes.c_void_p(koff.data_ptr()),
                                      ctypes.c_void_p(cstart.data_ptr()), len(sizes), chunk, ctypes.c_int64(total), 0,
                                      ctypes.c_void_p(gw.data_ptr()), None)
    assert rc == -4 and b"scratch" in lib.fvdbk_last_error()
    torch.cuda.synchronize()
    assert bool((gw == 7.0).all())  # nothing was launched
