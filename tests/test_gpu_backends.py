"""Every convolution route of the Python surface, the dense I/O and index queries beside it,
and the raw C ABI, on the GPU against the float64 oracle.

* ``SparseConv3d.backend`` in default / halo / dense / legacy / me / cutlass / lggs /
  igemm_mode0-2 (``fvdb/nn/modules.py:198-236``) and ``GridBatch.sparse_conv_halo``: forward and
  gradients within the fp32 bound of the oracle; halo and dense also within rel-L2 1e-3 of the
  default backend, the reference's own check (``tests/unit/test_nn.py:272-299``).
* ``build_cutlass`` / ``build_lggs`` / ``build_gather_scatter(use_me=True)`` / sorted and split
  implicit-GEMM builds through ``SparseConvPackInfo.sparse_conv_3d``.
* Sorted builds: ``reorder_loc[0][:N]`` is the order the contraction runs, equal to the
  oracle's stable tap-mask order.
* ``write_to_dense_*`` / ``read_from_dense_*`` (``tests/unit/test_accessors.py:22-60``) and
  ``ijk_to_index`` against the oracle's restatements.
* The C ABI (``include/fvdbk.h``) called through ctypes on raw device pointers: kernel map and
  implicit-GEMM contraction against the oracle.
"""
import ctypes
import math

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


def _fvdb():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb

    return fvdb


def rel_err(a, ref) -> float:
    a, ref = a.detach().double().cpu(), ref.detach().double().cpu()
    return (a - ref).abs().max().item() / ref.abs().max().item()


def rel_l2(a, b) -> float:
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def point_grid(fvdb, n_grids=4, n=1024, seed=0):
    """Random points in [0, 2)^3 at voxel size 0.025 (the reference's backend-test grid, shifted
    to non-negative coordinates)."""
    g = torch.Generator().manual_seed(seed)
    pts = [torch.rand(n, 3, generator=g) * 2.0 for _ in range(n_grids)]
    return fvdb.GridBatch.from_points(fvdb.JaggedTensor([p.cuda() for p in pts]), voxel_sizes=0.025,
                                      origins=[0.0] * 3)


def grid_parts(grid):
    jo = grid.joffsets.cpu().tolist()
    ijk = grid.ijk.jdata.cpu().numpy()
    return [ijk[jo[b]:jo[b + 1]] for b in range(grid.grid_count)]


BACKENDS = ["default", "halo", "dense", "legacy", "me", "cutlass", "lggs", "igemm_mode0", "igemm_mode1",
            "igemm_mode2"]


@pytest.mark.parametrize("cin,cout", [(16, 32), (32, 64), (128, 128)])
def test_conv_backends_vs_oracle(cin, cout):
    fvdb = _fvdb()
    grid = point_grid(fvdb)
    N = grid.total_voxels
    parts = grid_parts(grid)
    km = oracle.kernel_map(parts, parts, 3, 1)
    g = torch.Generator().manual_seed(cin + cout)
    x = torch.randn(N, cin, generator=g)
    R = torch.randn(N, cout, generator=g)
    conv = fvdb.nn.SparseConv3d(cin, cout, 3).cuda()
    w, b = conv.weight.detach().double().cpu(), conv.bias.detach().double().cpu()
    ref = oracle.sparse_conv(x.double(), w, km, N, 1) + b
    gi, gw = oracle.sparse_conv_backward(x.double(), R.double(), w, km, (3, 3, 3))
    outs = {}
    for be in BACKENDS:
        print(f"[backend] {be} {cin}->{cout} ({N} voxels)", flush=True)
        conv.backend = be
        conv.zero_grad()
        xx = x.cuda().requires_grad_(True)
        y = conv(fvdb.nn.VDBTensor(grid, grid.jagged_like(xx))).data.jdata
        (y * R.cuda()).sum().backward()
        e_out, e_gi, e_gw = rel_err(y, ref), rel_err(xx.grad, gi), rel_err(conv.weight.grad, gw)
        print(f"  out {e_out:.2e} dgrad {e_gi:.2e} wgrad {e_gw:.2e}", flush=True)
        tol = FP32_TOL if be != "dense" else 1e-4  # the dense route runs torch's conv3d
        assert e_out <= tol and e_gi <= tol and e_gw <= tol, be
        outs[be] = y.detach()
    for be in ("halo", "dense"):
        assert rel_l2(outs[be], outs["default"]) < 1e-3


@pytest.mark.parametrize("variant", [8, 64])
def test_sparse_conv_halo_vs_oracle(variant):
    fvdb = _fvdb()
    grid = point_grid(fvdb, n_grids=2, n=3000, seed=1)
    N = grid.total_voxels
    parts = grid_parts(grid)
    km = oracle.kernel_map(parts, parts, 3, 1)
    g = torch.Generator().manual_seed(variant)
    x = torch.randn(N, 24, generator=g)
    w = torch.randn(40, 24, 3, 3, 3, generator=g) / math.sqrt(24 * 27)
    R = torch.randn(N, 40, generator=g)
    xx, ww = x.cuda().requires_grad_(True), w.cuda().requires_grad_(True)
    y = grid.sparse_conv_halo(grid.jagged_like(xx), ww, variant)
    assert y.num_tensors == grid.grid_count
    (y.jdata * R.cuda()).sum().backward()
    ref = oracle.sparse_conv(x.double(), w.double(), km, N, 1)
    gi, gw = oracle.sparse_conv_backward(x.double(), R.double(), w.double(), km, (3, 3, 3))
    assert rel_err(y.jdata, ref) <= FP32_TOL
    assert rel_err(xx.grad, gi) <= FP32_TOL and rel_err(ww.grad, gw) <= FP32_TOL
    with pytest.raises(NotImplementedError):
        grid.sparse_conv_halo(grid.jagged_like(x.cuda()), torch.randn(40, 24, 5, 5, 5, device="cuda"))


def test_pack_builders_route_to_the_oracle_result():
    """SparseConvPackInfo builders of every backend give the oracle's result (forward and
    transposed), and a sorted build's reorder_loc is the order its contraction runs."""
    fvdb = _fvdb()
    grid = point_grid(fvdb, n_grids=2, n=4000, seed=2)
    N = grid.total_voxels
    parts = grid_parts(grid)
    g = torch.Generator().manual_seed(7)
    x = torch.randn(N, 32, generator=g)
    w = torch.randn(64, 32, 3, 3, 3, generator=g) / math.sqrt(32 * 27)
    km = oracle.kernel_map(parts, parts, 3, 1)
    ref = oracle.sparse_conv(x.double(), w.double(), km, N, 1)
    BE = fvdb.ConvPackBackend

    def run(build, backend):
        pack, _ = grid.sparse_conv_kernel_map(3, 1)
        build(pack)
        return pack, pack.sparse_conv_3d(grid.jagged_like(x.cuda()), w.cuda(), backend).jdata

    cases = [
        ("gather_scatter", lambda p: p.build_gather_scatter(), BE.GATHER_SCATTER),
        ("gather_scatter me", lambda p: p.build_gather_scatter(use_me=True), BE.GATHER_SCATTER),
        ("igemm", lambda p: p.build_implicit_gemm(), BE.IGEMM),
        ("igemm sorted", lambda p: p.build_implicit_gemm(sorted=True, training=True), BE.IGEMM),
        ("igemm sorted split 3", lambda p: p.build_implicit_gemm(sorted=True, split_mask_num=3), BE.IGEMM),
        ("cutlass", lambda p: p.build_cutlass(), BE.CUTLASS),
        ("lggs", lambda p: p.build_lggs(), BE.LGGS),
    ]
    for name, build, backend in cases:
        pack, y = run(build, backend)
        e = rel_err(y, ref)
        print(f"[pack] {name}: rel err {e:.2e}", flush=True)
        assert e <= FP32_TOL, name
        if name == "igemm sorted":
            n = N
            rl = pack.reorder_loc
            assert rl.shape == (1, (n + 127) // 128 * 128)
            assert np.array_equal(rl[0, :n].cpu().numpy(), oracle.row_order(km, 0))
            assert torch.equal(rl[0, :n], pack._order_fwd())
            assert np.array_equal(rl[0, n:].cpu().numpy(), np.arange(n, rl.shape[1]))
            m = pack.sorted_mask[0].long().cpu()
            assert bool((m[:-1] >= m[1:]).all())
            tm = pack.out_in_map_bwd[:n].cpu().numpy()
            assert np.array_equal(pack.reorder_loc_bwd[0, :n].cpu().numpy(), oracle.row_order(tm, 0))
        if name == "igemm sorted split 3":
            assert pack.reorder_loc.shape[0] == 3
            for s in range(3):
                m = pack.sorted_mask[s].long().cpu()
                assert bool((m[:-1] >= m[1:]).all())
    # the transposed conv through a gather-scatter build
    pack, tgt = grid.sparse_conv_kernel_map(2, 2)
    pack.build_gather_scatter()
    tparts = grid_parts(tgt)
    km2 = oracle.kernel_map(parts, tparts, 2, 2)
    xt = torch.randn(tgt.total_voxels, 16, generator=g)
    wt = torch.randn(16, 8, 2, 2, 2, generator=g)
    y = pack.sparse_transpose_conv_3d(tgt.jagged_like(xt.cuda()), wt.cuda(), BE.GATHER_SCATTER).jdata
    assert rel_err(y, oracle.sparse_conv(xt.double(), wt.double(), km2, N, 2, transposed=True)) <= FP32_TOL
    with pytest.raises(RuntimeError):
        pack.sparse_transpose_conv_3d(tgt.jagged_like(xt.cuda()), wt.cuda(), BE.CUTLASS)


# ---------------------------------------------------------------------------------------------
# dense I/O and index queries
# ---------------------------------------------------------------------------------------------


def test_dense_read_write_vs_oracle():
    fvdb = _fvdb()
    g = torch.Generator().manual_seed(3)
    raw = [torch.randint(0, 20, (600, 3), generator=g, dtype=torch.int32),
           torch.randint(-5, 14, (500, 3), generator=g, dtype=torch.int32)]
    grid = fvdb.GridBatch.from_ijk(fvdb.JaggedTensor([r.cuda() for r in raw]))
    parts = grid_parts(grid)
    N = grid.total_voxels
    data = torch.randn(N, 3, generator=g)
    bbox = grid.total_bbox.cpu().numpy()
    full = tuple(int(v) for v in bbox[1] - bbox[0] + 1)
    for min_coord, size in ((None, None), ([2, -3, 1], None), ([[0, 0, 0], [-6, 1, 2]], (9, 30, 12)),
                            ([-7, -7, -7], (5, 5, 5))):
        print(f"[dense] min_coord={min_coord} grid_size={size}", flush=True)
        d = grid.write_to_dense_xyzc(grid.jagged_like(data.cuda()), min_coord=min_coord, grid_size=size)
        lo = np.broadcast_to(np.asarray(bbox[0] if min_coord is None else min_coord).reshape(-1, 3), (2, 3))
        ref = oracle.write_to_dense(parts, data.numpy(), lo, size or full)
        assert d.shape == (2,) + tuple(size or full) + (3,)   # the frame size never depends on min_coord
        assert np.array_equal(d.cpu().numpy(), ref)
        dc = grid.write_to_dense_czyx(grid.jagged_like(data.cuda()), min_coord=min_coord, grid_size=size)
        assert torch.equal(dc, d.permute(0, 4, 3, 2, 1))
        back = grid.read_from_dense_xyzc(d, dense_origins=lo.copy())
        assert np.array_equal(back.jdata.cpu().numpy(), oracle.read_from_dense(parts, ref, lo))
        back_c = grid.read_from_dense_czyx(dc.contiguous(), dense_origins=lo.copy())
        assert torch.equal(back_c.jdata, back.jdata)
    # gradients flow through both directions
    x = data.cuda().requires_grad_(True)
    dense = grid.write_to_dense_xyzc(grid.jagged_like(x))
    grid.read_from_dense_xyzc(dense, dense_origins=bbox[0].tolist()).jdata.sum().backward()
    assert torch.equal(x.grad, torch.ones_like(x))
    empty = fvdb.GridBatch.from_zero_voxels(device="cuda")
    with pytest.raises(RuntimeError):
        empty.write_to_dense_xyzc(torch.zeros(0, 3, device="cuda"))
    with pytest.raises(ValueError):
        grid.write_to_dense_xyzc(torch.zeros(N, device="cuda"))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16, torch.float16, torch.uint8])
@pytest.mark.parametrize("row", [(8,), (3,), (2, 4)])
def test_dense_io_dtypes_layouts(dtype, row):
    """Every element size, rows of 16 bytes (16-byte path) and of other widths, several row
    dimensions, a strided (non-contiguous) dense input and the gradients of both layouts."""
    fvdb = _fvdb()
    g = torch.Generator().manual_seed(11)
    raw = [torch.randint(0, 24, (900, 3), generator=g, dtype=torch.int32),
           torch.randint(-8, 10, (400, 3), generator=g, dtype=torch.int32)]
    grid = fvdb.GridBatch.from_ijk(fvdb.JaggedTensor([r.cuda() for r in raw]))
    parts = grid_parts(grid)
    N = grid.total_voxels
    data = torch.randn((N,) + row, generator=g) * 40
    data = (data.abs() if dtype == torch.uint8 else data).to(dtype)
    lo = np.array([[1, -2, 0], [-9, -3, -1]])
    size = (20, 17, 23)
    print(f"[dense io] dtype={dtype} row={row}", flush=True)
    d = grid.write_to_dense_xyzc(grid.jagged_like(data.cuda()), min_coord=lo.tolist(), grid_size=size)
    flat = data.reshape(N, -1).double().numpy()
    ref = oracle.write_to_dense(parts, flat, lo, size)
    assert d.shape == (2,) + size + row and d.dtype == dtype
    assert np.array_equal(d.reshape(2, *size, -1).double().cpu().numpy(), ref)
    dc = grid.write_to_dense_czyx(grid.jagged_like(data.cuda()), min_coord=lo.tolist(), grid_size=size)
    assert dc.is_contiguous() and torch.equal(dc, d.reshape(2, *size, -1).permute(0, 4, 3, 2, 1))
    want = oracle.read_from_dense(parts, ref, lo)
    back = grid.read_from_dense_xyzc(d, dense_origins=lo.tolist())
    assert back.jdata.shape == (N,) + row
    assert np.array_equal(back.jdata.reshape(N, -1).double().cpu().numpy(), want)
    # a strided view: the xyzc frame seen as czyx without a copy
    view = d.reshape(2, *size, -1).permute(0, 4, 3, 2, 1)
    assert not view.is_contiguous()
    back_v = grid.read_from_dense_czyx(view, dense_origins=lo.tolist())
    assert np.array_equal(back_v.jdata.double().cpu().numpy(), want)
    if dtype.is_floating_point:
        w = torch.randn(dc.shape, generator=g).to(dtype).cuda()
        x = data.cuda().requires_grad_(True)
        (grid.write_to_dense_czyx(grid.jagged_like(x), min_coord=lo.tolist(), grid_size=size) * w).sum().backward()
        gw = grid.read_from_dense_czyx(w, dense_origins=lo.tolist()).jdata
        assert torch.equal(x.grad.reshape(N, -1), gw)
        dd = d.clone().requires_grad_(True)
        r = torch.randn((N,) + row, generator=g).to(dtype).cuda()
        (grid.read_from_dense_xyzc(dd, dense_origins=lo.tolist()).jdata * r).sum().backward()
        want_g = grid.write_to_dense_xyzc(grid.jagged_like(r), min_coord=lo.tolist(), grid_size=size)
        assert torch.equal(dd.grad, want_g)


def test_ijk_to_index_vs_oracle():
    fvdb = _fvdb()
    g = torch.Generator().manual_seed(5)
    raw = [torch.randint(0, 40, (3000, 3), generator=g, dtype=torch.int32),
           torch.randint(10, 50, (2000, 3), generator=g, dtype=torch.int32)]
    grid = fvdb.GridBatch.from_ijk(fvdb.JaggedTensor([r.cuda() for r in raw]))
    parts = grid_parts(grid)
    q = [torch.randint(-2, 52, (4000, 3), generator=g, dtype=torch.int32) for _ in range(2)]
    ref = oracle.ijk_to_index(parts, [t.numpy() for t in q])
    assert (ref >= 0).any() and (ref < 0).any()
    got = grid.ijk_to_index(fvdb.JaggedTensor([t.cuda() for t in q]), cumulative=True).jdata.cpu().numpy()
    assert np.array_equal(got, ref)
    local = grid.ijk_to_index(fvdb.JaggedTensor([t.cuda() for t in q]), cumulative=False).jdata.cpu().numpy()
    base = np.repeat(grid.joffsets.cpu().numpy()[:2], 4000)
    assert np.array_equal(local, np.where(ref >= 0, ref - base, -1))


# ---------------------------------------------------------------------------------------------
# the raw C ABI on device pointers
# ---------------------------------------------------------------------------------------------


class _Grid(ctypes.Structure):
    _fields_ = [("n_voxels", ctypes.c_int64), ("n_leaves", ctypes.c_int64), ("hash_cap", ctypes.c_int64),
                ("n_grids", ctypes.c_int32), ("pad", ctypes.c_int32), ("ijk", ctypes.c_void_p),
                ("leaf_keys", ctypes.c_void_p), ("leaf_masks", ctypes.c_void_p), ("leaf_word_base", ctypes.c_void_p),
                ("hash", ctypes.c_void_p)]


def test_c_abi_kernel_map_and_contraction():
    """``fvdbk_kmap_build`` and ``fvdbk_conv_implicit_gemm`` called through ctypes with device
    pointers (as the integration recipe binds them), on the null stream, against the oracle."""
    fvdb = _fvdb()
    lib = ctypes.CDLL(fvdb._lib.lib_path())
    lib.fvdbk_last_error.restype = ctypes.c_char_p
    grid = point_grid(fvdb, n_grids=2, n=5000, seed=9)
    N = grid.total_voxels
    i = grid._impl
    desc = _Grid(N, i.keys.shape[0], i.hash.shape[0], grid.grid_count, 0, i.ijk.data_ptr(), i.keys.data_ptr(),
                 i.masks.data_ptr(), i.wbase.data_ptr(), i.hash.data_ptr())
    ks = (ctypes.c_int32 * 3)(3, 3, 3)
    st = (ctypes.c_int32 * 3)(1, 1, 1)
    n_pad = (N + 127) // 128 * 128
    kmap = torch.empty((n_pad, 27), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    rc = lib.fvdbk_kmap_build(ctypes.byref(desc), ctypes.byref(desc), ks, st, 0, ctypes.c_void_p(kmap.data_ptr()),
                              ctypes.c_int64(n_pad), None)
    assert rc == 0, lib.fvdbk_last_error()
    torch.cuda.synchronize()
    parts = grid_parts(grid)
    km = oracle.kernel_map(parts, parts, 3, 1)
    assert np.array_equal(kmap[:N].cpu().numpy(), km) and bool((kmap[N:] == -1).all())
    g = torch.Generator().manual_seed(11)
    x = torch.randn(N, 64, generator=g)
    a = torch.randn(27, 48, 64, generator=g) / math.sqrt(27 * 64)   # [K][M][C]
    # the module weight [O][I][k0][k1][k2] whose taps these are: tap k = (k2 * 3 + k1) * 3 + k0
    w = a.reshape(3, 3, 3, 48, 64).permute(3, 4, 2, 1, 0).contiguous()
    codes = {torch.float32: 0, torch.bfloat16: 3}
    for dt, tol in ((torch.float32, FP32_TOL), (torch.bfloat16, 1e-2)):
        xd, ad = x.to("cuda", dt).contiguous(), a.to("cuda", dt).contiguous()
        out = torch.empty((N, 48), dtype=dt, device="cuda")
        torch.cuda.synchronize()
        rc = lib.fvdbk_conv_implicit_gemm(codes[dt], ctypes.c_void_p(xd.data_ptr()), ctypes.c_int64(N), 64,
                                          ctypes.c_void_p(ad.data_ptr()), ctypes.c_void_p(kmap.data_ptr()),
                                          ctypes.c_int64(N), 27, 48, ctypes.c_void_p(out.data_ptr()), 0, None)
        assert rc == 0, lib.fvdbk_last_error()
        torch.cuda.synchronize()
        e = rel_err(out, oracle.sparse_conv(xd.double().cpu(), w.to(dt).double(), km, N, 1))
        print(f"[abi] {dt}: rel err {e:.2e}", flush=True)
        assert e <= tol


def test_save_load_gridbatch_round_trip(tmp_path):
    """``save_gridbatch`` / ``load_gridbatch`` (``fvdb/grid_batch.py:2056-2200``): the grids come back in
    the same voxel order with their voxel sizes, origins, names and data (bf16 bits and fp32), all
    or by index / indices / name / names; a NanoVDB file is rejected."""
    fvdb = _fvdb()
    grid = point_grid(fvdb, n_grids=3, n=2000, seed=12)
    grid = fvdb.GridBatch.from_ijk(grid.ijk, voxel_sizes=[[0.1, 0.1, 0.1], [0.2, 0.2, 0.3], [0.05, 0.05, 0.05]],
                                   origins=[[0.0, 1.0, 2.0], [3.0, 4.0, 5.0], [-1.0, 0.0, 1.0]])
    g = torch.Generator().manual_seed(2)
    for dtype, compressed in ((torch.bfloat16, False), (torch.float32, True)):
        data = torch.randn(grid.total_voxels, 5, generator=g).to("cuda", dtype)
        path = str(tmp_path / f"grids_{dtype}.nvdb")
        fvdb.save_gridbatch(path, grid, grid.jagged_like(data), names=["a", "b", "c"], compressed=compressed)
        g2, d2, names = fvdb.load_gridbatch(path)
        assert names == ["a", "b", "c"]
        assert torch.equal(g2.ijk.jdata, grid.ijk.jdata) and torch.equal(g2.joffsets, grid.joffsets)
        assert torch.equal(d2.jdata, data)
        assert torch.allclose(g2.voxel_sizes, grid.voxel_sizes) and torch.allclose(g2.origins, grid.origins)
        jo = grid.joffsets.cpu().tolist()
        for kw, sel in (({"index": 1}, [1]), ({"indices": [2, 0]}, [2, 0]), ({"name": "c"}, [2]),
                        ({"names": ["b", "a"]}, [1, 0])):
            gs, ds, ns = fvdb.load_gridbatch(path, **kw)
            assert ns == [["a", "b", "c"][i] for i in sel]
            assert torch.equal(gs.ijk.jdata, torch.cat([grid.ijk.jdata[jo[i]:jo[i + 1]] for i in sel]))
            assert torch.equal(ds.jdata, torch.cat([data[jo[i]:jo[i + 1]] for i in sel]))
    # grids only: zero data channels
    path = str(tmp_path / "plain.nvdb")
    fvdb.save_gridbatch(path, grid, name="room")
    g3, d3, n3 = fvdb.load_gridbatch(path)
    assert n3 == ["room"] * 3 and tuple(d3.jdata.shape) == (grid.total_voxels, 0)
    with pytest.raises(ValueError):
        fvdb.load_gridbatch(path, index=0, name="room")
    nano = tmp_path / "real.nvdb"
    nano.write_bytes(b"NanoVDB0" + bytes(64))
    with pytest.raises(NotImplementedError):
        fvdb.load_gridbatch(str(nano))
