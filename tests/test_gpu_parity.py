"""GPU parity of the gfx950 path against the CPU oracle and the golden fixtures.

Integer results (voxel order, target grids, kernel maps, transposed maps, neighbour
pairs) must be bit-exact.  Floating-point tolerances, stated per test and measured as
max|gpu - ref| / max|ref| against the float64 oracle:

* fp64 on the GPU: 1e-12
* fp32:            1e-5   (the required bound)
* fp16:            2e-3
* bf16:            1e-2   (the required bound)
"""
import math
import os

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
REL_TOL = {torch.float64: 1e-12, torch.float32: 1e-5, torch.float16: 2e-3, torch.bfloat16: 1e-2}


def _fvdb():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb

    return fvdb


def rel_err(a: torch.Tensor, ref: torch.Tensor) -> float:
    a = a.detach().double().cpu()
    ref = ref.detach().double().cpu()
    den = ref.abs().max().item()
    if den == 0:
        return a.abs().max().item() if a.numel() else 0.0
    return (a - ref).abs().max().item() / den


def golden(name):
    z = np.load(os.path.join(GOLDEN, name))
    return {k: z[k] for k in z.files}


GOLDEN_FILES = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def grids_from_fixture(fvdb, z):
    offs = z["src_joffsets"]
    raw = [torch.from_numpy(z["src_ijk_raw"][offs[b]:offs[b + 1]]).cuda() for b in range(len(offs) - 1)]
    src = fvdb.GridBatch.from_ijk(fvdb.JaggedTensor(raw))
    return src


@pytest.mark.parametrize("fname", GOLDEN_FILES)
def test_grid_order_and_kmap_bit_exact(fname):
    fvdb = _fvdb()
    z = golden(fname)
    print(f"[case] {fname}", flush=True)
    src = grids_from_fixture(fvdb, z)
    assert np.array_equal(src.ijk.jdata.cpu().numpy(), z["src_ijk"])
    assert np.array_equal(src.joffsets.cpu().numpy(), z["src_voxel_joffsets"])
    ks = tuple(int(v) for v in z["kernel_size"])
    st = tuple(int(v) for v in z["stride"])
    pack, tgt = src.sparse_conv_kernel_map(ks, st)
    assert np.array_equal(tgt.ijk.jdata.cpu().numpy(), z["tgt_ijk"])
    assert np.array_equal(tgt.joffsets.cpu().numpy(), z["tgt_joffsets"])
    pack.build_implicit_gemm(training=True)
    km = pack.out_in_map
    n_t = tgt.total_voxels
    assert km.shape[0] == (n_t + 127) // 128 * 128
    assert np.array_equal(km[:n_t].cpu().numpy(), z["kmap"])
    assert bool((km[n_t:] == -1).all())
    tm = pack.out_in_map_bwd
    ref_t = oracle.transpose_map(z["kmap"], src.total_voxels)
    assert np.array_equal(tm[:src.total_voxels].cpu().numpy(), ref_t)
    pack.build_gather_scatter()
    nb_ref, sz_ref = oracle.neighbor_map(z["kmap"])
    assert np.array_equal(pack.neighborhood_map.cpu().numpy(), nb_ref)
    assert np.array_equal(pack.neighborhood_sizes.cpu().numpy(), sz_ref)


@pytest.mark.parametrize("fname", GOLDEN_FILES)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.float16, torch.bfloat16])
def test_forward_vs_golden(fname, dtype):
    fvdb = _fvdb()
    z = golden(fname)
    print(f"[case] {fname} {dtype}", flush=True)
    src = grids_from_fixture(fvdb, z)
    ks = tuple(int(v) for v in z["kernel_size"])
    st = tuple(int(v) for v in z["stride"])
    pack, tgt = src.sparse_conv_kernel_map(ks, st)
    pack.build_implicit_gemm()
    x = torch.from_numpy(z["feat"]).to("cuda", dtype)
    w = torch.from_numpy(z["weight"]).to("cuda", dtype)
    out = pack.sparse_conv_3d(src.jagged_like(x), w, fvdb.ConvPackBackend.IGEMM).jdata
    # reference from the rounded inputs, so only the contraction's error is measured
    ref = oracle.sparse_conv(x.double().cpu(), w.double().cpu(), z["kmap"], z["kmap"].shape[0], st)
    err = rel_err(out, ref)
    print(f"  rel err {err:.3e}", flush=True)
    assert err <= REL_TOL[dtype]
    if dtype == torch.float64:
        assert rel_err(out, torch.from_numpy(z["out"])) <= 1e-12


def test_mfma_fragment_layout_exact():
    """Small-integer data: every product and partial sum is exact in fp32, so bf16/f16/f32
    results must equal the fp64 result exactly -- any operand or output lane mix-up fails."""
    fvdb = _fvdb()
    g = torch.Generator().manual_seed(0)
    for cin, cout in [(32, 16), (64, 64), (8, 24), (3, 5), (96, 130)]:
        n = 300
        grid = fvdb.GridBatch.from_ijk(torch.randint(0, 9, (n, 3), generator=g, dtype=torch.int32).cuda())
        pack, _ = grid.sparse_conv_kernel_map(3, 1)
        pack.build_implicit_gemm()
        N = grid.total_voxels
        x = torch.randint(-3, 4, (N, cin), generator=g).double()
        w = torch.randint(-2, 3, (cout, cin, 3, 3, 3), generator=g).double()
        ref = oracle.sparse_conv(x, w, pack.out_in_map[:N].cpu().numpy(), N, 1)
        for dt in (torch.float32, torch.float16, torch.bfloat16):
            print(f"[case] cin={cin} cout={cout} {dt}", flush=True)
            out = pack.sparse_conv_3d(x.to("cuda", dt), w.to("cuda", dt), fvdb.ConvPackBackend.IGEMM).jdata
            if dt == torch.float32 or ref.abs().max() < 256:
                assert torch.equal(out.double().cpu(), ref), f"max diff {(out.double().cpu() - ref).abs().max()}"


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16, torch.float16])
@pytest.mark.parametrize("transposed", [False, True])
def test_backward_vs_oracle(dtype, transposed):
    fvdb = _fvdb()
    z = golden("sphere_k3s2_c4x4.npz") if not transposed else golden("sphere_k3s1_c8.npz")
    print(f"[case] backward {dtype} transposed={transposed}", flush=True)
    src = grids_from_fixture(fvdb, z)
    ks = tuple(int(v) for v in z["kernel_size"])
    st = tuple(int(v) for v in z["stride"])
    pack, tgt = src.sparse_conv_kernel_map(ks, st)
    pack.build_implicit_gemm(training=True)
    torch.manual_seed(3)
    cin, cout = 8, 12
    n_in = tgt.total_voxels if transposed else src.total_voxels
    n_out = src.total_voxels if transposed else tgt.total_voxels
    x = torch.randn(n_in, cin, dtype=torch.float64)
    w = (torch.randn(cin, cout, *ks) if transposed else torch.randn(cout, cin, *ks)).double() * 0.2
    R = torch.randn(n_out, cout, dtype=torch.float64)
    xd = x.to("cuda", dtype).requires_grad_(True)
    wd = w.to("cuda", dtype).requires_grad_(True)
    if transposed:
        out = pack.sparse_transpose_conv_3d(tgt.jagged_like(xd), wd, fvdb.ConvPackBackend.IGEMM).jdata
    else:
        out = pack.sparse_conv_3d(src.jagged_like(xd), wd, fvdb.ConvPackBackend.IGEMM).jdata
    (out * R.to("cuda", dtype)).sum().backward()
    xr, wr = xd.detach().double().cpu(), wd.detach().double().cpu()
    kmap = z["kmap"]
    ref_out = oracle.sparse_conv(xr, wr, kmap, n_out, st, transposed=transposed)
    gi, gw = oracle.sparse_conv_backward(xr, R.to(dtype).double(), wr, kmap, ks, transposed=transposed)
    tol = REL_TOL[dtype]
    e_out, e_gi, e_gw = rel_err(out, ref_out), rel_err(xd.grad, gi), rel_err(wd.grad, gw)
    print(f"  out {e_out:.2e} dgrad {e_gi:.2e} wgrad {e_gw:.2e}", flush=True)
    assert e_out <= tol and e_gi <= tol
    # wgrad of a 16-bit dtype is rounded once more to the dtype at the end
    assert e_gw <= (tol if dtype in (torch.float64, torch.float32) else 2 * tol)


@pytest.mark.parametrize("cin,cout", [(128, 128), (256, 64), (128, 24), (64, 128), (256, 256)])
@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_wide_channels_vs_oracle(cin, cout, dtype):
    """Wide layers (inputs run as (tap, 64-channel chunk) steps, outputs of 128k channels in
    128-channel tiles): forward, dgrad and wgrad against the float64 oracle for a 3x3x3 stride-1
    map (27 taps) and a 2x2x2 stride-2 map (8 taps)."""
    fvdb = _fvdb()
    z = golden("sphere_k3s1_c8.npz")
    src = grids_from_fixture(fvdb, z)
    g = torch.Generator().manual_seed(cin + cout)
    for ks, st in ((3, 1), (2, 2)):
        print(f"[case] wide cin={cin} cout={cout} {dtype} ks={ks} stride={st}", flush=True)
        pack, tgt = src.sparse_conv_kernel_map(ks, st)
        pack.build_implicit_gemm(training=True)
        n_in, n_out = src.total_voxels, tgt.total_voxels
        x = torch.randn(n_in, cin, generator=g, dtype=torch.float64)
        w = torch.randn(cout, cin, ks, ks, ks, generator=g, dtype=torch.float64) / math.sqrt(cin * ks ** 3)
        R = torch.randn(n_out, cout, generator=g, dtype=torch.float64)
        xd = x.to("cuda", dtype).requires_grad_(True)
        wd = w.to("cuda", dtype).requires_grad_(True)
        out = pack.sparse_conv_3d(src.jagged_like(xd), wd, fvdb.ConvPackBackend.IGEMM).jdata
        (out * R.to("cuda", dtype)).sum().backward()
        xr, wr = xd.detach().double().cpu(), wd.detach().double().cpu()
        kmap = pack.out_in_map[:n_out].cpu().numpy()
        ref_out = oracle.sparse_conv(xr, wr, kmap, n_out, (st, st, st))
        gi, gw = oracle.sparse_conv_backward(xr, R.to(dtype).double(), wr, kmap, (ks, ks, ks))
        tol = REL_TOL[dtype]
        e_out, e_gi, e_gw = rel_err(out, ref_out), rel_err(xd.grad, gi), rel_err(wd.grad, gw)
        print(f"  out {e_out:.2e} dgrad {e_gi:.2e} wgrad {e_gw:.2e}", flush=True)
        assert e_out <= tol and e_gi <= tol
        assert e_gw <= (tol if dtype == torch.float32 else 2 * tol)


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("cin,cout", [(128, 128), (256, 128), (128, 256)])
@pytest.mark.parametrize("transposed", [False, True])
def test_wgrad_split_k_vs_oracle(dtype, cin, cout, transposed):
    """16-bit weight gradients with channel counts in multiples of 128 (the split-K kernel with
    rows copied straight into LDS), at chunk sizes that give partials of one round, ragged last
    rounds and long partials, against the oracle's gather-scatter backward in float64 on the
    same rounded inputs: the fp32 partials and their compensated sum keep it within 1e-5."""
    fvdb = _fvdb()
    z = golden("sphere_k3s1_c8.npz")
    src = grids_from_fixture(fvdb, z)
    pack, _ = src.sparse_conv_kernel_map(3, 1)
    pack.build_implicit_gemm(training=True)
    pack.build_gather_scatter()
    pairs, sizes = pack._pairs()
    n = src.total_voxels
    g = torch.Generator().manual_seed(cin + 3 * cout + int(transposed))
    x = torch.randn(n, cin, generator=g).to(dtype)
    r = torch.randn(n, cout, generator=g).to(dtype)
    _, ref = oracle.gs_backward(x.double(), r.double(), torch.zeros(27, cin, cout, dtype=torch.float64),
                                pairs.cpu(), sizes, transposed)
    for chunk in (64, 448, 4096):
        print(f"[case] wgrad {dtype} {cin}->{cout} transposed={transposed} chunk={chunk}", flush=True)
        koff, cstart, total = pack._wgrad_tables(chunk)
        gk = fvdb.ext.wgrad_tabled(x.cuda(), r.cuda(), pairs, koff, cstart, total, transposed, chunk)
        e = rel_err(gk, ref)
        print(f"  rel err {e:.2e}", flush=True)
        assert gk.shape == ref.shape and e <= 1e-5


def dense_box_grid(fvdb, dims):
    return fvdb.GridBatch.from_dense(1, dims, device="cuda")


GRAD_TOL = {torch.float64: 1e-12, torch.float32: 1e-5, torch.float16: 4e-3, torch.bfloat16: 2e-2}


def _build(pack, backend):
    if backend == "gather_scatter":
        pack.build_gather_scatter()
        return "GATHER_SCATTER"
    pack.build_implicit_gemm(sorted=backend == "igemm_sorted", training=True)
    return "IGEMM"


@pytest.mark.parametrize("backend", ["gather_scatter", "igemm", "igemm_sorted"])
@pytest.mark.parametrize("ks", [2, 3, 4, 5, 6])
@pytest.mark.parametrize("stride", [1, 2, 4])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.float16, torch.bfloat16])
def test_dense_conv3d_equivalence(ks, stride, dtype, backend):
    """``tests/unit/test_conv.py:361-461`` on a 32^3 dense grid (4 -> 8 channels): forward, dgrad and
    wgrad against ``conv3d`` and its autograd in float64 on the CPU, fed the same rounded inputs.
    Target voxels outside the dense output's range (the sparse target grid of a strided conv can
    reach one coarse voxel past it) are left out, as the reference crops them."""
    fvdb = _fvdb()
    print(f"[case] dense {backend} ks={ks} stride={stride} {dtype}", flush=True)
    g = torch.Generator().manual_seed(ks * 10 + stride)
    D = 32
    grid = dense_box_grid(fvdb, (D, D, D))
    x = torch.randn(grid.total_voxels, 4, generator=g, dtype=torch.float64).to(dtype)
    w = torch.empty(8, 4, ks, ks, ks, dtype=torch.float64)
    torch.nn.init.kaiming_uniform_(w, a=math.sqrt(5), generator=g)
    w = w.to(dtype)
    pack, tgt = grid.sparse_conv_kernel_map(ks, stride)
    be = getattr(fvdb.ConvPackBackend, _build(pack, backend))
    xd, wd = x.cuda().requires_grad_(True), w.cuda().requires_grad_(True)
    out = pack.sparse_conv_3d(xd, wd, be).jdata
    # dense reference: [C, z, y, x] volume, float64
    ijk = grid.ijk.jdata.long().cpu()
    xr = x.double().requires_grad_(True)
    wr = w.double().requires_grad_(True)
    dense = torch.zeros(4, D, D, D, dtype=torch.float64)
    dense[:, ijk[:, 2], ijk[:, 1], ijk[:, 0]] = xr.t()
    ref = torch.nn.functional.conv3d(dense[None], wr, padding=(ks - 1) // 2, stride=stride)[0]
    t = tgt.ijk.jdata.long().cpu()
    keep = (t[:, 0] < ref.shape[3]) & (t[:, 1] < ref.shape[2]) & (t[:, 2] < ref.shape[1]) & (t >= 0).all(1)
    ref_rows = ref[:, t[keep, 2], t[keep, 1], t[keep, 0]].t()
    tol = REL_TOL[dtype]
    e = rel_err(out.detach().cpu()[keep], ref_rows)
    # gradients of a loss that weighs every kept output row
    R = torch.randn(ref_rows.shape, generator=g, dtype=torch.float64).to(dtype).double()
    (ref_rows * R).sum().backward()
    Rf = torch.zeros(out.shape, dtype=dtype)
    Rf[keep] = R.to(dtype)
    (out * Rf.cuda()).sum().backward()
    e_gi, e_gw = rel_err(xd.grad, xr.grad), rel_err(wd.grad, wr.grad)
    print(f"  out {e:.2e} dgrad {e_gi:.2e} wgrad {e_gw:.2e}", flush=True)
    assert e <= tol and e_gi <= GRAD_TOL[dtype] and e_gw <= GRAD_TOL[dtype]


@pytest.mark.parametrize("backend", ["gather_scatter", "igemm"])
@pytest.mark.parametrize("ks", [2, 3, 4, 5, 6])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.float16, torch.bfloat16])
def test_transposed_vs_conv_transpose3d(ks, dtype, backend):
    """``tests/unit/test_conv.py:579-675``: transposed conv (stride 1) from the target grid back onto
    a 16^3 dense source grid against ``conv_transpose3d`` in float64, forward and gradients."""
    fvdb = _fvdb()
    print(f"[case] transposed {backend} ks={ks} {dtype}", flush=True)
    g = torch.Generator().manual_seed(ks)
    D = 16
    grid = dense_box_grid(fvdb, (D, D, D))
    pack, tgt = grid.sparse_conv_kernel_map(ks, 1)
    be = getattr(fvdb.ConvPackBackend, _build(pack, backend))
    x = torch.randn(tgt.total_voxels, 4, generator=g, dtype=torch.float64).to(dtype)
    w = (torch.randn(4, 6, ks, ks, ks, generator=g, dtype=torch.float64) / math.sqrt(4 * ks ** 3)).to(dtype)
    xd, wd = x.cuda().requires_grad_(True), w.cuda().requires_grad_(True)
    out = pack.sparse_transpose_conv_3d(xd, wd, be).jdata
    ijk = tgt.ijk.jdata.long().cpu()
    xr, wr = x.double().requires_grad_(True), w.double().requires_grad_(True)
    dense = torch.zeros(4, D, D, D, dtype=torch.float64)
    dense[:, ijk[:, 2], ijk[:, 1], ijk[:, 0]] = xr.t()
    ref = torch.nn.functional.conv_transpose3d(dense[None], wr, padding=(ks - 1) // 2)[0][:, :D, :D, :D]
    s = grid.ijk.jdata.long().cpu()
    ref_rows = ref[:, s[:, 2], s[:, 1], s[:, 0]].t()
    R = torch.randn(ref_rows.shape, generator=g, dtype=torch.float64).to(dtype)
    (ref_rows * R.double()).sum().backward()
    (out * R.cuda()).sum().backward()
    e, e_gi, e_gw = rel_err(out, ref_rows), rel_err(xd.grad, xr.grad), rel_err(wd.grad, wr.grad)
    print(f"  out {e:.2e} dgrad {e_gi:.2e} wgrad {e_gw:.2e}", flush=True)
    assert e <= REL_TOL[dtype] and e_gi <= GRAD_TOL[dtype] and e_gw <= GRAD_TOL[dtype]


def test_error_inputs_raise_value_error():
    """``tests/unit/test_conv.py:677-716``."""
    fvdb = _fvdb()
    torch.manual_seed(0)
    grid = fvdb.GridBatch.from_points(fvdb.JaggedTensor(torch.randn(1000, 3, device="cuda")), voxel_sizes=0.025,
                                      origins=[0.0, 0.0, 0.0])
    n = int(grid.num_voxels[0].item())
    for dtype in (torch.float32, torch.bfloat16):
        for transposed in (False, True):
            def run(feats, kernels):
                kmap, _ = grid.sparse_conv_kernel_map(kernels.shape[-1], 1)
                kmap.build_gather_scatter()
                f = kmap.sparse_transpose_conv_3d if transposed else kmap.sparse_conv_3d
                return f(feats, kernels, fvdb.ConvPackBackend.GATHER_SCATTER)

            with pytest.raises(ValueError):
                run(torch.randn(n, device="cuda", dtype=dtype), torch.rand(8, 0, 5, 5, 5, device="cuda", dtype=dtype))
            with pytest.raises(ValueError):
                run(torch.randn(n, 5, device="cuda", dtype=dtype), torch.rand(8, 4, 5, 5, 5, device="cuda", dtype=dtype))
            with pytest.raises(ValueError):
                run(torch.randn(n + 7, 4, device="cuda", dtype=dtype),
                    torch.rand(8, 4, 5, 5, 5, device="cuda", dtype=dtype))


def test_conv_empty_grid():
    """``tests/unit/test_basic_ops.py:1694-1711``: an empty grid gives a [0, C] output."""
    fvdb = _fvdb()
    grid = fvdb.GridBatch.from_zero_voxels(device="cuda")
    pack, tgt = grid.sparse_conv_kernel_map(3, 1)
    pack.build_implicit_gemm()
    out = pack.sparse_conv_3d(torch.empty(0, 4, device="cuda"), torch.randn(8, 4, 3, 3, 3, device="cuda"),
                              fvdb.ConvPackBackend.IGEMM)
    assert tuple(out.jdata.shape) == (0, 8)


def test_batch_invariance_bit_exact():
    """A grid's output rows are the same alone and inside a batch of other grids."""
    fvdb = _fvdb()
    z = golden("batch_k3s1_c4x6.npz")
    g = torch.Generator().manual_seed(5)
    offs = z["src_joffsets"]
    raw = [torch.from_numpy(z["src_ijk_raw"][offs[b]:offs[b + 1]]).cuda() for b in range(len(offs) - 1)]
    batch = fvdb.GridBatch.from_ijk(fvdb.JaggedTensor(raw))
    w = torch.randn(64, 32, 3, 3, 3, generator=g).cuda()
    for dt in (torch.bfloat16, torch.float32):
        x = torch.randn(batch.total_voxels, 32, generator=g).to("cuda", dt)
        pb, _ = batch.sparse_conv_kernel_map(3, 1)
        pb.build_implicit_gemm()
        yb = pb.sparse_conv_3d(batch.jagged_like(x), w.to(dt), fvdb.ConvPackBackend.IGEMM).jdata
        jo = batch.joffsets.cpu().tolist()
        for b in range(batch.grid_count):
            single = fvdb.GridBatch.from_ijk(raw[b])
            ps, _ = single.sparse_conv_kernel_map(3, 1)
            ps.build_implicit_gemm()
            ys = ps.sparse_conv_3d(x[jo[b]:jo[b + 1]], w.to(dt), fvdb.ConvPackBackend.IGEMM).jdata
            assert torch.equal(ys, yb[jo[b]:jo[b + 1]])


def test_finite_difference_fp64():
    """Central differences in fp64 on the GPU against the hand-written dgrad/wgrad."""
    fvdb = _fvdb()
    torch.manual_seed(7)
    rng = np.random.default_rng(7)
    ijk = torch.from_numpy(rng.integers(0, 14, size=(1500, 3)).astype(np.int32)).cuda()
    grid = fvdb.GridBatch.from_ijk(ijk)
    pack, _ = grid.sparse_conv_kernel_map(3, 1)
    pack.build_implicit_gemm(training=True)
    N = grid.total_voxels
    x = torch.randn(N, 8, dtype=torch.float64, device="cuda", requires_grad=True)
    w = (0.2 * torch.randn(8, 8, 3, 3, 3, dtype=torch.float64, device="cuda")).requires_grad_(True)
    R = torch.randn(N, 8, dtype=torch.float64, device="cuda")

    def loss(xx, ww):
        return (pack.sparse_conv_3d(xx, ww, fvdb.ConvPackBackend.IGEMM).jdata * R).sum()

    loss(x, w).backward()
    h = 1e-6
    for idx in [(0, 0), (5, 3), (N - 1, 7), (N // 2, 2)]:
        xp = x.detach().clone(); xp[idx] += h
        xm = x.detach().clone(); xm[idx] -= h
        fd = (loss(xp, w.detach()) - loss(xm, w.detach())).item() / (2 * h)
        assert abs(fd - x.grad[idx].item()) <= 1e-6 * max(1.0, abs(fd))
    for idx in [(0, 0, 0, 0, 0), (3, 5, 1, 2, 0), (7, 7, 2, 2, 2), (1, 6, 1, 1, 1)]:
        wp = w.detach().clone(); wp[idx] += h
        wm = w.detach().clone(); wm[idx] -= h
        fd = (loss(x.detach(), wp) - loss(x.detach(), wm)).item() / (2 * h)
        assert abs(fd - w.grad[idx].item()) <= 1e-6 * max(1.0, abs(fd))


def test_module_and_kmap_cache():
    fvdb = _fvdb()
    torch.manual_seed(0)
    grid = dense_box_grid(fvdb, (8, 8, 8))
    x = fvdb.nn.VDBTensor(grid, grid.jagged_like(torch.randn(grid.total_voxels, 16, device="cuda")))
    c1 = fvdb.nn.SparseConv3d(16, 16, 3).cuda()
    c2 = fvdb.nn.SparseConv3d(16, 32, 3).cuda()
    y = c1(x)
    assert y.kmap is not None and y.grid.is_same(grid)
    z = c2(y)
    assert z.kmap is y.kmap
    down = fvdb.nn.SparseConv3d(32, 32, 2, stride=2).cuda()
    d = down(z)
    assert d.kmap is None and d.grid.total_voxels == 64
    up = fvdb.nn.SparseConv3d(32, 16, 2, stride=2, transposed=True).cuda()
    u = up(d, out_grid=grid)
    assert u.grid.is_same(grid) and tuple(u.data.jdata.shape) == (grid.total_voxels, 16)
    (u.data.jdata.sum()).backward()
    assert c1.weight.grad is not None and up.weight.grad is not None


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16, torch.float32])
@pytest.mark.parametrize("cin,cout", [(64, 64), (32, 128), (20, 24), (128, 64), (64, 128)])
def test_sorted_rows_bit_exact(dtype, cin, cout):
    """Rows taken in tap-mask order (with per-tile tap skipping) give the unsorted result bit for bit,
    forward and backward, for a stride-1 3x3x3 conv and a 2x2x2 stride-2 conv and its transpose."""
    fvdb = _fvdb()
    z = golden("sphere_k3s1_c8.npz")
    g = torch.Generator().manual_seed(11)
    src = grids_from_fixture(fvdb, z)
    for ks, st in ((3, 1), (2, 2)):
        pu, tgt = src.sparse_conv_kernel_map(ks, st)
        pu.build_implicit_gemm(training=True)
        ps, _ = src.sparse_conv_kernel_map(ks, st)
        ps.build_implicit_gemm(sorted=True, training=True)
        # row orders: permutations, masks non-increasing
        for order, kmap, n in ((ps._order_fwd(), ps._map_fwd(), tgt.total_voxels),
                               (ps._order_bwd(), ps._map_bwd(), src.total_voxels)):
            o = order.long().cpu()
            assert torch.equal(torch.sort(o).values, torch.arange(n))
            bits = (kmap[:n] >= 0).long().cpu() << torch.arange(kmap.shape[1]).clamp(max=62)
            m = bits.sum(1)[o]
            assert bool((m[:-1] >= m[1:]).all())
        for transposed in (False, True):
            n_in = tgt.total_voxels if transposed else src.total_voxels
            shape = (cin, cout, ks, ks, ks) if transposed else (cout, cin, ks, ks, ks)
            w = (0.1 * torch.randn(*shape, generator=g)).to("cuda", dtype)
            x = torch.randn(n_in, cin, generator=g).to("cuda", dtype)
            outs = []
            for p in (pu, ps):
                xx = x.clone().requires_grad_(True)
                ww = w.clone().requires_grad_(True)
                f = p.sparse_transpose_conv_3d if transposed else p.sparse_conv_3d
                y = f(xx, ww, fvdb.ConvPackBackend.IGEMM).jdata
                gy = torch.linspace(-1, 1, y.numel(), device="cuda").reshape(y.shape).to(dtype)
                y.backward(gy)
                outs.append((y.detach(), xx.grad, ww.grad))
            for a, b in zip(*outs):
                assert torch.equal(a, b)


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("cout,large", [(64, False), (192, False), (192, True), (256, True)])
@pytest.mark.parametrize("ks,st", [(3, 1), (2, 2)])
def test_direct_lds_tile_bit_exact(dtype, cout, large, ks, st):
    """The direct-to-LDS contraction (16-bit, 64 input channels, 27 or 8 taps, rows in tree
    order, outputs in whole 64-channel tiles) equals the tap-skipping kernel run over an
    identity row order bit for bit (same accumulation order; a skipped tap adds exact zeros),
    with and without the fused bias + ReLU, on a ragged last tile, and the fp64 oracle within
    the 16-bit bound.  The large grid (>= 1024 tiles of 256 rows) takes the 8-wavefront form;
    2x2x2 layers with 128k outputs there take 128-channel workgroups."""
    fvdb = _fvdb()
    from fvdb._lib import ext
    from fvdb.utils.synthetic import sphere_points, voxelize

    n_points, vs = (1_500_000, 1 / 512) if large else (60_000, 1 / 96)
    grid = voxelize(sphere_points(seed=5, n_points=n_points, device="cuda"), vs)
    pack, tgt = grid.sparse_conv_kernel_map(ks, st)
    n_in, n = grid.total_voxels, tgt.total_voxels
    kmap = pack._map_fwd()
    print(f"[direct-lds] {dtype} {ks}^3 s{st} cout={cout} rows={n} (last tile {n % 128} rows)", flush=True)
    g = torch.Generator().manual_seed(21)
    x = torch.randn(n_in, 64, generator=g).to("cuda", dtype)
    w = (0.05 * torch.randn(cout, 64, ks, ks, ks, generator=g)).to("cuda", dtype)
    a = fvdb._autograd.weight_to_kio(w, False).transpose(1, 2).contiguous()  # [K, M, C]
    bias = torch.randn(cout, generator=g).to("cuda", dtype)
    ident = torch.arange(n, dtype=torch.int32, device="cuda")
    for b, relu in ((None, False), (bias, True), (bias, False), (None, True)):
        y = ext.implicit_gemm(x, a, kmap, n, None, b, relu)
        y_skip = ext.implicit_gemm(x, a, kmap, n, ident, b, relu)
        assert torch.equal(y, y_skip), (b is not None, relu)
    if large:
        assert (n + 255) // 256 * (cout // 64) >= 1024, n
        if ks == 2 and cout % 128 == 0:
            assert (n + 255) // 256 * (cout // 128) >= 1024, n
        return
    ref = oracle.sparse_conv(x.double().cpu(), w.double().cpu(), kmap[:n].cpu().numpy(), n, st)
    assert rel_err(ext.implicit_gemm(x, a, kmap, n), ref) < REL_TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("cin,cout", [(64, 64), (64, 128), (128, 128), (128, 64), (128, 192), (64, 320)])
def test_direct_lds_ordered_rows_bit_exact(dtype, cin, cout):
    """The direct-to-LDS contraction over rows in a given order with compacted steps (16-bit
    3x3x3, 64 or 128 input channels, outputs in multiples of 64 channels) equals the tree-order
    result bit for bit -- same accumulation order, and a skipped step adds exact zeros --
    through the global and windowed tap-mask orders, a random permutation and the identity,
    with and without the fused bias + ReLU, on a ragged last tile and with a whole 256-row
    tile of rows without neighbours (a tile with no steps), for the forward and the
    transposed map; and the fp64 oracle within the 16-bit bound."""
    fvdb = _fvdb()
    from fvdb._lib import ext
    from fvdb.utils.synthetic import sphere_points, voxelize

    grid = voxelize(sphere_points(seed=6, n_points=300_000, device="cuda"), 1 / 200)
    pack, _ = grid.sparse_conv_kernel_map(3, 1)
    pack.build_implicit_gemm(training=True)
    n = grid.total_voxels
    assert n % 256 != 0 and n > 1024, n
    print(f"[direct-lds-ordered] {dtype} {cin}->{cout} rows={n} (last tile {n % 256} rows)", flush=True)
    g = torch.Generator().manual_seed(23)
    x = torch.randn(n, cin, generator=g).to("cuda", dtype)
    w = (0.05 * torch.randn(cout, cin, 3, 3, 3, generator=g)).to("cuda", dtype)
    a = fvdb._autograd.weight_to_koi(w, False)   # [K, O, I]
    bias = torch.randn(cout, generator=g).to("cuda", dtype)
    perm = torch.randperm(n, generator=g).to("cuda", torch.int32)
    for m in (pack._map_fwd(), pack._map_bwd()):
        m = m.clone()
        m[256:512] = -1   # rows 256..511 have no neighbour: one whole tile of the identity order
        orders = {"identity": torch.arange(n, dtype=torch.int32, device="cuda"), "random": perm,
                  "global": ext.row_order(m, n, 0), "windowed": ext.row_order(m, n, 2048)}
        for b, relu in ((None, False), (bias, True)):
            y = ext.implicit_gemm(x, a, m, n, None, b, relu)
            assert bool(torch.isfinite(y).all())
            for name, o in orders.items():
                assert torch.equal(ext.implicit_gemm(x, a, m, n, o, b, relu), y), (name, b is not None, relu)
    ref = oracle.sparse_conv(x.double().cpu(), w.double().cpu(), pack._map_fwd()[:n].cpu().numpy(), n, 1)
    y = ext.implicit_gemm(x, a, pack._map_fwd(), n, orders["global"])
    assert rel_err(y, ref) < REL_TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("cin", [32, 64])
@pytest.mark.parametrize("ks,st", [(3, 1), (2, 2)])
def test_resident_weights_bit_exact(dtype, cin, ks, st):
    """The narrow-layer contraction with its weights resident in LDS (16-bit, 32 output
    channels, 32 or 64 input channels, 27 or 8 taps, rows in tree order; persistent
    workgroups whose wavefronts loop over 32-row tiles) equals the tap-skipping kernel over an
    identity row order bit for bit, with and without the fused bias + ReLU, for the forward
    map (coarse rows for 2x2x2 stride 2) and the transposed map (one tap per fine row), on a
    grid with more 32-row tiles than the launch has wavefronts and a ragged last tile; and the
    fp64 oracle within the 16-bit bound."""
    fvdb = _fvdb()
    from fvdb._lib import ext
    from fvdb.utils.synthetic import sphere_points, voxelize

    grid = voxelize(sphere_points(seed=7, n_points=1_500_000, device="cuda"), 1 / 400)
    pack, tgt = grid.sparse_conv_kernel_map(ks, st)
    pack.build_implicit_gemm(training=True)
    n_src, n_tgt = grid.total_voxels, tgt.total_voxels
    g = torch.Generator().manual_seed(24)
    w = (0.05 * torch.randn(32, cin, ks, ks, ks, generator=g)).to("cuda", dtype)
    a = fvdb._autograd.weight_to_koi(w, False)   # [K, 32, cin]
    bias = torch.randn(32, generator=g).to("cuda", dtype)
    for m, n, n_in in ((pack._map_fwd(), n_tgt, n_src), (pack._map_bwd(), n_src, n_tgt)):
        print(f"[resident] {dtype} {ks}^3 s{st} {cin}->32 rows={n} (last tile {n % 32} rows)", flush=True)
        if ks == 3:
            assert (n + 31) // 32 > 256 * 16, n
        x = torch.randn(n_in, cin, generator=g).to("cuda", dtype)
        ident = torch.arange(n, dtype=torch.int32, device="cuda")
        for b, relu in ((None, False), (bias, True)):
            y = ext.implicit_gemm(x, a, m, n, None, b, relu)
            assert torch.equal(ext.implicit_gemm(x, a, m, n, ident, b, relu), y), (b is not None, relu)
    small = voxelize(sphere_points(seed=8, n_points=60_000, device="cuda"), 1 / 96)
    sp, st_grid = small.sparse_conv_kernel_map(ks, st)
    x = torch.randn(small.total_voxels, cin, generator=g).to("cuda", dtype)
    n = st_grid.total_voxels
    ref = oracle.sparse_conv(x.double().cpu(), w.double().cpu(), sp._map_fwd()[:n].cpu().numpy(), n, st)
    assert rel_err(ext.implicit_gemm(x, a, sp._map_fwd(), n), ref) < REL_TOL[dtype]


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float16])
@pytest.mark.parametrize("cout", [256, 192])
def test_direct_lds_wide_bit_exact(dtype, cout):
    """The direct-to-LDS contraction for 128 input channels (steps of one tap and one
    64-channel chunk; 128-channel workgroups for 256 outputs, 64-channel ones for 192) equals the tap-skipping kernel over an
    identity row order bit for bit, with and without the fused bias + ReLU, on a layer
    large enough to take it (>= 1024 workgroups of 256 rows), forward and as dgrad
    (transposed map)."""
    fvdb = _fvdb()
    from fvdb._lib import ext
    from fvdb.utils.synthetic import sphere_points, voxelize

    grid = voxelize(sphere_points(seed=5, n_points=1_500_000, device="cuda"), 1 / 512)
    pack, _ = grid.sparse_conv_kernel_map(3, 1)
    pack.build_implicit_gemm(training=True)
    n = grid.total_voxels
    assert (n + 255) // 256 * (cout // 128 if cout % 128 == 0 else cout // 64) >= 1024, n
    print(f"[direct-lds-wide] {dtype} rows={n} (last tile {n % 256} rows)", flush=True)
    g = torch.Generator().manual_seed(22)
    x = torch.randn(n, 128, generator=g).to("cuda", dtype)
    a = (0.05 * torch.randn(27, cout, 128, generator=g)).to("cuda", dtype)
    bias = torch.randn(cout, generator=g).to("cuda", dtype)
    ident = torch.arange(n, dtype=torch.int32, device="cuda")
    for m in (pack._map_fwd(), pack._map_bwd()):
        for b, relu in ((None, False), (bias, True)):
            y = ext.implicit_gemm(x, a, m, n, None, b, relu)
            y_skip = ext.implicit_gemm(x, a, m, n, ident, b, relu)
            assert torch.equal(y, y_skip), (b is not None, relu)
            assert bool(torch.isfinite(y).all())


def test_row_orders_match_oracle():
    """Global and windowed row orders equal the oracle's (a deterministic sort: descending
    tap mask, ties in row order, windows in place), on one block and on many, and the
    contraction through them equals the unordered one bit for bit."""
    fvdb = _fvdb()
    from fvdb.utils.synthetic import sphere_points, voxelize

    z = golden("sphere_k3s1_c8.npz")
    small = grids_from_fixture(fvdb, z)
    big = voxelize(sphere_points(seed=4, n_points=200_000, device="cuda"), 1 / 160)
    g = torch.Generator().manual_seed(3)
    for grid in (small, big):
        for ks, st in ((3, 1), (2, 2)):
            pack, tgt = grid.sparse_conv_kernel_map(ks, st)
            pack.build_implicit_gemm(training=True)
            # (map, rows, voxels its entries index)
            for kmap, n, n_in in ((pack._map_fwd(), tgt.total_voxels, grid.total_voxels),
                                  (pack._map_bwd(), grid.total_voxels, tgt.total_voxels)):
                km = kmap[:n].cpu().numpy()
                assert int(km.max()) < n_in
                x = torch.randn(n_in, 64, generator=g).to("cuda", torch.bfloat16)
                a = (0.1 * torch.randn(kmap.shape[1], 32, 64, generator=g)).to("cuda", torch.bfloat16)
                ref = fvdb.ext.implicit_gemm(x, a, kmap, n)
                for win in (0, 256, 2048, 8192):
                    o = fvdb.ext.row_order(kmap, n, win)
                    assert np.array_equal(o.cpu().numpy(), oracle.row_order(km, win)), (n, ks, win)
                    assert torch.equal(fvdb.ext.implicit_gemm(x, a, kmap, n, o), ref)
    kmap = small.sparse_conv_kernel_map(3, 1)[0]._map_fwd()
    for bad in (-1, 1, 100, 16384):
        with pytest.raises(ValueError):
            fvdb.ext.row_order(kmap, small.total_voxels, bad)


def test_automatic_row_order_policy():
    """Without a sorted build the pack orders rows by tap mask per window where that pays
    (fp32; 16-bit at 128+ input channels), and the results do not depend on it:
    bit-identical to tree order, forward and backward."""
    fvdb = _fvdb()
    from fvdb.utils.synthetic import sphere_points, voxelize

    grid = voxelize(sphere_points(seed=2, n_points=100_000, device="cuda"), 1 / 128)
    p3, _ = grid.sparse_conv_kernel_map(3, 1)
    p2, coarse = grid.sparse_conv_kernel_map(2, 2)
    p3.build_implicit_gemm(training=True)
    p2.build_implicit_gemm(training=True)
    bf, f32 = torch.bfloat16, torch.float32
    assert p3._order_for(True, 64, bf) is None and p3._order_for(True, 64, torch.float64) is None
    assert p3._order_for(True, 64, f32) is not None and p3._order_for(True, 128, bf) is not None
    assert p2._order_for(True, 64, bf) is None          # coarse rows: 1..8 taps each
    assert p2._order_for(False, 64, bf) is None         # fine rows (one tap each), narrow 16-bit
    from fvdb.sparse_conv_pack_info import wants_row_order
    # 16-bit 3x3x3: 128 inputs, and 64 inputs with outputs in multiples of 128, take a row order
    # (direct-to-LDS kernel with compacted steps); 64 -> 64 / 192 stay in tree order
    assert wants_row_order(27, 128, bf, 128, 300_000) and wants_row_order(27, 128, bf, 64, 300_000)
    assert wants_row_order(27, 64, bf, 128, 300_000) and wants_row_order(27, 64, bf, 256, 1_000)
    assert not wants_row_order(27, 64, bf, 64, 300_000) and not wants_row_order(27, 64, bf, 192, 300_000)
    assert not wants_row_order(8, 64, bf, 128, 300_000) and not wants_row_order(27, 32, bf, 128, 300_000)
    assert wants_row_order(27, 128, f32, 128, 300_000) and wants_row_order(8, 128, bf, 128, 300_000)
    o = p2._order_for(False, 64, f32)                  # fine rows in fp32: windowed order
    assert o is not None
    mb = p2._map_bwd()[:grid.total_voxels].cpu().numpy()
    assert np.array_equal(o.cpu().numpy(), oracle.row_order(mb, 8192))
    # both workgroup sizes of the order kernel (2048- and 8192-row blocks) against the oracle
    for w in (256, 2048, 4096, 8192):
        got = fvdb.ext.row_order(p2._map_bwd(), grid.total_voxels, w)
        assert np.array_equal(got.cpu().numpy(), oracle.row_order(mb, w)), w
    g = torch.Generator().manual_seed(8)
    for dtype in (bf, f32):
        for pack, n_in, transposed, cin, cout in ((p3, grid.total_voxels, False, 128, 64),
                                                  (p2, coarse.total_voxels, True, 64, 32)):
            shape = (cin, cout, *pack.kernel_size) if transposed else (cout, cin, *pack.kernel_size)
            w = (0.1 * torch.randn(*shape, generator=g)).to("cuda", dtype)
            x = torch.randn(n_in, cin, generator=g).to("cuda", dtype)
            outs = []
            for auto in (True, False):
                if not auto:
                    pack._order_for = lambda fwd, c, dt, co=0: None  # instance attribute shadows the method
                try:
                    xx, ww = x.clone().requires_grad_(True), w.clone().requires_grad_(True)
                    f = pack.sparse_transpose_conv_3d if transposed else pack.sparse_conv_3d
                    y = f(xx, ww, fvdb.ConvPackBackend.IGEMM).jdata
                    y.backward(torch.linspace(-1, 1, y.numel(), device="cuda").reshape(y.shape).to(dtype))
                    outs.append((y.detach(), xx.grad, ww.grad))
                finally:
                    pack.__dict__.pop("_order_for", None)
            for a, b in zip(*outs):
                assert torch.equal(a, b)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("cin,cout", [(8, 12), (64, 64), (128, 128)])
def test_fused_bias_relu_epilogue(dtype, cin, cout):
    """act(conv + bias) in the contraction's epilogue equals conv, then + bias, then ReLU:
    exactly in fp64 (forward and all gradients), within one extra rounding otherwise."""
    fvdb = _fvdb()
    z = golden("sphere_k3s1_c8.npz")
    src = grids_from_fixture(fvdb, z)
    pack, _ = src.sparse_conv_kernel_map(3, 1)
    pack.build_implicit_gemm(training=True)
    g = torch.Generator().manual_seed(cin * 7 + cout)
    n = src.total_voxels
    x = torch.randn(n, cin, generator=g, dtype=torch.float64).to("cuda", dtype)
    w = (torch.randn(cout, cin, 3, 3, 3, generator=g, dtype=torch.float64) / math.sqrt(27 * cin)).to("cuda", dtype)
    b = (0.3 * torch.randn(cout, generator=g, dtype=torch.float64)).to("cuda", dtype)
    R = torch.randn(n, cout, generator=g, dtype=torch.float64).to("cuda", dtype)
    outs = []
    mask = None
    for fused in (True, False):
        xx, ww, bb = (t.clone().requires_grad_(True) for t in (x, w, b))
        if fused:
            y = pack.sparse_conv_3d(src.jagged_like(xx), ww, fvdb.ConvPackBackend.IGEMM, bias=bb, relu=True).jdata
            mask = (y > 0).detach()
        else:
            pre = pack.sparse_conv_3d(src.jagged_like(xx), ww, fvdb.ConvPackBackend.IGEMM).jdata + bb
            # fp64: the true composition; otherwise the fused output's ReLU mask, since a
            # value within one rounding of zero may land on either side of the ReLU
            y = torch.relu(pre) if dtype == torch.float64 else pre * mask
        (y * R).sum().backward()
        outs.append((y.detach(), xx.grad, ww.grad, bb.grad))
    for name, a, r in zip(("out", "dgrad", "wgrad", "bgrad"), *outs):
        if dtype == torch.float64:
            assert torch.allclose(a, r, rtol=1e-12, atol=1e-12), name
        else:
            tol = 1e-5 if dtype == torch.float32 else 2e-2
            if name == "out":
                r = torch.relu(r)
            assert rel_err(a, r) <= tol, f"{name}: {rel_err(a, r):.3e}"
    # the module path: SparseConv3d(..., fuse_relu=True) == ReLU(SparseConv3d(...))
    conv = fvdb.nn.SparseConv3d(cin, cout, 3).cuda().to(dtype)
    v = fvdb.nn.VDBTensor(src, src.jagged_like(x))
    a = conv(v, fuse_relu=True).data.jdata
    r = fvdb.nn.ReLU()(conv(v)).data.jdata
    if dtype == torch.float64:
        assert torch.allclose(a, r, rtol=1e-12, atol=1e-12)
    else:
        assert rel_err(a, r) <= (1e-5 if dtype == torch.float32 else 2e-2)
