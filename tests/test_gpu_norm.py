"""Group normalization per grid on the native kernels, and batch-norm folding into the
convolution, against float64 references on the CPU.

* ``fvdb.nn.GroupNorm`` against the reference's semantics restated in float64
  (``fvdb/nn/modules.py:407-432``: ``torch.nn.functional.group_norm`` on each grid's [1, C, n]
  view), on uneven batches whose grids span several 1024-row blocks, for 3 to 96 channels,
  with and without affine parameters and with the fused ReLU; forward and the gradients of
  the input, weight and bias.  Tolerances (max |err| / max |ref|): fp64 1e-12, fp32 1e-5,
  fp16 2e-3, bf16 1e-2 (gradients twice that for 16-bit: they are rounded once more).
* The reference's dense check (``tests/unit/test_nn.py:153-188``): a 20^3 dense grid against
  ``torch.nn.GroupNorm`` on the dense tensor.
* A grid's output is bit-identical alone and inside a batch.
* ``fuse_conv_bn``: SparseConv3d followed by BatchNorm in inference mode equals the folded
  convolution.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-12, torch.float32: 1e-5, torch.float16: 2e-3, torch.bfloat16: 1e-2}


def _fvdb():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb

    return fvdb


def rel_err(a, ref) -> float:
    a, ref = a.detach().double().cpu(), ref.detach().double().cpu()
    den = ref.abs().max().item()
    return (a - ref).abs().max().item() / (den if den > 0 else 1.0)


def batch_grid(fvdb, sizes, seed=0):
    g = torch.Generator().manual_seed(seed)
    parts = []
    for n in sizes:
        side = max(2, int(round(n ** (1 / 3) * 1.6)))
        ijk = torch.randint(0, side, (n * 3, 3), generator=g, dtype=torch.int32)
        parts.append(torch.unique(ijk, dim=0)[:n])
    return fvdb.GridBatch.from_ijk(fvdb.JaggedTensor([p.cuda() for p in parts]))


def reference_group_norm(x, counts, G, w, b, eps, relu):
    """float64, per grid: torch's group_norm on the grid's [1, C, n] view."""
    outs, o = [], 0
    for n in counts:
        f = x[o:o + n]
        if n:
            y = torch.nn.functional.group_norm(f.t().reshape(1, f.shape[1], n), G, w, b, eps)
            f = y.reshape(f.shape[1], n).t()
        outs.append(torch.relu(f) if relu else f)
        o += n
    return torch.cat(outs)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.float16, torch.bfloat16])
@pytest.mark.parametrize("C,G", [(64, 8), (96, 32), (3, 1), (16, 16)])
@pytest.mark.parametrize("affine,relu", [(True, False), (False, False), (True, True)])
def test_group_norm_vs_reference(dtype, C, G, affine, relu):
    fvdb = _fvdb()
    grid = batch_grid(fvdb, [5000, 1024, 37, 2500], seed=C + G)
    counts = grid.num_voxels.tolist()
    g = torch.Generator().manual_seed(C * 3 + G)
    x = (torch.randn(grid.total_voxels, C, generator=g, dtype=torch.float64) * 2 + 0.7).to(dtype)
    gn = fvdb.nn.GroupNorm(G, C, affine=affine).cuda().to(dtype)
    if affine:
        with torch.no_grad():
            gn.weight.copy_(torch.randn(C, generator=g).to(dtype))
            gn.bias.copy_(torch.randn(C, generator=g).to(dtype))
    R = torch.randn(grid.total_voxels, C, generator=g, dtype=torch.float64).to(dtype)
    xd = x.cuda().requires_grad_(True)
    y = gn(fvdb.nn.VDBTensor(grid, grid.jagged_like(xd)), fuse_relu=relu).data.jdata
    (y * R.cuda()).sum().backward()
    xr = x.double().requires_grad_(True)
    wr = gn.weight.detach().double().cpu().requires_grad_(True) if affine else None
    br = gn.bias.detach().double().cpu().requires_grad_(True) if affine else None
    ref = reference_group_norm(xr, counts, G, wr, br, gn.eps, relu)
    (ref * R.double()).sum().backward()
    tol = TOL[dtype]
    gtol = tol if dtype in (torch.float64, torch.float32) else 2 * tol
    e = rel_err(y, ref)
    e_dx = rel_err(xd.grad, xr.grad)
    print(f"[group norm] {dtype} C={C} G={G} affine={affine} relu={relu}: out {e:.2e} dx {e_dx:.2e}", flush=True)
    assert e <= tol and e_dx <= gtol
    if affine:
        assert rel_err(gn.weight.grad, wr.grad) <= gtol and rel_err(gn.bias.grad, br.grad) <= gtol


@pytest.mark.parametrize("dtype,C,G", [(torch.float32, 1032, 8), (torch.bfloat16, 2064, 16)])
def test_group_norm_wide_rows(dtype, C, G):
    """Rows of more 16-byte segments than one workgroup covers (258 here): the statistics
    kernels loop over segment chunks and the element-wise kernels take a second grid
    dimension; forward and backward against float64 per grid."""
    fvdb = _fvdb()
    grid = batch_grid(fvdb, [900, 40, 700], seed=G)
    counts = grid.num_voxels.tolist()
    g = torch.Generator().manual_seed(C)
    x = (torch.randn(grid.total_voxels, C, generator=g, dtype=torch.float64) * 1.5 - 0.4).to(dtype)
    gn = fvdb.nn.GroupNorm(G, C).cuda().to(dtype)
    with torch.no_grad():
        gn.weight.copy_(torch.randn(C, generator=g).to(dtype))
        gn.bias.copy_(torch.randn(C, generator=g).to(dtype))
    R = torch.randn(grid.total_voxels, C, generator=g, dtype=torch.float64).to(dtype)
    print(f"[group norm wide] {dtype} C={C} G={G}", flush=True)
    xd = x.cuda().requires_grad_(True)
    y = gn(fvdb.nn.VDBTensor(grid, grid.jagged_like(xd))).data.jdata
    (y * R.cuda()).sum().backward()
    xr = x.double().requires_grad_(True)
    wr = gn.weight.detach().double().cpu().requires_grad_(True)
    br = gn.bias.detach().double().cpu().requires_grad_(True)
    ref = reference_group_norm(xr, counts, G, wr, br, gn.eps, False)
    (ref * R.double()).sum().backward()
    tol = TOL[dtype]
    gtol = tol if dtype == torch.float32 else 2 * tol
    assert rel_err(y, ref) <= tol and rel_err(xd.grad, xr.grad) <= gtol
    assert rel_err(gn.weight.grad, wr.grad) <= gtol and rel_err(gn.bias.grad, br.grad) <= gtol


@pytest.mark.parametrize("dtype", [torch.float32, torch.bfloat16])
def test_group_norm_empty_grids_in_batch(dtype):
    """Grids without voxels inside and at the end of a batch: the per-thread grid walk steps
    over them and the other grids' results equal their results alone."""
    fvdb = _fvdb()
    g = torch.Generator().manual_seed(5)
    parts = []
    for n in (300, 0, 500, 0):
        ijk = torch.unique(torch.randint(0, 12, (n * 3, 3), generator=g, dtype=torch.int32), dim=0)[:n]
        parts.append(ijk.reshape(-1, 3).cuda())
    grid = fvdb.GridBatch.from_ijk(fvdb.JaggedTensor(parts))
    counts = grid.num_voxels.tolist()
    assert counts[1] == 0 and counts[3] == 0
    C, G = 32, 4
    x = (torch.randn(grid.total_voxels, C, generator=g, dtype=torch.float64) + 0.3).to(dtype)
    R = torch.randn(grid.total_voxels, C, generator=g, dtype=torch.float64).to(dtype)
    gn = fvdb.nn.GroupNorm(G, C).cuda().to(dtype)
    xd = x.cuda().requires_grad_(True)
    y = gn(fvdb.nn.VDBTensor(grid, grid.jagged_like(xd))).data.jdata
    (y * R.cuda()).sum().backward()
    xr = x.double().requires_grad_(True)
    ref = reference_group_norm(xr, counts, G, gn.weight.detach().double().cpu(), gn.bias.detach().double().cpu(),
                               gn.eps, False)
    (ref * R.double()).sum().backward()
    tol = TOL[dtype]
    assert rel_err(y, ref) <= tol and rel_err(xd.grad, xr.grad) <= 2 * tol
    jo = grid.joffsets.cpu().tolist()
    for b in (0, 2):
        single = fvdb.GridBatch.from_ijk(grid.ijk.jdata[jo[b]:jo[b + 1]])
        ys = gn(fvdb.nn.VDBTensor(single, single.jagged_like(xd.detach()[jo[b]:jo[b + 1]]))).data.jdata
        assert torch.equal(ys, y.detach()[jo[b]:jo[b + 1]])


def test_group_norm_dense_grid_vs_torch():
    """``tests/unit/test_nn.py:153-188``: 20^3 dense grid, 64/128 channels, 8-64 groups."""
    fvdb = _fvdb()
    grid = fvdb.GridBatch.from_dense(1, [20, 20, 20], voxel_sizes=0.05, device="cuda")
    ijk = grid.ijk.jdata.long().cpu()
    for affine in (True, False):
        for C in (64, 128):
            for G in (8, 16, 32, 64):
                x = torch.randn(grid.total_voxels, C, dtype=torch.float32)
                ours = fvdb.nn.GroupNorm(G, C, affine=affine).cuda()
                xd = x.cuda().requires_grad_(True)
                y = ours(fvdb.nn.VDBTensor(grid, grid.jagged_like(xd))).data.jdata
                y.sum().backward()
                dense = torch.zeros(1, C, 20, 20, 20)
                dense[0, :, ijk[:, 0], ijk[:, 1], ijk[:, 2]] = x.t()
                dense.requires_grad_(True)
                yt = torch.nn.GroupNorm(G, C, affine=affine)(dense)
                yt.sum().backward()
                ref = yt[0, :, ijk[:, 0], ijk[:, 1], ijk[:, 2]].t()
                gref = dense.grad[0, :, ijk[:, 0], ijk[:, 1], ijk[:, 2]].t()
                assert rel_err(y, ref) <= 1e-5
                assert (xd.grad.cpu() - gref).abs().max().item() <= 1e-5


def test_group_norm_batch_invariant():
    fvdb = _fvdb()
    grid = batch_grid(fvdb, [3000, 800, 4100], seed=9)
    jo = grid.joffsets.cpu().tolist()
    x = torch.randn(grid.total_voxels, 32, device="cuda")
    gn = fvdb.nn.GroupNorm(8, 32).cuda()
    yb = gn(fvdb.nn.VDBTensor(grid, grid.jagged_like(x))).data.jdata
    for b in range(grid.grid_count):
        single = fvdb.GridBatch.from_ijk(grid.ijk.jdata[jo[b]:jo[b + 1]])
        ys = gn(fvdb.nn.VDBTensor(single, single.jagged_like(x[jo[b]:jo[b + 1]]))).data.jdata
        assert torch.equal(ys, yb[jo[b]:jo[b + 1]])


@pytest.mark.parametrize("transposed", [False, True])
def test_fuse_conv_bn(transposed):
    fvdb = _fvdb()
    g = torch.Generator().manual_seed(4)
    grid = batch_grid(fvdb, [3000, 1200], seed=4)
    cin, cout = 16, 24
    conv = fvdb.nn.SparseConv3d(cin, cout, 3, transposed=transposed).cuda()
    bn = fvdb.nn.BatchNorm(cout).cuda().eval()
    with torch.no_grad():
        bn.running_mean.copy_(torch.randn(cout, generator=g))
        bn.running_var.copy_(torch.rand(cout, generator=g) + 0.5)
        bn.weight.copy_(torch.randn(cout, generator=g))
        bn.bias.copy_(torch.randn(cout, generator=g))
    x = torch.randn(grid.total_voxels, cin, generator=g).cuda()
    v = fvdb.nn.VDBTensor(grid, grid.jagged_like(x))
    out_grid = grid if transposed else None
    ref = bn(conv(v, out_grid) if transposed else conv(v)).data.jdata
    fused = fvdb.nn.fuse_conv_bn(conv, bn)
    y = (fused(v, out_grid) if transposed else fused(v)).data.jdata
    assert rel_err(y, ref) <= 1e-5
    y_relu = (fused(v, out_grid, fuse_relu=True) if transposed else fused(v, fuse_relu=True)).data.jdata
    assert torch.equal(y_relu, torch.relu(y)) or rel_err(y_relu, torch.relu(ref)) <= 1e-5
    with pytest.raises(ValueError):
        fvdb.nn.fuse_conv_bn(conv, bn.train())
    assert np.isfinite(y.detach().cpu().numpy()).all()
