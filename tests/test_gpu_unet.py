"""The U-Net workload under test: sparse networks on the GPU against mathematics-fixed
references.

* ``test_simple_net_sparse_vs_dense``: the reference's ``test_simple_net``
  (``tests/unit/test_nn.py:301-345``): a U-Net of GroupNorm / SparseConv3d / ReLU blocks with
  max pooling and nearest upsampling, on a fully dense grid, against the same network built
  from ``torch.nn`` dense modules -- outputs and every parameter gradient within rel-L2 1e-2
  (the reference's bound), tighter in fp64.
* ``test_sparse_unet_vs_dense_network``: this package's 4-level ``SparseUNet``
  (32/64/128/256: 3x3x3 convs, 2x2x2 stride-2 down, transposed up, ``jcat`` skips, bias and
  ReLU in the epilogue) on dense boxes against ``Conv3d`` / ``ConvTranspose3d`` with the same
  weights, forward and parameter gradients.
* ``test_indoor_unet_vs_oracle``: the same network on an indoor room (the workload's
  generator, small) against a layer-by-layer float64 forward built from the oracle's kernel
  maps, target grids and gather-scatter.
* ``test_unet_workload_full_size``: the workload itself (4 indoor rooms at 2 cm, ~3M voxels):
  each room's output alone equals its rows in the batch bit for bit, and bf16 stays within
  1e-2 of fp32.
"""
import numpy as np
import pytest
import torch
import torch.nn as nn

import oracle

pytestmark = pytest.mark.gpu


def _fvdb():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb

    return fvdb


def rel_l2(a, b) -> float:
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def max_rel(a, b) -> float:
    a, b = a.detach().double().cpu(), b.detach().double().cpu()
    return float((a - b).abs().max() / b.abs().max())


# ---------------------------------------------------------------------------------------------
# the reference's simple net (GroupNorm -> conv -> ReLU blocks, max pool, nearest upsampling)
# ---------------------------------------------------------------------------------------------


class _Block(nn.Sequential):
    """GroupNorm(affine=False) -> conv 3x3x3 (no bias) -> ReLU."""

    def __init__(self, cin, cout, sparse, fvnn):
        super().__init__()
        groups = 8 if cin >= 8 else 1
        if sparse:
            self.add_module("GroupNorm", fvnn.GroupNorm(groups, cin, affine=False))
            self.add_module("Conv", fvnn.SparseConv3d(cin, cout, 3, 1, bias=False))
            self.add_module("ReLU", fvnn.ReLU())
        else:
            self.add_module("GroupNorm", nn.GroupNorm(groups, cin, affine=False))
            self.add_module("Conv", nn.Conv3d(cin, cout, 3, 1, padding=1, bias=False))
            self.add_module("ReLU", nn.ReLU())


class _Double(nn.Sequential):
    def __init__(self, cin, cout, encoder, sparse, fvnn):
        super().__init__()
        if encoder:
            mid = max(cout // 2, cin)
            self.add_module("MaxPool", fvnn.MaxPool(2) if sparse else nn.MaxPool3d(2))
        else:
            mid = cout
        self.add_module("SingleConv1", _Block(cin, mid, sparse, fvnn))
        self.add_module("SingleConv2", _Block(mid, cout, sparse, fvnn))


class _SimpleNet(nn.Module):
    def __init__(self, sparse, fvnn):
        super().__init__()
        self.enc0 = _Double(3, 64, True, sparse, fvnn)
        self.enc1 = _Double(64, 128, True, sparse, fvnn)
        self.enc2 = _Double(128, 256, True, sparse, fvnn)
        self.dec2 = _Double(256, 128, False, sparse, fvnn)
        self.dec1 = _Double(128, 64, False, sparse, fvnn)
        self.dec0 = _Double(64, 1, False, sparse, fvnn)
        self.upsample = fvnn.UpsamplingNearest(2) if sparse else nn.Upsample(scale_factor=2)

    def forward(self, x):
        x = self.enc2(self.enc1(self.enc0(x)))
        x = self.dec2(self.upsample(x))
        x = self.dec1(self.upsample(x))
        return self.dec0(self.upsample(x))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("batch,res", [(1, 16), (2, 32)])
def test_simple_net_sparse_vs_dense(dtype, batch, res):
    fvdb = _fvdb()
    fvnn = fvdb.nn
    print(f"[simple net] {dtype} batch={batch} res={res}", flush=True)
    torch.manual_seed(0)
    dense = _SimpleNet(False, fvnn).double().train()
    sparse = _SimpleNet(True, fvnn).cuda().to(dtype).train()
    with torch.no_grad():
        for name, p in sparse.named_parameters():
            p.copy_(dense.state_dict()[name])
    x = torch.randn(batch, 3, res, res, res, dtype=torch.float64)
    x_d = x.to(dtype).double().requires_grad_(False)
    y_dense = dense(x_d)
    y_dense.sum().backward()
    # [B, C, z, y, x] -> [B, x, y, z, C]
    v = fvnn.vdbtensor_from_dense(x.to(dtype).permute(0, 4, 3, 2, 1).contiguous().cuda())
    y_sparse = sparse(v)
    y_sparse.data.jdata.sum().backward()
    y_sd = y_sparse.to_dense().permute(0, 4, 3, 2, 1)
    bound = 1e-2 if dtype == torch.float32 else 1e-9
    e = rel_l2(y_sd, y_dense)
    print(f"  output rel-L2 {e:.2e}", flush=True)
    assert e < bound
    grads = {n: p.grad for n, p in dense.named_parameters() if p.grad is not None}
    assert grads
    worst = max(rel_l2(dict(sparse.named_parameters())[n].grad, g) for n, g in grads.items())
    print(f"  max parameter-gradient rel-L2 {worst:.2e}", flush=True)
    assert worst < bound


# ---------------------------------------------------------------------------------------------
# the 4-level SparseUNet against a dense network with the same weights
# ---------------------------------------------------------------------------------------------


class _DenseUNet(nn.Module):
    """SparseUNet's structure from torch.nn dense modules (same parameter names)."""

    def __init__(self, sparse_unet):
        super().__init__()
        ch = sparse_unet.channels

        def conv(m):
            if m.transposed:
                d = nn.ConvTranspose3d(m.in_channels, m.out_channels, m.kernel_size, m.stride, bias=m.bias is not None)
            else:
                pad = 1 if m.kernel_size == (3, 3, 3) else 0
                d = nn.Conv3d(m.in_channels, m.out_channels, m.kernel_size, m.stride, padding=pad,
                              bias=m.bias is not None)
            with torch.no_grad():
                d.weight.copy_(m.weight.detach().double().cpu())
                if m.bias is not None:
                    d.bias.copy_(m.bias.detach().double().cpu())
            return d

        self.channels = ch
        self.enc = nn.ModuleList([nn.ModuleList([conv(c) for c in cs]) for cs in sparse_unet.enc])
        self.down = nn.ModuleList([conv(c) for c in sparse_unet.down])
        self.up = nn.ModuleList([conv(c) for c in sparse_unet.up])
        self.dec = nn.ModuleList([nn.ModuleList([conv(c) for c in cs]) for cs in sparse_unet.dec])

    def forward(self, x):
        skips = []
        for l, convs in enumerate(self.enc):
            if l > 0:
                x = torch.relu(self.down[l - 1](x))
            for c in convs:
                x = torch.relu(c(x))
            skips.append(x)
        L = len(self.channels)
        for i, l in enumerate(range(L - 2, -1, -1)):
            x = torch.cat([torch.relu(self.up[i](x)), skips[l]], 1)
            for c in self.dec[i]:
                x = torch.relu(c(x))
        return x


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_sparse_unet_vs_dense_network(dtype):
    fvdb = _fvdb()
    print(f"[sparse unet vs dense] {dtype}", flush=True)
    torch.manual_seed(1)
    model = fvdb.nn.SparseUNet(in_channels=8, channels=(32, 64, 128, 256)).cuda().to(dtype).train()
    dense = _DenseUNet(model).double()
    B, D = 2, 16
    x = torch.randn(B, 8, D, D, D, dtype=torch.float64)
    xr = x.to(dtype).double()
    y_dense = dense(xr)
    y_dense.sum().backward()
    v = fvdb.nn.vdbtensor_from_dense(xr.to(dtype).permute(0, 4, 3, 2, 1).contiguous().cuda())
    y = model(v)
    y.data.jdata.sum().backward()
    y_sd = y.to_dense().permute(0, 4, 3, 2, 1)
    bound = 1e-5 if dtype == torch.float32 else 1e-10
    e = max_rel(y_sd, y_dense)
    print(f"  output max rel {e:.2e}", flush=True)
    assert e < bound
    sp = dict(model.named_parameters())
    worst = 0.0
    for n, p in dense.named_parameters():
        worst = max(worst, rel_l2(sp[n].grad, p.grad))
    print(f"  max parameter-gradient rel-L2 {worst:.2e}", flush=True)
    assert worst < (1e-4 if dtype == torch.float32 else 1e-9)


# ---------------------------------------------------------------------------------------------
# indoor room through the 4-level network vs a layer-by-layer float64 oracle forward
# ---------------------------------------------------------------------------------------------


def oracle_unet_forward(model, ijk: np.ndarray, x: torch.Tensor, dtype) -> torch.Tensor:
    """Float64 forward of SparseUNet from the oracle's target grids, kernel maps and
    gather-scatter (weights and input rounded to `dtype` first)."""
    def w(m):
        return m.weight.detach().to(dtype).double().cpu()

    def b(m):
        return m.bias.detach().to(dtype).double().cpu()

    L = len(model.channels)
    grids = [oracle.sort_unique(ijk)]
    for _ in range(L - 1):
        grids.append(oracle.conv_grid(grids[-1], 2, 2))
    conv_maps = [oracle.kernel_map([g], [g], 3, 1) for g in grids]
    down_maps = [oracle.kernel_map([grids[l]], [grids[l + 1]], 2, 2) for l in range(L - 1)]
    h = x.to(dtype).double().cpu()
    skips = []
    for l in range(L):
        if l > 0:
            m = model.down[l - 1]
            h = torch.relu(oracle.sparse_conv(h, w(m), down_maps[l - 1], len(grids[l]), 2) + b(m))
        for m in model.enc[l]:
            h = torch.relu(oracle.sparse_conv(h, w(m), conv_maps[l], len(grids[l]), 1) + b(m))
        skips.append(h)
    for i, l in enumerate(range(L - 2, -1, -1)):
        m = model.up[i]
        up = torch.relu(oracle.sparse_conv(h, w(m), down_maps[l], len(grids[l]), 2, transposed=True) + b(m))
        h = torch.cat([up, skips[l]], 1)
        for m in model.dec[i]:
            h = torch.relu(oracle.sparse_conv(h, w(m), conv_maps[l], len(grids[l]), 1) + b(m))
    return h


def test_indoor_unet_vs_oracle():
    fvdb = _fvdb()
    from fvdb.utils.synthetic import indoor_ijk

    torch.manual_seed(2)
    model = fvdb.nn.SparseUNet(in_channels=32, channels=(32, 64, 128, 256)).cuda().eval()
    raw = indoor_ijk(9, n_boxes=6, room=(90, 90, 40), device="cuda")
    grid = fvdb.GridBatch.from_ijk(raw, voxel_sizes=0.02)
    ijk = grid.ijk.jdata.cpu().numpy()
    assert np.array_equal(ijk, oracle.sort_unique(raw.cpu().numpy()))
    x = torch.randn(grid.total_voxels, 32, generator=torch.Generator().manual_seed(3))
    with torch.no_grad():
        for dt, tol in ((torch.float32, 1e-5), (torch.bfloat16, 1e-2)):
            m = model.to(dt)
            y = m(fvdb.nn.VDBTensor(grid, grid.jagged_like(x.to("cuda", dt)))).data.jdata
            ref = oracle_unet_forward(m, ijk, x, dt)
            e = max_rel(y, ref)
            print(f"[indoor unet] {grid.total_voxels} voxels {dt}: max rel {e:.2e}", flush=True)
            assert e <= tol
        model.float()


def test_unet_workload_full_size():
    """4 indoor rooms at 2 cm (~3M voxels), the 4-level network in bf16: a room's output is the
    same alone and in the batch (bit for bit), and bf16 is within 1e-2 of fp32."""
    fvdb = _fvdb()
    from fvdb.utils.synthetic import check_root_tile, indoor_ijk

    parts = [indoor_ijk(s, device="cuda") for s in (1, 2, 3, 4)]
    grid = fvdb.GridBatch.from_ijk(fvdb.JaggedTensor(parts), voxel_sizes=0.02)
    check_root_tile(grid.ijk.jdata)
    N = grid.total_voxels
    print(f"[unet workload] {N} voxels {grid.num_voxels.tolist()}", flush=True)
    torch.manual_seed(0)
    model = fvdb.nn.SparseUNet(in_channels=32, channels=(32, 64, 128, 256)).cuda().eval()
    x = torch.randn((N, 32), generator=torch.Generator(device="cuda").manual_seed(5), device="cuda")
    jo = grid.joffsets.cpu().tolist()
    with torch.no_grad():
        y32 = model(fvdb.nn.VDBTensor(grid, grid.jagged_like(x))).data.jdata
        mb = model.to(torch.bfloat16)
        yb = mb(fvdb.nn.VDBTensor(grid, grid.jagged_like(x.bfloat16()))).data.jdata
        e = max_rel(yb, y32)
        print(f"  bf16 vs fp32 max rel {e:.2e}, rel-L2 {rel_l2(yb, y32):.2e}", flush=True)
        assert e <= 1e-2
        for b in range(grid.grid_count):
            g1 = fvdb.GridBatch.from_ijk(parts[b], voxel_sizes=0.02)
            y1 = mb(fvdb.nn.VDBTensor(g1, g1.jagged_like(x[jo[b]:jo[b + 1]].bfloat16()))).data.jdata
            assert torch.equal(y1, yb[jo[b]:jo[b + 1]]), f"room {b} differs alone vs in the batch"
    model.float()
