"""The contraction over a channel concatenation [a | b] read from both arrays
(``fvdbk_conv_implicit_gemm_cat2``, ``SparseConv3d(..., concat=b)``) against the same
contraction of ``torch.cat([a, b], 1)``.

The two-source kernels are the instances the one-source dispatch picks for those shapes,
with the same accumulation order, so the outputs must be identical bit for bit; shapes
and dtypes without a two-source kernel fall back to the concatenation (also identical).
Gradients flow to both halves and to the weight exactly as through the concatenation.
"""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _fvdb():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb

    return fvdb


def _grid(fvdb, seed=3, n_points=300_000):
    from fvdb.utils.synthetic import sphere_points, voxelize

    return voxelize(sphere_points(seed=seed, n_points=n_points, device="cuda"), 1 / 200)


# (dtype, c_half, M, row order): the three two-source forms, then shapes that fall back
CASES = [
    (torch.bfloat16, 32, 32, False), (torch.float16, 32, 32, False),
    (torch.bfloat16, 64, 64, True), (torch.bfloat16, 64, 128, True), (torch.float16, 64, 64, True),
    (torch.bfloat16, 128, 128, True), (torch.bfloat16, 128, 64, False), (torch.float16, 128, 128, False),
    (torch.bfloat16, 64, 64, False), (torch.bfloat16, 48, 32, False), (torch.bfloat16, 128, 16, True),
    (torch.float32, 32, 32, False), (torch.float32, 64, 64, True), (torch.float64, 16, 8, False),
]


@pytest.mark.parametrize("dtype,c_half,M,ordered", CASES)
@pytest.mark.parametrize("epilogue", [False, True])
def test_two_source_contraction_bit_identical(dtype, c_half, M, ordered, epilogue):
    fvdb = _fvdb()
    grid = _grid(fvdb)
    N = grid.total_voxels
    pack, _ = grid.sparse_conv_kernel_map(3, 1)
    pack.build_implicit_gemm()
    m = pack._map_fwd()
    order = fvdb.ext.row_order(m, N, 8192) if ordered else None
    g = torch.Generator(device="cuda").manual_seed(c_half * 7 + M)
    x0 = torch.randn((N, c_half), generator=g, device="cuda").to(dtype)
    x1 = torch.randn((N, c_half), generator=g, device="cuda").to(dtype)
    a = (torch.randn((27, M, 2 * c_half), generator=g, device="cuda") / (27 * c_half) ** 0.5).to(dtype)
    bias = torch.randn((M,), generator=g, device="cuda").to(dtype) if epilogue else None
    print(f"[cat2] {dtype} {c_half}+{c_half}->{M} order={ordered} epilogue={epilogue} N={N}", flush=True)
    ref = fvdb.ext.implicit_gemm(torch.cat([x0, x1], 1), a, m, N, order, bias, epilogue)
    got = fvdb.ext.implicit_gemm_cat2(x0, x1, a, m, N, order, bias, epilogue)
    assert got.shape == ref.shape and torch.equal(got, ref)


@pytest.mark.parametrize("dtype", [torch.bfloat16, torch.float32])
def test_module_concat_equals_jcat_with_gradients(dtype):
    fvdb = _fvdb()
    grid = _grid(fvdb, seed=5, n_points=200_000)
    N = grid.total_voxels
    g = torch.Generator(device="cuda").manual_seed(11)
    a = torch.randn((N, 64), generator=g, device="cuda").to(dtype)
    b = torch.randn((N, 64), generator=g, device="cuda").to(dtype)
    torch.manual_seed(0)
    conv = fvdb.nn.SparseConv3d(128, 64, 3).cuda().to(dtype)
    R = torch.randn((N, 64), generator=g, device="cuda").to(dtype)

    def run(use_concat):
        xa, xb = a.clone().requires_grad_(True), b.clone().requires_grad_(True)
        conv.zero_grad()
        va = fvdb.nn.VDBTensor(grid, grid.jagged_like(xa))
        vb = fvdb.nn.VDBTensor(grid, grid.jagged_like(xb))
        if use_concat:
            y = conv(va, fuse_relu=True, concat=vb)
        else:
            y = conv(fvdb.jcat([va, vb], dim=1), fuse_relu=True)
        (y.data.jdata.float() * R.float()).sum().backward()
        return y.data.jdata.detach(), xa.grad, xb.grad, conv.weight.grad.clone(), conv.bias.grad.clone()

    ref, got = run(False), run(True)
    for r, o in zip(ref, got):
        assert torch.equal(r, o)
