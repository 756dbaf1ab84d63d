"""GPU parity of pooling, nearest upsampling and injection against the CPU oracle.

Window maps, pooled maxima, upsampled and injected values are bit-exact; average pooling
is bit-exact against the oracle's identical arithmetic (fp32 accumulation in tap order,
then T(sum) * T(scale)); gradients are checked by their definitions and by fp64 gradcheck.
"""
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import oracle

pytestmark = pytest.mark.gpu


def _fvdb():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb

    return fvdb


def grids(fvdb, rng, sizes, hi):
    raw = [torch.from_numpy(rng.integers(-hi // 4, hi, size=(n, 3)).astype(np.int32)).cuda() for n in sizes]
    g = fvdb.GridBatch.from_ijk(fvdb.JaggedTensor(raw))
    offs = g.joffsets.cpu().tolist()
    per = [g.ijk.jdata[offs[b]:offs[b + 1]].cpu().numpy() for b in range(len(sizes))]
    return g, per


CASES = [((2, 2, 2), (2, 2, 2)), ((3, 3, 3), (2, 2, 2)), ((2, 2, 2), (1, 1, 1)), ((2, 1, 3), (1, 2, 2))]


@pytest.mark.parametrize("f,s", CASES)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32, torch.float16, torch.bfloat16])
@pytest.mark.parametrize("C", [5, 16])
def test_pool_vs_oracle(f, s, dtype, C):
    """C = 5: the per-element kernels; C = 16: rows of whole 16-byte segments in every dtype
    (the vector kernel)."""
    fvdb = _fvdb()
    rng = np.random.default_rng(sum(f) + 10 * sum(s) + C)
    print(f"[case] pool f={f} s={s} {dtype} C={C}", flush=True)
    fine, fine_np = grids(fvdb, rng, [700, 300], 40)
    x = torch.from_numpy(rng.standard_normal((fine.total_voxels, C))).to("cuda", dtype)
    out_max, coarse = fine.max_pool(f, fine.jagged_like(x), stride=s)
    out_avg, coarse2 = fine.avg_pool(f, fine.jagged_like(x), stride=s, coarse_grid=coarse)
    offs = coarse.joffsets.cpu().tolist()
    coarse_np = [coarse.ijk.jdata[offs[b]:offs[b + 1]].cpu().numpy() for b in range(2)]
    for b in range(2):
        assert np.array_equal(coarse_np[b], oracle.sort_unique(np.floor_divide(fine_np[b], np.asarray(s)).astype(np.int32)))
    wmap = oracle.window_map(coarse_np, fine_np, f, s)
    from fvdb._pooling import window_map
    gmap = window_map(coarse, fine, f, s)[: coarse.total_voxels].cpu().numpy()
    assert np.array_equal(gmap, wmap)
    ref_max, ref_arg = oracle.max_pool(x.cpu(), wmap)
    assert torch.equal(out_max.jdata.cpu(), ref_max)
    ref_avg = oracle.avg_pool(x.cpu(), wmap, 1.0 / (f[0] * f[1] * f[2]))
    assert torch.equal(out_avg.jdata.cpu(), ref_avg)
    # backward by definition
    xg = x.clone().requires_grad_(True)
    gy = torch.from_numpy(rng.standard_normal((coarse.total_voxels, C))).to("cuda", dtype)
    fine.max_pool(f, fine.jagged_like(xg), stride=s, coarse_grid=coarse)[0].jdata.backward(gy)
    want = torch.zeros((fine.total_voxels, C), dtype=torch.float64)
    ok = ref_arg >= 0
    want.view(-1).index_add_(0, (ref_arg * C + torch.arange(C))[ok], gy.cpu().double()[ok])
    # windows overlap when a stride is below the pool factor: several contributions then
    # meet in one voxel and are summed by atomics in an unspecified order
    overlap = any(a < b for a, b in zip(s, f))
    tol = 0.0 if not overlap else {torch.float64: 1e-12, torch.float32: 1e-6}.get(dtype, 1e-2)
    assert torch.allclose(xg.grad.cpu().double(), want.to(dtype).double(), rtol=tol, atol=tol)
    xg2 = x.clone().requires_grad_(True)
    fine.avg_pool(f, fine.jagged_like(xg2), stride=s, coarse_grid=coarse)[0].jdata.backward(gy)
    gs = (gy.cpu() * torch.tensor(1.0 / (f[0] * f[1] * f[2]), dtype=dtype)).double()
    want2 = torch.zeros((fine.total_voxels, C), dtype=torch.float64)
    rows, taps = np.nonzero(wmap >= 0)
    want2.index_add_(0, torch.from_numpy(wmap[rows, taps].astype(np.int64)), gs[torch.from_numpy(rows)])
    assert torch.allclose(xg2.grad.cpu().double(), want2.to(dtype).double(), rtol=tol, atol=tol)


def test_pool_dense_box_matches_torch():
    fvdb = _fvdb()
    g = fvdb.GridBatch.from_dense(1, (8, 6, 10), device="cuda")
    ijk = g.ijk.jdata.long()
    dense = torch.randn(3, 8, 6, 10, dtype=torch.float64, device="cuda")
    x = dense[:, ijk[:, 0], ijk[:, 1], ijk[:, 2]].t().contiguous()
    for f, s in (((2, 2, 2), (2, 2, 2)), ((2, 3, 2), (2, 3, 2))):
        mx, coarse = g.max_pool(f, g.jagged_like(x), stride=s)
        av, _ = g.avg_pool(f, g.jagged_like(x), stride=s, coarse_grid=coarse)
        cidx = coarse.ijk.jdata.long()
        ref_m = F.max_pool3d(dense[None], f, s)[0][:, cidx[:, 0], cidx[:, 1], cidx[:, 2]].t()
        ref_a = F.avg_pool3d(dense[None], f, s)[0][:, cidx[:, 0], cidx[:, 1], cidx[:, 2]].t()
        assert torch.equal(mx.jdata, ref_m)
        assert torch.allclose(av.jdata, ref_a, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("f", [(2, 2, 2), (3, 1, 2)])
@pytest.mark.parametrize("dtype,C", [(torch.float64, 4), (torch.float32, 3), (torch.float16, 3),
                                     (torch.bfloat16, 8)])
def test_subdivide_and_inject_vs_oracle(f, dtype, C):
    """Rows copied in 16-byte units (f64 x 4, bf16 x 8), 4-byte units (f32 x 3) and bytes
    (f16 x 3)."""
    fvdb = _fvdb()
    rng = np.random.default_rng(3)
    print(f"[case] subdivide/inject f={f} {dtype} C={C}", flush=True)
    coarse, coarse_np = grids(fvdb, rng, [200, 90], 16)
    vals = torch.from_numpy(rng.standard_normal((coarse.total_voxels, C))).to("cuda", dtype)
    up, fine = coarse.subdivide(f, coarse.jagged_like(vals))
    offs = fine.joffsets.cpu().tolist()
    fine_np = [fine.ijk.jdata[offs[b]:offs[b + 1]].cpu().numpy() for b in range(2)]
    pidx = torch.from_numpy(oracle.parent_index(fine_np, coarse_np, f))
    assert bool((pidx >= 0).all())  # every child has its parent
    assert torch.equal(up.jdata.cpu(), vals.cpu()[pidx])
    # a foreign fine grid: children of missing parents read zero
    other, other_np = grids(fvdb, rng, [500, 200], 40)
    up2, _ = coarse.subdivide(f, coarse.jagged_like(vals), fine_grid=other)
    pidx2 = torch.from_numpy(oracle.parent_index(other_np, coarse_np, f))
    want = torch.zeros((other.total_voxels, C), dtype=dtype)
    ok = pidx2 >= 0
    want[ok] = vals.cpu()[pidx2[ok]]
    assert torch.equal(up2.jdata.cpu(), want)
    # inject coarse values onto the foreign grid
    inj = other.inject_from(coarse, coarse.jagged_like(vals), default_value=-7.0)
    iidx = torch.from_numpy(oracle.inject_index(other_np, coarse_np))
    want = torch.full((other.total_voxels, C), -7.0, dtype=dtype)
    ok = iidx >= 0
    want[ok] = vals.cpu()[iidx[ok]]
    assert torch.equal(inj.jdata.cpu(), want)


def test_gradcheck_fp64():
    fvdb = _fvdb()
    rng = np.random.default_rng(4)
    fine, _ = grids(fvdb, rng, [120], 10)
    coarse = fine.coarsened_grid(2)
    x = torch.from_numpy(rng.standard_normal((fine.total_voxels, 2))).cuda().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: fine.max_pool(2, fine.jagged_like(t), coarse_grid=coarse)[0].jdata, (x,))
    assert torch.autograd.gradcheck(lambda t: fine.avg_pool(2, fine.jagged_like(t), coarse_grid=coarse)[0].jdata, (x,))
    assert torch.autograd.gradcheck(lambda t: fine.avg_pool(3, fine.jagged_like(t), stride=2)[0].jdata, (x,))
    c = torch.from_numpy(rng.standard_normal((coarse.total_voxels, 2))).cuda().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: coarse.subdivide(2, coarse.jagged_like(t), fine_grid=fine)[0].jdata, (c,))
    d = torch.from_numpy(rng.standard_normal((coarse.total_voxels, 2))).cuda().requires_grad_(True)
    other, _ = grids(fvdb, rng, [100], 8)
    e = torch.from_numpy(rng.standard_normal((other.total_voxels, 2))).cuda().requires_grad_(True)
    assert torch.autograd.gradcheck(
        lambda s_, t_: coarse.inject_from(other, other.jagged_like(s_), dst=coarse.jagged_like(t_)).jdata, (e, d))


def test_modules_chain():
    fvdb = _fvdb()
    torch.manual_seed(0)
    g = fvdb.GridBatch.from_dense(2, (8, 8, 8), device="cuda")
    x = fvdb.nn.VDBTensor(g, g.jagged_like(torch.randn(g.total_voxels, 4, device="cuda")))
    down = fvdb.nn.MaxPool(2)(x)
    assert down.grid.total_voxels == 2 * 64
    avg = fvdb.nn.AvgPool(2)(x, ref_coarse_data=down)
    assert avg.grid.is_same(down.grid)
    up = fvdb.nn.UpsamplingNearest(2)(down, ref_fine_data=x)
    assert up.grid.is_same(g) and torch.equal(fvdb.nn.MaxPool(2)(up, down).data.jdata, down.data.jdata)
    filled = fvdb.nn.FillFromGrid(default_value=1.5)(down, x)
    assert filled.grid.is_same(g)
    # an empty window maps to 0 in the module (the reference's -inf -> 0)
    sparse = fvdb.GridBatch.from_ijk(torch.tensor([[0, 0, 0], [5, 5, 5]], dtype=torch.int32, device="cuda"))
    ref = fvdb.GridBatch.from_ijk(torch.tensor([[0, 0, 0], [1, 1, 1], [2, 2, 2]], dtype=torch.int32, device="cuda"))
    v = fvdb.nn.VDBTensor(sparse, sparse.jagged_like(torch.ones(2, 1, device="cuda")))
    out = fvdb.nn.MaxPool(2)(v, ref)
    assert out.data.jdata.reshape(-1).tolist() == [1.0, 0.0, 1.0]


@pytest.mark.parametrize("factor", [(2, 2, 2), (1, 2, 4), (4, 1, 2), (8, 8, 8), (1, 1, 1), (32, 16, 2)])
def test_coarsen_from_leaves_bit_exact(factor):
    """Coarsening by power-of-two factors is built from the fine grid's leaves (one
    de-duplication entry per leaf, coarse masks ORed from the fine masks).  Every output
    of the index (voxels in tree order, grid offsets, leaf keys, masks, word bases) equals
    the build from the coarsened voxels bit for bit, and the voxels equal the oracle's
    floor(ijk / factor) set in tree order, on grids with negative coordinates, several
    grids, and one of 700k voxels on sphere shells."""
    fvdb = _fvdb()
    from fvdb._lib import ext
    from fvdb.utils.synthetic import sphere_points, voxelize

    rng = np.random.default_rng(sum(factor))
    small, small_np = grids(fvdb, rng, [900, 1, 400], 60)
    big = voxelize(sphere_points(seed=9, n_points=700_000, device="cuda"), 1 / 300)
    for g, per in ((small, small_np), (big, None)):
        print(f"[case] coarsen factor={factor} voxels={g.total_voxels}", flush=True)
        fast = g.coarsened_grid(factor)
        cijk, cj = ext.conv_grid_candidates(g._impl.ijk, g._impl.jidx, list(factor), list(factor))
        ref = g._derived(cijk, cj)
        a, b = fast._impl, ref._impl
        for name in ("ijk", "jidx", "joffsets", "keys", "masks", "wbase"):
            assert torch.equal(getattr(a, name), getattr(b, name)), name
        if per is not None:
            offs = fast.joffsets.cpu().tolist()
            for i, p in enumerate(per):
                want = oracle.sort_unique(np.floor_divide(p, np.asarray(factor, dtype=np.int32)))
                assert np.array_equal(fast.ijk.jdata[offs[i]:offs[i + 1]].cpu().numpy(), want)
    # conv target grid with stride == kernel size takes the same path
    t = small.conv_grid(2, 2)
    want = small.coarsened_grid(2)
    assert torch.equal(t.ijk.jdata, want.ijk.jdata) and torch.equal(t._impl.masks, want._impl.masks)


def test_ijk_to_inv_index_and_inject_from_ijk():
    """ijk_to_inv_index: per voxel, the position of the query naming it (per list or
    cumulative), -1 where none does, the largest position among duplicate queries;
    inject_from_ijk: rows of src land on the voxels at src_ijk, the rest keep dst or the
    default.  Checked against dictionaries built on the host."""
    fvdb = _fvdb()
    rng = np.random.default_rng(12)
    g, per = grids(fvdb, rng, [500, 300], 24)
    offs = g.joffsets.cpu().tolist()
    q_np = []
    for b, p in enumerate(per):
        pick = p[rng.integers(0, len(p), size=200)]                      # hits, with duplicates
        miss = rng.integers(100, 120, size=(40, 3)).astype(np.int32)     # outside the grid
        q = np.concatenate([pick, miss])[rng.permutation(240)]
        q_np.append(q)
    q = fvdb.JaggedTensor([torch.from_numpy(x).cuda() for x in q_np])
    for cumulative in (False, True):
        got = g.ijk_to_inv_index(q, cumulative=cumulative).jdata.cpu().numpy()
        want = np.full(g.total_voxels, -1, dtype=np.int64)
        base = 0
        for b, (p, qq) in enumerate(zip(per, q_np)):
            where = {tuple(c): offs[b] + i for i, c in enumerate(p.tolist())}
            for e, c in enumerate(qq.tolist()):
                v = where.get(tuple(c))
                if v is not None:
                    want[v] = max(want[v], e + (base if cumulative else 0))
            base += len(qq)
        assert np.array_equal(got, want), cumulative
    # unique coordinates per list so that the expected result is unambiguous
    uq = [np.unique(x, axis=0) for x in q_np]
    sq = fvdb.JaggedTensor([torch.from_numpy(x).cuda() for x in uq])
    sv = fvdb.JaggedTensor([torch.from_numpy(rng.standard_normal((len(x), 2))).cuda() for x in uq])
    out = g.inject_from_ijk(sq, sv, default_value=-5.0).jdata.cpu().numpy()
    want = np.full((g.total_voxels, 2), -5.0)
    vals = [t.cpu().numpy() for t in [sv.jdata[sv.joffsets[i]:sv.joffsets[i + 1]] for i in range(len(uq))]]
    for b, (p, x, v) in enumerate(zip(per, uq, vals)):
        where = {tuple(c): offs[b] + i for i, c in enumerate(p.tolist())}
        for c, row in zip(x.tolist(), v):
            if tuple(c) in where:
                want[where[tuple(c)]] = row
    assert np.array_equal(out, want)
    dst = g.jagged_like(torch.zeros(g.total_voxels, 2, dtype=torch.float64, device="cuda"))
    res = g.inject_from_ijk(sq, sv, dst=dst)
    assert res is dst and np.array_equal(dst.jdata.cpu().numpy(), np.where(want == -5.0, 0.0, want))


def test_grid_pickle_round_trip():
    """pickle.dumps / pickle.loads of a GridBatch gives the same grids: voxel coordinates in the
    same linear order, the same grid offsets, voxel sizes and origins, and the same answers
    from ijk_to_index; a single grid and a batch holding an empty grid both round-trip."""
    import pickle
    fvdb = _fvdb()
    rng = np.random.default_rng(31)
    g, _ = grids(fvdb, rng, [400, 250, 90], 40)
    g = fvdb.GridBatch.from_ijk(g.ijk, voxel_sizes=[[0.5, 0.25, 1.0], [1.0, 1.0, 1.0], [2.0, 2.0, 0.125]],
                                origins=[[0.0, 1.0, -2.0], [0.5, 0.5, 0.5], [3.0, 0.0, 0.0]])
    empty = torch.empty((0, 3), dtype=torch.int32, device="cuda")
    single = fvdb.GridBatch.from_ijk(g.ijk.jdata[:150].clone())
    with_empty = fvdb.GridBatch.from_ijk(fvdb.JaggedTensor([g.ijk.jdata[:60].clone(), empty]))
    for src in (g, single, with_empty):
        back = pickle.loads(pickle.dumps(src))
        assert back.grid_count == src.grid_count and back.device == src.device
        assert torch.equal(back.ijk.jdata, src.ijk.jdata)
        assert torch.equal(back.joffsets, src.joffsets)
        assert torch.equal(back.voxel_sizes, src.voxel_sizes) and torch.equal(back.origins, src.origins)
        q = fvdb.JaggedTensor([src.ijk.jdata.flip(0)] * src.grid_count)
        assert torch.equal(back.ijk_to_index(q).jdata, src.ijk_to_index(q).jdata)
