"""Pins of the oracle's pooling, upsampling and injection (CPU only).

Every check compares the oracle with an independent closed form: torch's dense
``max_pool3d`` / ``avg_pool3d`` / nearest ``interpolate`` on a dense tensor that holds
-inf (max) or 0 (average, upsampling) at missing voxels, or a set intersection.
"""
import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import oracle


def sparse_set(rng, n, lo, hi):
    return oracle.sort_unique(rng.integers(lo, hi, size=(n, 3)).astype(np.int32))


def dense_of(ijk, vals, shape, fill):
    d = torch.full((vals.shape[1],) + tuple(shape), fill, dtype=vals.dtype)
    t = torch.from_numpy(ijk.astype(np.int64))
    d[:, t[:, 0], t[:, 1], t[:, 2]] = vals.t()
    return d


def coarse_of(ijk, s):
    return oracle.sort_unique(np.floor_divide(ijk, np.asarray(s)).astype(np.int32))


CASES = [((2, 2, 2), (2, 2, 2)), ((3, 3, 3), (2, 2, 2)), ((2, 2, 2), (1, 1, 1)), ((2, 1, 3), (1, 2, 2)),
         ((1, 1, 1), (1, 1, 1)), ((4, 2, 2), (4, 2, 1))]


@pytest.mark.parametrize("f,s", CASES)
@pytest.mark.parametrize("dense", [True, False])
def test_max_avg_pool_vs_dense_torch(f, s, dense):
    rng = np.random.default_rng(sum(f) * 7 + sum(s) + dense)
    if dense:
        g = np.stack(np.meshgrid(*[np.arange(n) for n in (7, 6, 9)], indexing="ij"), -1).reshape(-1, 3)
        fine = oracle.sort_unique(g.astype(np.int32))
    else:
        fine = sparse_set(rng, 300, 0, 12)
    coarse = coarse_of(fine, s)
    x = torch.from_numpy(rng.standard_normal((len(fine), 3)))
    wmap = oracle.window_map([coarse], [fine], f, s)
    ext = [max(int(coarse[:, a].max()) * s[a] + f[a], int(fine[:, a].max()) + 1) for a in range(3)]
    idx = torch.from_numpy(coarse.astype(np.int64))
    # max: -inf where no voxel
    dm = F.max_pool3d(dense_of(fine, x, ext, -math.inf)[None], f, s)[0]
    ref_max = dm[:, idx[:, 0], idx[:, 1], idx[:, 2]].t()
    out, arg = oracle.max_pool(x, wmap)
    assert torch.equal(out, ref_max)
    # average: zeros where no voxel, divisor = window volume
    da = F.avg_pool3d(dense_of(fine, x, ext, 0.0)[None], f, s)[0]
    ref_avg = da[:, idx[:, 0], idx[:, 1], idx[:, 2]].t()
    got = oracle.avg_pool(x, wmap, 1.0 / (f[0] * f[1] * f[2]))
    assert torch.allclose(got, ref_avg, rtol=1e-12, atol=1e-12)
    # argmax points at a voxel holding the maximum, and is -1 exactly when the window is empty
    empty = (wmap < 0).all(1)
    assert bool(((arg < 0).all(1) == torch.from_numpy(empty)).all())
    ok = arg >= 0
    assert torch.equal(x[arg.clamp(min=0), torch.arange(3)][ok], out[ok])


def test_max_pool_ties_take_the_first_tap():
    fine = oracle.sort_unique(np.array([[0, 0, 0], [0, 0, 1], [1, 0, 0], [1, 1, 1]], np.int32))
    coarse = coarse_of(fine, (2, 2, 2))
    wmap = oracle.window_map([coarse], [fine], 2, 2)
    x = torch.tensor([[1.0], [5.0], [5.0], [2.0]], dtype=torch.float64)
    out, arg = oracle.max_pool(x, wmap)
    # taps run x outermost: (0,0,1) -> tap 1 precedes (1,0,0) -> tap 4
    idx = {tuple(c): i for i, c in enumerate(fine.tolist())}
    assert out.item() == 5.0 and arg.item() == idx[(0, 0, 1)]
    assert wmap[0].tolist().index(idx[(0, 0, 1)]) == 1 and wmap[0].tolist().index(idx[(1, 0, 0)]) == 4


@pytest.mark.parametrize("f", [(2, 2, 2), (3, 1, 2)])
def test_upsample_parent_vs_nearest_interpolate(f):
    rng = np.random.default_rng(5)
    coarse = sparse_set(rng, 150, 0, 8)
    fine = sparse_set(rng, 600, 0, 8 * max(f))
    pidx = oracle.parent_index([fine], [coarse], f)
    vals = torch.from_numpy(rng.standard_normal((len(coarse), 2)))
    ext = [int(fine[:, a].max()) // f[a] + 1 for a in range(3)]
    dc = dense_of(coarse, vals, ext, 0.0)
    up = F.interpolate(dc[None], scale_factor=f, mode="nearest")[0]
    t = torch.from_numpy(fine.astype(np.int64))
    ref = up[:, t[:, 0], t[:, 1], t[:, 2]].t()
    got = torch.zeros((len(fine), 2), dtype=torch.float64)
    ok = torch.from_numpy(pidx >= 0)
    got[ok] = vals[torch.from_numpy(pidx)[ok]]
    assert torch.equal(got, ref)


def test_upsample_negative_coordinates_floor():
    coarse = oracle.sort_unique(np.array([[-1, -1, -1], [0, 0, 0]], np.int32))
    fine = oracle.sort_unique(np.array([[-1, -2, -1], [-3, 0, 0], [1, 1, 1], [-2, -2, -2]], np.int32))
    pidx = oracle.parent_index([fine], [coarse], 2)
    cidx = {tuple(c): i for i, c in enumerate(coarse.tolist())}
    want = [cidx.get(tuple(int(math.floor(v / 2)) for v in c), -1) for c in fine.tolist()]
    assert pidx.tolist() == want


def test_inject_index_is_set_intersection():
    rng = np.random.default_rng(9)
    a = [sparse_set(rng, 200, 0, 10), sparse_set(rng, 50, 0, 6)]
    b = [sparse_set(rng, 300, 0, 10), sparse_set(rng, 80, 0, 6)]
    idx = oracle.inject_index(b, a)
    base_a = [0, len(a[0])]
    k = 0
    for g in range(2):
        sa = {tuple(c): base_a[g] + i for i, c in enumerate(a[g].tolist())}
        for c in b[g].tolist():
            assert idx[k] == sa.get(tuple(c), -1)
            k += 1
