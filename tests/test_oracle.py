"""Pins the CPU oracle before anything is compared against it.

Each check is chosen so that a plausible mistake in the oracle (a swapped axis, a
wrong kernel start, a transposed weight, a dropped offset) fails at least one of
them: brute force with Python sets/dicts on tiny grids, closed forms on dense boxes,
dense ``torch.nn.functional.conv3d`` / ``conv_transpose3d`` in float64 (the
reference's own strategy, ``tests/unit/test_conv.py:361-675``), and the committed
golden fixtures.
"""
import itertools
import math
import os

import numpy as np
import pytest
import torch

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def dense_box(dims, lo=(0, 0, 0)):
    g = np.stack(np.meshgrid(*[np.arange(l, l + d) for l, d in zip(lo, dims)], indexing="ij"), -1)
    return g.reshape(-1, 3).astype(np.int32)


def random_voxels(rng, n, extent, lo=0):
    return rng.integers(lo, lo + extent, size=(n, 3)).astype(np.int32)


# --------------------------------------------------------------------------------------
# voxel order
# --------------------------------------------------------------------------------------


def test_order_single_leaf_closed_form():
    g = oracle.sort_unique(dense_box((8, 8, 8)))
    idx = g[:, 0] * 64 + g[:, 1] * 8 + g[:, 2]
    assert np.array_equal(idx, np.arange(512))


def test_order_leaves_then_bits():
    g = oracle.sort_unique(dense_box((16, 16, 16)))
    leaf = (g[:, 0] >> 3) * 4 + (g[:, 1] >> 3) * 2 + (g[:, 2] >> 3)
    bit = (g[:, 0] & 7) * 64 + (g[:, 1] & 7) * 8 + (g[:, 2] & 7)
    key = leaf * 512 + bit
    assert np.array_equal(key, np.arange(4096))


def test_order_lower_and_upper_nodes():
    # a coordinate in a later lower node (x >= 128) sorts after every coordinate of
    # the first lower node even when its leaf index inside the node is smaller
    assert oracle.cmp_order([128, 0, 0], [127, 127, 127]) > 0
    # lower-node child (0,1,0) sorts after (0,0,15): the y digit decides before z
    assert oracle.cmp_order([0, 8, 0], [7, 0, 127]) > 0
    assert oracle.cmp_order([8, 0, 0], [0, 127, 127]) > 0
    assert oracle.cmp_order([4096, 0, 0], [4095, 4095, 4095]) > 0
    # negative coordinates: unsigned root key -> after every non-negative tile
    assert oracle.cmp_order([-1, 0, 0], [4095, 0, 0]) > 0


def test_order_is_a_bijection_and_dedups():
    rng = np.random.default_rng(0)
    a = random_voxels(rng, 3000, 40, lo=-20)
    g = oracle.sort_unique(a)
    assert len({tuple(x) for x in a.tolist()}) == g.shape[0]
    for i in range(g.shape[0] - 1):
        assert oracle.cmp_order(g[i], g[i + 1]) < 0


# --------------------------------------------------------------------------------------
# target grid
# --------------------------------------------------------------------------------------


def brute_conv_grid(src, ks, st):
    ks = (ks,) * 3 if isinstance(ks, int) else tuple(ks)
    st = (st,) * 3 if isinstance(st, int) else tuple(st)
    out = set()
    if st == (1, 1, 1) or st == ks:
        for x, y, z in src.tolist():
            out.add((x // st[0], y // st[1], z // st[2]))
    else:
        rng = []
        for d in range(3):
            rng.append(range(0, ks[d]) if ks[d] % 2 == 0 else range(-(ks[d] - 1) // 2, (ks[d] - 1) // 2 + 1))
        for x, y, z in src.tolist():
            for di in rng[0]:
                for dj in rng[1]:
                    for dk in rng[2]:
                        X, Y, Z = x + dk, y + dj, z + di
                        # C remainder: sign of the dividend
                        if math.fmod(X, st[2]) != 0 or math.fmod(Y, st[1]) != 0 or math.fmod(Z, st[0]) != 0:
                            continue
                        out.add((int(X / st[2]), int(Y / st[1]), int(Z / st[0])))
    return out


@pytest.mark.parametrize("ks,st", [(3, 1), (2, 2), (3, 2), (5, 2), (4, 2), ((1, 3, 2), (1, 2, 1)), ((5, 3, 1), (1, 3, 2)), (3, 3)])
def test_conv_grid_bruteforce(ks, st):
    rng = np.random.default_rng(1)
    src = oracle.sort_unique(random_voxels(rng, 300, 12, lo=-6))
    got = oracle.conv_grid(src, ks, st)
    assert {tuple(x) for x in got.tolist()} == brute_conv_grid(src, ks, st)
    assert got.shape[0] == len({tuple(x) for x in got.tolist()})


# --------------------------------------------------------------------------------------
# kernel map
# --------------------------------------------------------------------------------------


def brute_kmap(src, tgt, ks, st):
    ks = (ks,) * 3 if isinstance(ks, int) else tuple(ks)
    st = (st,) * 3 if isinstance(st, int) else tuple(st)
    lut = {tuple(c): i for i, c in enumerate(src.tolist())}
    start = [math.floor(-k / 2 + 1) for k in ks]
    K = ks[0] * ks[1] * ks[2]
    out = np.full((tgt.shape[0], K), -1, np.int32)
    for t, (x, y, z) in enumerate(tgt.tolist()):
        k = 0
        # outer loop over kernel dim 2 which offsets x
        for a in range(ks[2]):
            for b in range(ks[1]):
                for c in range(ks[0]):
                    q = (x * st[2] + start[2] + a, y * st[1] + start[1] + b, z * st[0] + start[0] + c)
                    out[t, k] = lut.get(q, -1)
                    k += 1
    return out


@pytest.mark.parametrize("ks,st", [(3, 1), (2, 2), (3, 2), (4, 1), (5, 2), ((1, 3, 2), (1, 2, 1)), ((4, 1, 4), (3, 3, 3))])
def test_kmap_bruteforce(ks, st):
    rng = np.random.default_rng(2)
    src = oracle.sort_unique(random_voxels(rng, 400, 14, lo=-7))
    tgt = src if st == 1 else oracle.conv_grid(src, ks, st)
    got = oracle.kernel_map([src], [tgt], ks, st)
    assert np.array_equal(got, brute_kmap(src, tgt, ks, st))


def test_kmap_batched_offsets():
    rng = np.random.default_rng(3)
    grids = [oracle.sort_unique(random_voxels(rng, n, 10)) for n in (50, 200, 5)]
    km = oracle.kernel_map(grids, grids, 3, 1)
    base = 0
    row = 0
    for g in grids:
        part = km[row:row + g.shape[0]]
        ref = brute_kmap(g, g, 3, 1)
        assert np.array_equal(np.where(ref >= 0, ref + base, -1), part)
        base += g.shape[0]
        row += g.shape[0]


def test_kmap_dense_closed_form_and_centre_identity():
    g = oracle.sort_unique(dense_box((9, 10, 11)))
    km = oracle.kernel_map([g], [g], 3, 1)
    assert np.array_equal(km[:, 13], np.arange(g.shape[0]))
    pos = {tuple(c): i for i, c in enumerate(g.tolist())}
    for t in [0, 17, g.shape[0] - 1, 500]:
        x, y, z = g[t]
        k = 0
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                for c in (-1, 0, 1):
                    q = (x + a, y + b, z + c)
                    inside = 0 <= q[0] < 9 and 0 <= q[1] < 10 and 0 <= q[2] < 11
                    assert km[t, k] == (pos[q] if inside else -1)
                    k += 1


def test_neighbor_map_invariants():
    rng = np.random.default_rng(4)
    g = oracle.sort_unique(random_voxels(rng, 500, 12))
    km = oracle.kernel_map([g], [g], 3, 1)
    nb, sizes = oracle.neighbor_map(km)
    assert sizes.sum() == (km >= 0).sum() == nb.shape[0]
    off = 0
    for k in range(27):
        seg = nb[off:off + sizes[k]]
        assert np.all(np.diff(seg[:, 1]) > 0)
        assert np.array_equal(km[seg[:, 1], k], seg[:, 0])
        off += sizes[k]
    t = oracle.transpose_map(km, g.shape[0])
    # transpose is an involution on valid entries
    assert np.array_equal(oracle.transpose_map(t, g.shape[0]), km)


# --------------------------------------------------------------------------------------
# gather-scatter against dense conv3d in float64
# --------------------------------------------------------------------------------------


def to_dense_czyx(ijk, feat, dims):
    C = feat.shape[1]
    d = torch.zeros((C, dims[2], dims[1], dims[0]), dtype=feat.dtype)
    ii = torch.from_numpy(ijk.astype(np.int64))
    d[:, ii[:, 2], ii[:, 1], ii[:, 0]] = feat.t()
    return d


@pytest.mark.parametrize("ks,st", [(3, 1), (2, 1), (4, 1), (5, 1), (3, 2), (2, 2), ((1, 3, 2), 1), ((5, 3, 1), 1)])
def test_gs_forward_backward_vs_dense_conv3d(ks, st):
    torch.manual_seed(0)
    dims = (10, 9, 8)
    g = oracle.sort_unique(dense_box(dims))
    ksv = (ks,) * 3 if isinstance(ks, int) else ks
    tgt = g if st == 1 else oracle.conv_grid(g, ks, st)
    km = oracle.kernel_map([g], [tgt], ks, st)
    x = torch.randn(g.shape[0], 3, dtype=torch.float64, requires_grad=True)
    w = torch.randn(5, 3, *ksv, dtype=torch.float64, requires_grad=True)
    out = oracle.sparse_conv(x, w, km, tgt.shape[0], st)
    dense = to_dense_czyx(g, x, dims)
    pad = [(k - 1) // 2 for k in ksv]
    ref = torch.nn.functional.conv3d(dense[None], w, padding=pad, stride=st)[0]  # [O, Z, Y, X]
    ti = torch.from_numpy(tgt.astype(np.int64))
    keep = (ti[:, 0] < ref.shape[3]) & (ti[:, 1] < ref.shape[2]) & (ti[:, 2] < ref.shape[1])
    ref_rows = ref[:, ti[keep, 2], ti[keep, 1], ti[keep, 0]].t()
    assert torch.allclose(out[keep], ref_rows.detach(), atol=1e-12, rtol=1e-12)
    # backward: loss = sum(out * R) over the rows that the dense conv also has
    R = torch.randn_like(out)
    R[~keep] = 0
    (ref_rows * R[keep]).sum().backward()
    gi, gw = oracle.sparse_conv_backward(x, R, w, km, ks)
    assert torch.allclose(gi, x.grad, atol=1e-12, rtol=1e-12)
    assert torch.allclose(gw, w.grad, atol=1e-11, rtol=1e-11)


@pytest.mark.parametrize("ks", [2, 3, 4])
def test_gs_transposed_vs_conv_transpose3d(ks):
    torch.manual_seed(1)
    dims = (8, 8, 8)
    g = oracle.sort_unique(dense_box(dims))
    km = oracle.kernel_map([g], [g], ks, 1)
    x = torch.randn(g.shape[0], 4, dtype=torch.float64)
    w = torch.randn(4, 6, ks, ks, ks, dtype=torch.float64)  # [I, O, k0, k1, k2]
    out = oracle.sparse_conv(x, w, km, g.shape[0], 1, transposed=True)
    dense = to_dense_czyx(g, x, dims)
    # transposed conv = adjoint of the forward conv: padding mirrors the forward's
    pad = (ks - 1) // 2
    ref = torch.nn.functional.conv_transpose3d(dense[None], w, padding=pad)[0]
    ref = ref[:, :8, :8, :8]
    start = math.floor(-ks / 2 + 1)
    # output voxel s receives from targets t with s = t + start + a; conv_transpose3d with
    # padding p places in[t] * W[a] at t + a - p: equal when start == -p, which holds
    assert start == -pad
    ii = torch.from_numpy(g.astype(np.int64))
    ref_rows = ref[:, ii[:, 2], ii[:, 1], ii[:, 0]].t()
    assert torch.allclose(out, ref_rows, atol=1e-12, rtol=1e-12)


def test_gs_bruteforce_tiny_with_batch():
    rng = np.random.default_rng(5)
    torch.manual_seed(2)
    grids = [oracle.sort_unique(random_voxels(rng, n, 6)) for n in (30, 60)]
    km = oracle.kernel_map(grids, grids, 3, 1)
    N = sum(g.shape[0] for g in grids)
    x = torch.randn(N, 2, dtype=torch.float64)
    w = torch.randn(3, 2, 3, 3, 3, dtype=torch.float64)
    out = oracle.sparse_conv(x, w, km, N, 1)
    ref = torch.zeros(N, 3, dtype=torch.float64)
    wk = w.permute(4, 3, 2, 1, 0).reshape(27, 2, 3)
    for t in range(N):
        for k in range(27):
            s = km[t, k]
            if s >= 0:
                ref[t] += x[s] @ wk[k]
    assert torch.allclose(out, ref, atol=1e-12, rtol=1e-12)


# --------------------------------------------------------------------------------------
# golden fixtures (written by tests/golden/make_golden.py from this oracle and
# committed; they pin the oracle against silent change)
# --------------------------------------------------------------------------------------


def _golden_files():
    if not os.path.isdir(GOLDEN):
        return []
    return sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))


@pytest.mark.parametrize("fname", _golden_files())
def test_golden_fixture(fname):
    z = np.load(os.path.join(GOLDEN, fname))
    ks = tuple(int(v) for v in z["kernel_size"])
    st = tuple(int(v) for v in z["stride"])
    offs = z["src_joffsets"]
    srcs = [oracle.sort_unique(z["src_ijk_raw"][offs[b]:offs[b + 1]]) for b in range(len(offs) - 1)]
    assert np.array_equal(np.concatenate(srcs), z["src_ijk"])
    toffs = z["tgt_joffsets"]
    tgts = [z["tgt_ijk"][toffs[b]:toffs[b + 1]] for b in range(len(toffs) - 1)]
    if st != (1, 1, 1):
        for s, t in zip(srcs, tgts):
            assert np.array_equal(oracle.conv_grid(s, ks, st), t)
    km = oracle.kernel_map(srcs, tgts, ks, st)
    assert np.array_equal(km, z["kmap"])
    x = torch.from_numpy(z["feat"])
    w = torch.from_numpy(z["weight"])
    out = oracle.sparse_conv(x, w, km, km.shape[0], st)
    assert torch.allclose(out, torch.from_numpy(z["out"]), atol=1e-12, rtol=1e-12)


# --------------------------------------------------------------------------------------
# voxelisation of points, ijk -> index, dense read/write
# --------------------------------------------------------------------------------------


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_points_to_ijk_closed_form(dtype):
    """Points at voxel centres and at +-0.3 voxel from them map to that voxel; voxel sizes are
    powers of two so every value is exact.  A sign slip on the origin, a multiplied instead of
    divided voxel size or a floor instead of a round fails here."""
    rng = np.random.default_rng(0)
    for vs, og in ((0.25, (0.0, 0.0, 0.0)), (0.125, (-1.5, 2.0, 0.375)), (2.0, (3.0, -4.0, 10.0))):
        i = rng.integers(-40, 40, size=(200, 3))
        for d in (-0.3, 0.0, 0.3):
            p = (np.asarray(og) + (i + d) * vs).astype(dtype)
            ijk, tie = oracle.points_to_ijk(p, vs, og)
            assert np.array_equal(ijk, i), (vs, og, d)
            assert (tie >= 0.19).all()
    # exactly half-way rounds up (NanoVDB's floor(x + 0.5)); flagged as a tie
    ijk, tie = oracle.points_to_ijk(np.array([[0.5, -0.5, 1.5]], dtype), 1.0, 0.0)
    assert ijk.tolist() == [[1, 0, 2]] and (tie == 0).all()


def test_points_to_ijk_vs_exact_rationals():
    """float64 points away from ties: the same voxel as exact rational arithmetic gives."""
    from fractions import Fraction

    rng = np.random.default_rng(1)
    p = rng.uniform(-3, 3, size=(300, 3))
    vs, og = 1 / 96, (0.2, -0.7, 1.1)
    ijk, tie = oracle.points_to_ijk(p, vs, og)
    keep = (tie > 1e-6).all(1)
    assert keep.sum() > 280
    for row, q in zip(ijk[keep], p[keep]):
        exact = [math.floor((Fraction(float(q[d])) - Fraction(og[d])) / Fraction(vs) + Fraction(1, 2)) for d in range(3)]
        assert row.tolist() == exact


def test_ijk_to_index_invariants():
    rng = np.random.default_rng(2)
    grids = [oracle.sort_unique(random_voxels(rng, 300, 12)), oracle.sort_unique(random_voxels(rng, 200, 9, lo=-4))]
    n0 = len(grids[0])
    # every voxel finds itself (cumulative index across grids)
    assert np.array_equal(oracle.ijk_to_index(grids, grids), np.arange(n0 + len(grids[1])))
    # coordinates outside a grid's set are -1; grid 1's voxels are not looked up in grid 0
    far = [np.array([[100, 100, 100]], np.int32), grids[0][:5]]
    got = oracle.ijk_to_index(grids, far)
    in1 = {tuple(c) for c in grids[1].tolist()}
    exp = [-1] + [n0 + int(np.flatnonzero((grids[1] == c).all(1))[0]) if tuple(c) in in1 else -1
                  for c in grids[0][:5].tolist()]
    assert got.tolist() == exp


def test_dense_write_read_closed_form():
    """A dense box written at its own minimum is a reshape of the rows (x, y, z order); reading
    it back is the identity; a shifted frame crops and pads with zeros."""
    box = dense_box((4, 3, 5), lo=(2, -1, 7))
    g = oracle.sort_unique(box)
    rng = np.random.default_rng(3)
    data = rng.standard_normal((len(g), 2))
    lo = np.array([[2, -1, 7]])
    dense = oracle.write_to_dense([g], data, lo, (4, 3, 5))
    for n, c in enumerate(g.tolist()):
        assert np.array_equal(dense[0, c[0] - 2, c[1] + 1, c[2] - 7], data[n])
    assert np.array_equal(oracle.read_from_dense([g], dense, lo), data)
    shifted = oracle.write_to_dense([g], data, np.array([[3, -1, 7]]), (4, 3, 5))
    assert np.array_equal(shifted[0, :3], dense[0, 1:]) and not shifted[0, 3].any()
    back = oracle.read_from_dense([g], shifted, np.array([[3, -1, 7]]))
    assert np.array_equal(back[g[:, 0] >= 3], data[g[:, 0] >= 3]) and not back[g[:, 0] < 3].any()
