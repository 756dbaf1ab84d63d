"""GPU parity on the benchmark workloads themselves, at their full sizes.

* Every workload generator lands in the root tile [0, 4096)^3, where the voxel order is
  the 3-level tree order the oracle restates and the golden fixtures pin (no grid of a
  workload relies on the order across root tiles).
* Voxelisation of points (``from_points``) against the oracle's restatement of
  ``BuildGridFromPoints.cu:28-44`` in float32 and float64.
* Kernel maps, target grids and neighbour maps of the single-grid, precision and
  differentiation workloads bit-exact against the oracle.
* The single-grid workload's fp32 32 -> 32 forward within 1e-5 of the float64 oracle, and the
  precision workload's two convolutions within the bf16 / fp16 / fp32 bounds on a sample grid.

Tolerances: max|gpu - ref| / max|ref| against the float64 oracle fed the same rounded
inputs -- fp32 1e-5, fp16 2e-3, bf16 1e-2.
"""
import math

import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu

REL_TOL = {torch.float32: 1e-5, torch.float16: 2e-3, torch.bfloat16: 1e-2}


def _fvdb():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb

    return fvdb


def rel_err(a, ref) -> float:
    a = a.detach().double().cpu()
    ref = ref.detach().double().cpu()
    return (a - ref).abs().max().item() / ref.abs().max().item()


def grid_parts(grid):
    jo = grid.joffsets.cpu().tolist()
    ijk = grid.ijk.jdata.cpu().numpy()
    return [ijk[jo[b]:jo[b + 1]] for b in range(grid.grid_count)]


def sphere_batch(fvdb):
    """The precision workload: 8 sphere grids, voxel sizes log-uniform in [1/512, 1/128]."""
    from fvdb.utils.synthetic import log_uniform, sphere_points, voxelize

    vs = log_uniform(0, 1 / 512, 1 / 128, 8)
    pts = [sphere_points(seed=s + 1, n_points=2_000_000, device="cuda") for s in range(8)]
    return voxelize(pts, vs)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_from_points_vs_oracle(dtype):
    """Per-point voxel coordinates of a 3-grid batch (different voxel sizes and origins) equal
    the oracle's restatement, away from rounding ties (|frac - 0.5| < 1e-3, where a fused
    multiply-add and two roundings may differ: parity unpinned there, as NanoVDB's
    ``Vec3::round`` tie rule is); the built grids hold exactly the unique coordinates."""
    fvdb = _fvdb()
    g = torch.Generator().manual_seed(4)
    vs = [0.013, 0.5 / 37, 0.021]
    og = [[-0.41, 0.2, 0.003], [0.0, 0.0, 0.0], [1.7, -2.3, 0.9]]
    pts = [(torch.rand(20_000, 3, generator=g, dtype=torch.float64) * 4 - 2).to(dtype) for _ in range(3)]
    np_ijk, keep = [], []
    for p, v, o in zip(pts, vs, og):
        ijk, tie = oracle.points_to_ijk(p.numpy(), v, o)
        np_ijk.append(ijk)
        keep.append((tie > 1e-3).all(1))
    jt = fvdb.JaggedTensor([p.cuda() for p in pts])
    xf = torch.cat([1.0 / torch.tensor(vs, dtype=torch.float64)[:, None].expand(3, 3),
                    -torch.tensor(og, dtype=torch.float64) / torch.tensor(vs, dtype=torch.float64)[:, None]], 1)
    gpu = fvdb.ext.points_to_ijk(jt.jdata, jt.jidx.to(torch.int32), xf.cuda()).cpu().numpy()
    ref = np.concatenate(np_ijk)
    k = np.concatenate(keep)
    print(f"[from_points] {dtype}: {k.sum()} of {len(k)} points away from ties", flush=True)
    assert k.mean() > 0.99
    assert np.array_equal(gpu[k], ref[k])
    # the grid build: points with a tie coordinate left out, the rest voxelised
    kept = [p[kk] for p, kk in zip(pts, keep)]
    grid = fvdb.GridBatch.from_points(fvdb.JaggedTensor([p.cuda() for p in kept]),
                                      voxel_sizes=[[v, v, v] for v in vs], origins=og)
    for b, part in enumerate(grid_parts(grid)):
        assert np.array_equal(part, oracle.sort_unique(np_ijk[b][keep[b]]))


def test_workload_grids_in_root_tile_and_kmaps_bit_exact():
    """Each workload's grids lie in [0, 4096)^3; their kernel maps, strided target grids and
    transposed maps equal the oracle's bit for bit, at full size."""
    fvdb = _fvdb()
    from fvdb.utils.synthetic import (indoor_ijk, planes_spheres_points, sphere_points, sphere_points_for_voxels,
                                      voxelize)

    def check(name, grid, cases):
        ijk = grid.ijk.jdata
        assert int(ijk.min()) >= 0 and int(ijk.max()) < 4096, name
        parts = grid_parts(grid)
        for b, part in enumerate(parts):
            assert np.array_equal(part, oracle.sort_unique(part)), (name, b)
        for ks, st in cases:
            print(f"[workload] {name}: {grid.total_voxels} voxels, ks={ks} stride={st}", flush=True)
            pack, tgt = grid.sparse_conv_kernel_map(ks, st)
            pack.build_implicit_gemm(training=True)
            tparts = [oracle.conv_grid(p, ks, st) for p in parts] if st != 1 else parts
            assert np.array_equal(tgt.ijk.jdata.cpu().numpy(), np.concatenate(tparts)), name
            km = oracle.kernel_map(parts, tparts, ks, st)
            n_t = tgt.total_voxels
            assert np.array_equal(pack.out_in_map[:n_t].cpu().numpy(), km), name
            assert np.array_equal(pack.out_in_map_bwd[:grid.total_voxels].cpu().numpy(),
                                  oracle.transpose_map(km, grid.total_voxels)), name

    check("single grid (2M sphere points at 1/256)",
          voxelize(sphere_points(seed=0, n_points=2_000_000, device="cuda"), 1 / 256), [(3, 1)])
    check("precision batch (8 sphere grids)", sphere_batch(fvdb), [(3, 1), (2, 2)])
    check("differentiation grid (planes + spheres at 1/1024)",
          voxelize(planes_spheres_points(seed=7, n_points=2_400_000, n_planes=2, n_spheres=2, device="cuda"),
                   1 / 1024), [(3, 1)])
    parts = [indoor_ijk(s, device="cuda") for s in (1, 2, 3, 4)]
    check("U-Net batch (4 indoor rooms at 2 cm)", fvdb.GridBatch.from_ijk(fvdb.JaggedTensor(parts)),
          [(3, 1), (2, 2)])
    # serving requests: the smallest and the largest of the log-uniform 10k-2M range
    for nv in (10_000, 2_000_000):
        pts, vs = sphere_points_for_voxels(1000 + nv % 7, nv, device="cuda")
        check(f"serving request (~{nv} voxels)", voxelize(pts, vs), [(3, 1)])


def test_single_grid_fp32_forward_vs_oracle():
    """The single-grid workload at full size (2M sphere points at 1/256, 3x3x3 32 -> 32): fp32
    (the contraction in tap-mask order per window, as the pack picks it) within 1e-5 and bf16
    within 1e-2 of the float64 oracle."""
    fvdb = _fvdb()
    from fvdb.utils.synthetic import sphere_points, voxelize

    grid = voxelize(sphere_points(seed=0, n_points=2_000_000, device="cuda"), 1 / 256)
    N = grid.total_voxels
    pack, _ = grid.sparse_conv_kernel_map(3, 1)
    pack.build_implicit_gemm()
    assert pack._order_for(True, 32, torch.float32, 32) is not None  # the workload's fp32 path
    g = torch.Generator().manual_seed(0)
    x = torch.randn(N, 32, generator=g)
    w = torch.randn(32, 32, 3, 3, 3, generator=g) * math.sqrt(2.0 / (32 * 27))
    km = pack.out_in_map[:N].cpu().numpy()
    for dt in (torch.float32, torch.bfloat16):
        xd, wd = x.to("cuda", dt), w.to("cuda", dt)
        out = pack.sparse_conv_3d(grid.jagged_like(xd), wd, fvdb.ConvPackBackend.IGEMM).jdata
        ref = oracle.sparse_conv(xd.double().cpu(), wd.double().cpu(), km, N, 1)
        e = rel_err(out, ref)
        print(f"[single grid] {N} voxels {dt}: rel err {e:.2e}", flush=True)
        assert e <= REL_TOL[dt]


def test_precision_workload_vs_oracle():
    """The precision workload's conv A (3x3x3 64 -> 64) and conv B (2x2x2 stride 2, 64 -> 128) on
    grid 0 of the full batch, in fp32, fp16 and bf16, against the float64 oracle (conv B fed
    the same dtype's conv A output)."""
    fvdb = _fvdb()
    grid = sphere_batch(fvdb)
    g = torch.Generator().manual_seed(1234)
    N = grid.total_voxels
    x = torch.randn(N, 64, generator=g)
    wa = torch.randn(64, 64, 3, 3, 3, generator=g) * math.sqrt(2.0 / (64 * 27))
    wb = torch.randn(128, 64, 2, 2, 2, generator=g) * math.sqrt(2.0 / (64 * 8))
    pa, _ = grid.sparse_conv_kernel_map(3, 1)
    pa.build_implicit_gemm()
    pb, coarse = grid.sparse_conv_kernel_map(2, 2)
    pb.build_implicit_gemm()
    jo, jb = grid.joffsets.cpu().tolist(), coarse.joffsets.cpu().tolist()
    lo, hi = jo[0], jo[1]
    ijk = grid.ijk.jdata[lo:hi].cpu().numpy()
    km_a = oracle.kernel_map([ijk], [ijk], 3, 1)
    tgt = oracle.conv_grid(ijk, 2, 2)
    km_b = oracle.kernel_map([ijk], [tgt], 2, 2)
    for dt in (torch.float32, torch.float16, torch.bfloat16):
        IG = fvdb.ConvPackBackend.IGEMM
        ya = pa.sparse_conv_3d(grid.jagged_like(x.to("cuda", dt)), wa.to("cuda", dt), IG)
        yb = pb.sparse_conv_3d(ya, wb.to("cuda", dt), IG).jdata
        ya_s = ya.jdata[lo:hi]
        ref_a = oracle.sparse_conv(x[lo:hi].to(dt).double(), wa.to(dt).double(), km_a, ijk.shape[0], 1)
        ref_b = oracle.sparse_conv(ya_s.double().cpu(), wb.to(dt).double(), km_b, tgt.shape[0], 2)
        ea, eb = rel_err(ya_s, ref_a), rel_err(yb[jb[0]:jb[1]], ref_b)
        print(f"[precision] grid 0 ({hi - lo} voxels) {dt}: conv A {ea:.2e} conv B {eb:.2e}", flush=True)
        assert ea <= REL_TOL[dt] and eb <= REL_TOL[dt]
