"""Seeded synthetic sparse grids for the benchmark workloads.

* ``sphere_points``: points sampled uniformly on random spheres (centres U[0,1]^3,
  radii U[0.1, 0.3]) -- the single-grid and batched sphere workloads.
* ``planes_spheres_points``: random planes plus spheres inside the unit cube.
* ``indoor_ijk``: axis-aligned room walls/floor plus random boxes (surfaces only),
  voxelised directly on a 2 cm lattice of a room of up to 10 x 10 x 3 m.

Points are generated on the device with a seeded ``torch.Generator`` and voxelised by
``GridBatch.from_points`` (``ijk = round((p - origin) / voxel_size)``) through ``voxelize``, whose
origin puts every generated point at a non-negative voxel index: the sphere generators reach
down to -0.3 (centres >= 0, radii <= 0.3), so with origin 0 a grid would straddle the 4096^3
root tiles around 0.  ``voxelize`` keeps each workload grid inside the root tile [0, 4096)^3, as
the reference's benchmark dataset shifts its grids to non-negative coordinates
(``tests/benchmark/fvdb_benchmark/dataset.py:77-80``).
"""
from __future__ import annotations

import math
from typing import List, Sequence, Tuple

import torch


# Lowest coordinate of any point the generators below produce (sphere centres in [0, 1]^3,
# radii up to 0.3; planes are clipped to the unit cube).
POINTS_LO = -0.3
ROOT_TILE = 4096


def nonnegative_origin(voxel_size: float, lo: float = POINTS_LO) -> List[float]:
    """Origin (world position of voxel (0, 0, 0)) at which every point with coordinates >= `lo`
    rounds to a voxel index >= 1."""
    o = float(lo) - float(voxel_size)
    return [o, o, o]


def voxelize(points, voxel_sizes, device=None):
    """GridBatch of `points` (a tensor or a list of tensors, one per grid) at `voxel_sizes`
    (one float, or one per grid) with ``nonnegative_origin`` origins.  Raises if a voxel
    lands outside [0, 4096)^3."""
    import fvdb

    pts = points if isinstance(points, (list, tuple)) else [points]
    vs = [float(v) for v in voxel_sizes] if isinstance(voxel_sizes, (list, tuple)) else [float(voxel_sizes)] * len(pts)
    if len(vs) != len(pts):
        raise ValueError("one voxel size per grid")
    arg = fvdb.JaggedTensor(list(pts)) if isinstance(points, (list, tuple)) else points
    grid = fvdb.GridBatch.from_points(arg, voxel_sizes=[[v, v, v] for v in vs],
                                      origins=[nonnegative_origin(v) for v in vs], device=device)
    check_root_tile(grid.ijk.jdata)
    return grid


def check_root_tile(ijk: torch.Tensor) -> None:
    """Raises ValueError unless every coordinate lies in [0, 4096)."""
    if ijk.numel() and (int(ijk.min()) < 0 or int(ijk.max()) >= ROOT_TILE):
        raise ValueError(f"voxel coordinates [{int(ijk.min())}, {int(ijk.max())}] leave the root tile [0, 4096)")


def _gen(seed: int, device) -> torch.Generator:
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def sphere_points(seed: int, n_points: int = 2_000_000, n_spheres: int = 8, device="cuda") -> torch.Tensor:
    g = _gen(seed, device)
    c = torch.rand((n_spheres, 3), generator=g, device=device, dtype=torch.float64)
    r = 0.1 + 0.2 * torch.rand((n_spheres,), generator=g, device=device, dtype=torch.float64)
    per = n_points // n_spheres
    d = torch.randn((n_spheres, per, 3), generator=g, device=device, dtype=torch.float64)
    d = d / d.norm(dim=-1, keepdim=True)
    p = c[:, None, :] + r[:, None, None] * d
    return p.reshape(-1, 3).float()


def planes_spheres_points(seed: int, n_points: int = 8_000_000, n_planes: int = 4, n_spheres: int = 4,
                          device="cuda") -> torch.Tensor:
    g = _gen(seed, device)
    per = n_points // (n_planes + n_spheres)
    parts = []
    for _ in range(n_planes):
        n = torch.randn(3, generator=g, device=device, dtype=torch.float64)
        n = n / n.norm()
        o = torch.rand(3, generator=g, device=device, dtype=torch.float64)
        a = torch.linalg.cross(n, torch.tensor([0.0, 0.0, 1.0], device=device, dtype=torch.float64))
        if a.norm() < 1e-3:
            a = torch.tensor([1.0, 0.0, 0.0], device=device, dtype=torch.float64)
        a = a / a.norm()
        b = torch.linalg.cross(n, a)
        uv = torch.rand((per, 2), generator=g, device=device, dtype=torch.float64) - 0.5
        p = o + uv[:, :1] * a + uv[:, 1:] * b
        parts.append(p[((p >= 0) & (p <= 1)).all(1)])
    for _ in range(n_spheres):
        c = torch.rand(3, generator=g, device=device, dtype=torch.float64)
        r = 0.1 + 0.2 * torch.rand((), generator=g, device=device, dtype=torch.float64)
        d = torch.randn((per, 3), generator=g, device=device, dtype=torch.float64)
        parts.append(c + r * d / d.norm(dim=-1, keepdim=True))
    return torch.cat(parts).float()


def indoor_ijk(seed: int, n_boxes: int = 20, room=None, device="cuda") -> torch.Tensor:
    """Surface voxels of a room (floor, ceiling, 4 walls) and random boxes on a 2 cm lattice.

    The room is at most 10 x 10 x 3 m (500 x 500 x 150 cells); without an explicit
    `room` its footprint is drawn per seed (5-10 m a side), which spreads the voxel
    counts of a batch over roughly 0.4M-1.2M.
    """
    g = _gen(seed, device)
    if room is None:
        dims = torch.rand(2, generator=g, device=device)
        room = (int(250 + 250 * dims[0].item()), int(250 + 250 * dims[1].item()), 150)
    X, Y, Z = room
    faces: List[torch.Tensor] = []

    def rect(axis: int, at: int, lo: Sequence[int], hi: Sequence[int]):
        # all lattice cells of an axis-aligned rectangle at coordinate `at` along `axis`
        dims = [d for d in range(3) if d != axis]
        a = torch.arange(lo[0], hi[0] + 1, device=device, dtype=torch.int32)
        b = torch.arange(lo[1], hi[1] + 1, device=device, dtype=torch.int32)
        A, B = torch.meshgrid(a, b, indexing="ij")
        out = torch.empty((A.numel(), 3), dtype=torch.int32, device=device)
        out[:, axis] = at
        out[:, dims[0]] = A.reshape(-1)
        out[:, dims[1]] = B.reshape(-1)
        faces.append(out)

    def box(lo, hi):
        for axis in range(3):
            dims = [d for d in range(3) if d != axis]
            for at in (lo[axis], hi[axis]):
                rect(axis, at, [lo[dims[0]], lo[dims[1]]], [hi[dims[0]], hi[dims[1]]])

    # room shell (open at the top half of one wall to vary topology a little)
    rect(2, 0, [0, 0], [X - 1, Y - 1])
    rect(2, Z - 1, [0, 0], [X - 1, Y - 1])
    rect(0, 0, [0, 0], [Y - 1, Z - 1])
    rect(0, X - 1, [0, 0], [Y - 1, Z - 1])
    rect(1, 0, [0, 0], [X - 1, Z - 1])
    rect(1, Y - 1, [0, 0], [X - 1, Z // 2])
    for _ in range(n_boxes):
        size = (torch.rand(3, generator=g, device=device) *
                torch.tensor([X * 0.24, Y * 0.24, 100.0], device=device) + 10).int()
        lo = (torch.rand(3, generator=g, device=device) * (torch.tensor([X, Y, Z], device=device) - size)).int()
        lo[2] = 0 if bool(torch.rand((), generator=g, device=device) < 0.6) else lo[2]
        hi = lo + size
        hi = torch.minimum(hi, torch.tensor([X - 1, Y - 1, Z - 1], device=device, dtype=torch.int32))
        box(lo.tolist(), hi.tolist())
    return torch.cat(faces)


def log_uniform(seed: int, lo: float, hi: float, n: int) -> List[float]:
    g = torch.Generator().manual_seed(int(seed))
    u = torch.rand(n, generator=g, dtype=torch.float64)
    return torch.exp(math.log(lo) + u * (math.log(hi) - math.log(lo))).tolist()


def sphere_points_for_voxels(seed: int, target_voxels: int, n_spheres: int = 8, device="cuda"):
    """Points on random spheres and a voxel size that gives about `target_voxels` voxels.

    A thin spherical shell of area A voxelised at size v covers about 1.16 A / v^2
    voxels (measured); 6 A / v^2 points cover it almost completely.
    Returns (points [n, 3] float32, voxel size).
    """
    g = _gen(seed, device)
    c = torch.rand((n_spheres, 3), generator=g, device=device, dtype=torch.float64)
    r = 0.1 + 0.2 * torch.rand((n_spheres,), generator=g, device=device, dtype=torch.float64)
    area = float((4 * math.pi * r ** 2).sum())
    vs = math.sqrt(1.16 * area / max(int(target_voxels), 1))
    per = max(1, int(6 * area / vs ** 2) // n_spheres)
    d = torch.randn((n_spheres, per, 3), generator=g, device=device, dtype=torch.float64)
    d = d / d.norm(dim=-1, keepdim=True)
    p = c[:, None, :] + r[:, None, None] * d
    return p.reshape(-1, 3).float(), vs
