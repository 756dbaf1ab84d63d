"""Writes the golden fixtures under tests/golden from the CPU oracle.

The reference ships no golden vectors for the sparse-conv path (its conv tests draw
seeded random data and compare with dense torch conv3d or torchsparse), so these
fixtures are produced here by the oracle after it has been pinned by the
closed-form, brute-force and dense-conv3d checks in tests/test_oracle.py.  They
freeze the oracle's voxel order, target grid, kernel map and float64 outputs, and
they are the data the GPU parity tests compare against bit for bit (integer) or
within tolerance (float).

Usage:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402


def sphere_shell(rng, n_points, voxel_size, shift):
    c = rng.uniform(0, 1, 3)
    r = rng.uniform(0.1, 0.3)
    d = rng.normal(size=(n_points, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p = c + r * d
    return (np.floor(p / voxel_size + 0.5).astype(np.int64) + shift).astype(np.int32)


def write_case(name, raw_grids, ks, stride, cin, cout, seed):
    rng = np.random.default_rng(seed)
    srcs = [oracle.sort_unique(g) for g in raw_grids]
    st = (stride,) * 3 if isinstance(stride, int) else tuple(stride)
    if st == (1, 1, 1):
        tgts = srcs
    else:
        tgts = [oracle.conv_grid(s, ks, stride) for s in srcs]
    km = oracle.kernel_map(srcs, tgts, ks, stride)
    ksv = (ks,) * 3 if isinstance(ks, int) else tuple(ks)
    n_src = sum(s.shape[0] for s in srcs)
    feat = rng.normal(size=(n_src, cin))
    w = rng.normal(size=(cout, cin) + ksv) * np.sqrt(2.0 / (cin * np.prod(ksv)))
    out = oracle.sparse_conv(torch.from_numpy(feat), torch.from_numpy(w), km, km.shape[0], stride).numpy()
    raw = np.concatenate(raw_grids)
    offs = np.cumsum([0] + [g.shape[0] for g in raw_grids]).astype(np.int64)
    toffs = np.cumsum([0] + [t.shape[0] for t in tgts]).astype(np.int64)
    np.savez_compressed(
        os.path.join(HERE, name),
        src_ijk_raw=raw,
        src_joffsets=offs,
        src_ijk=np.concatenate(srcs),
        src_voxel_joffsets=np.cumsum([0] + [s.shape[0] for s in srcs]).astype(np.int64),
        tgt_ijk=np.concatenate(tgts),
        tgt_joffsets=toffs,
        kernel_size=np.asarray(ksv, np.int32),
        stride=np.asarray(st, np.int32),
        kmap=km,
        feat=feat,
        weight=w,
        out=out,
    )
    print(f"{name}: src {n_src} voxels, tgt {km.shape[0]}, K {km.shape[1]}, pairs {(km >= 0).sum()}")


def main():
    rng = np.random.default_rng(1234)
    # two small sphere shells at 1/64 (about 1-3k voxels), one grid per case plus a batch
    g1 = np.concatenate([sphere_shell(rng, 6000, 1 / 64, 8) for _ in range(2)])
    g2 = sphere_shell(rng, 3000, 1 / 48, 4)
    g3 = rng.integers(0, 6, size=(40, 3)).astype(np.int32)
    write_case("sphere_k3s1_c8.npz", [g1], 3, 1, 8, 8, 1)
    write_case("batch_k3s1_c4x6.npz", [g2, g3, g1], 3, 1, 4, 6, 2)
    write_case("sphere_k2s2_c8x16.npz", [g1, g2], 2, 2, 8, 16, 3)
    write_case("sphere_k3s2_c4x4.npz", [g2], 3, 2, 4, 4, 4)
    write_case("aniso_k513_s132.npz", [g3, g2], (5, 1, 3), (1, 3, 2), 3, 5, 5)
    # negative coordinates spanning two root tiles in x: order follows the unsigned root key
    neg = rng.integers(-20, 20, size=(500, 3)).astype(np.int32)
    write_case("negative_k3s1_c2.npz", [neg], 3, 1, 2, 3, 6)


if __name__ == "__main__":
    main()
