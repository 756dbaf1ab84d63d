#!/usr/bin/env python
"""Sparse-convolution benchmark on one MI355X.

Workload (the batched precision workload): a batch of 8 grids, each 2M points sampled
on 8 random spheres (seeds 1..8, centres U[0,1]^3, radii U[0.1,0.3]) voxelised at a
voxel size drawn log-uniformly from [1/512, 1/128] (about 50k-600k voxels per grid);
conv A = 3x3x3 stride-1 submanifold conv 64 -> 64, then conv B = 2x2x2 stride-2
downsampling conv 64 -> 128 on conv A's output.  Features N(0,1), Kaiming-normal
weights.

One step = the whole path over the batch: kernel map A, conv A, target grid of B
(floor(ijk/2)), kernel map B, conv B.  `value` is input voxels per second of the BF16
step (MFMA, fp32 accumulation).  Also reported: the FP32 step, conv-only timings and
their BF16/FP32 ratio, effective TFLOPS (2*pairs*Cin*Cout per second), the error of
BF16/FP32 outputs against the float64 CPU oracle on a sample grid, the roofline of
the dominant kernel (conv A's implicit GEMM, timed with HIP events on its stream),
and the CPU baseline (the oracle's restatement of the reference CPU path, fp32).

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "fvdb-core_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "sparse-conv voxels/s + TFLOPS, BF16 vs FP32 at ≤1e-2 rel err; serve p50/p99"
PEAK = {"hbm_GBs": 8000.0, "bf16_TFLOPs": 2500.0, "fp16_TFLOPs": 2500.0, "fp32_TFLOPs": 157.3}
COPY_BYTES = 256 * 1024 * 1024  # calibration copy for the traffic counters
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "conv_tile_bf16_sphere_batch8_traffic.json")


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_workload(torch, fvdb, device):
    from fvdb.utils.synthetic import log_uniform, sphere_points

    vs = log_uniform(0, 1 / 512, 1 / 128, 8)
    pts = [sphere_points(seed=s + 1, n_points=2_000_000, device=device) for s in range(8)]
    grid = fvdb.GridBatch.from_points(fvdb.JaggedTensor(pts), voxel_sizes=[[v, v, v] for v in vs], origins=0.0)
    del pts
    g = torch.Generator(device=device).manual_seed(1234)
    N = grid.total_voxels
    x = torch.randn((N, 64), generator=g, device=device)
    wa = torch.randn((64, 64, 3, 3, 3), generator=g, device=device) * math.sqrt(2.0 / (64 * 27))
    wb = torch.randn((128, 64, 2, 2, 2), generator=g, device=device) * math.sqrt(2.0 / (64 * 8))
    return grid, vs, x, wa, wb


def run_step(fvdb, grid, x, wa, wb):
    IG = fvdb.ConvPackBackend.IGEMM
    pa, _ = grid.sparse_conv_kernel_map(3, 1)
    pa.build_implicit_gemm()
    ya = pa.sparse_conv_3d(grid.jagged_like(x), wa, IG)
    pb, coarse = grid.sparse_conv_kernel_map(2, 2)
    pb.build_implicit_gemm()
    yb = pb.sparse_conv_3d(ya, wb, IG)
    return pa, pb, ya, yb


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def event_avg_ms(torch, fn, iters):
    """Average duration of `fn`'s launches, HIP events on the current stream (the
    stream the extension launches on)."""
    fn()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(iters):
        s.record()
        fn()
        e.record()
        e.synchronize()
        times.append(s.elapsed_time(e))
    return statistics.median(times), times


def rel_errors(torch, out, ref):
    d = (out.double().cpu() - ref).abs()
    return {"max_rel": float(d.max() / ref.abs().max()),
            "rel_l2": float(torch.linalg.vector_norm(out.double().cpu() - ref) / torch.linalg.vector_norm(ref))}


def check_errors(torch, fvdb, grid, x, wa, wb, sample_b):
    """BF16/FP32 vs the float64 oracle on grid `sample_b` of the batch."""
    import numpy as np

    import oracle

    jo = grid.joffsets.cpu().tolist()
    lo, hi = jo[sample_b], jo[sample_b + 1]
    ijk = grid.ijk.jdata[lo:hi].cpu().numpy()
    km_a = oracle.kernel_map([ijk], [ijk], 3, 1)
    tgt = oracle.conv_grid(ijk, 2, 2)
    km_b = oracle.kernel_map([ijk], [tgt], 2, 2)
    res = {}
    for dt, name in ((torch.bfloat16, "bf16"), (torch.float16, "fp16"), (torch.float32, "fp32")):
        pa, pb, ya, yb = run_step(fvdb, grid, x.to(dt), wa.to(dt), wb.to(dt))
        xs = x.to(dt)[lo:hi].double().cpu()
        ref_a = oracle.sparse_conv(xs, wa.to(dt).double().cpu(), km_a, ijk.shape[0], 1)
        ya_s = ya.jdata[lo:hi]
        e_a = rel_errors(torch, ya_s, ref_a)
        # conv B fed with this dtype's conv-A output, so only conv B's error is measured
        ref_b = oracle.sparse_conv(ya_s.double().cpu(), wb.to(dt).double().cpu(), km_b, tgt.shape[0], 2)
        jb = pb.target_grid.joffsets.cpu().tolist()
        yb_s = yb.jdata[jb[sample_b]:jb[sample_b + 1]]
        e_b = rel_errors(torch, yb_s, ref_b)
        # the GPU kernel map of the sample grid must equal the oracle's bit for bit
        gm = pa.out_in_map[lo:hi].cpu().numpy()
        gm = np.where(gm >= 0, gm - lo, -1)
        res[name] = {"convA": e_a, "convB": e_b, "kmap_bit_exact": bool(np.array_equal(gm, km_a))}
    return res, int(hi - lo)


def cpu_baseline(torch, fvdb, grid, x, wa, wb, budget_s=10.0):
    """The oracle's restatement of the reference CPU path (C kernel map, torch-CPU
    gather-GEMM-scatter) in fp32 over grids of the batch, smallest first, until about
    `budget_s` seconds of CPU work."""
    import oracle

    jo = grid.joffsets.cpu().tolist()
    sizes = [(jo[b + 1] - jo[b], b) for b in range(grid.grid_count)]
    sizes.sort()
    wa_c, wb_c = wa.float().cpu(), wb.float().cpu()
    tot_t, tot_v, used = 0.0, 0, []
    warm = True
    for n, b in sizes:
        ijk = grid.ijk.jdata[jo[b]:jo[b + 1]].cpu().numpy()
        xs = x[jo[b]:jo[b + 1]].float().cpu()

        def step():
            km_a = oracle.kernel_map([ijk], [ijk], 3, 1)
            ya = oracle.sparse_conv(xs, wa_c, km_a, ijk.shape[0], 1, dtype=torch.float32)
            tgt = oracle.conv_grid(ijk, 2, 2)
            km_b = oracle.kernel_map([ijk], [tgt], 2, 2)
            return oracle.sparse_conv(ya, wb_c, km_b, tgt.shape[0], 2, dtype=torch.float32)

        if warm:
            step()
            warm = False
        t0 = time.perf_counter()
        step()
        dt = time.perf_counter() - t0
        tot_t += dt
        tot_v += n
        used.append(b)
        if tot_t >= budget_s:
            break
    return {"value": tot_v / tot_t, "unit": "voxels/s", "cores": torch.get_num_threads(), "kind": "port",
            "sample": f"grids {used} of the batch ({tot_v} voxels), one fp32 step each: C kernel maps + "
                      f"torch-CPU gather-GEMM-scatter (reference CPU algorithm), {tot_t:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only loop the dominant kernel (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    import torch

    import fvdb

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log("[bench] building workload")
    grid, vs, x, wa, wb = build_workload(torch, fvdb, dev)
    N = grid.total_voxels
    counts = grid.num_voxels.tolist()
    log(f"[bench] {grid.grid_count} grids, {N} voxels ({counts})")

    xb, wab, wbb = x.bfloat16(), wa.bfloat16(), wb.bfloat16()
    pa, pb, ya, yb = run_step(fvdb, grid, xb, wab, wbb)
    P_a = int((pa.out_in_map >= 0).sum())
    P_b = int((pb.out_in_map >= 0).sum())
    N_b = pb.target_grid.total_voxels
    flops = 2.0 * P_a * 64 * 64 + 2.0 * P_b * 64 * 128
    flops_a = 2.0 * P_a * 64 * 64

    # dominant kernel: conv A implicit GEMM (bf16), maps prebuilt
    kio = fvdb._autograd.weight_to_kio(wab, False).transpose(1, 2).contiguous()
    fmap = pa._map_fwd()
    conv_a = lambda: fvdb.ext.implicit_gemm(xb, kio, fmap, N)  # noqa: E731
    if args.kernel_only:
        # a plain device copy of known size first: its FETCH_SIZE / WRITE_SIZE readings
        # calibrate the HBM-traffic counters of the conv kernel
        src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        for i in range(args.steps):
            conv_a()
            torch.cuda.synchronize()
            log(f"[bench] kernel iter {i}")
        return 0

    log("[bench] timing bf16 step")
    t_bf16 = time_steps(torch, lambda: run_step(fvdb, grid, xb, wab, wbb), args.steps, args.warmup)
    log("[bench] timing fp32 step")
    t_fp32 = time_steps(torch, lambda: run_step(fvdb, grid, x, wa, wb), args.steps, args.warmup)

    # conv-only (maps built once), per dtype
    def conv_only(dt):
        xa, w1, w2 = x.to(dt), wa.to(dt), wb.to(dt)
        IG = fvdb.ConvPackBackend.IGEMM
        xj = grid.jagged_like(xa)
        return lambda: pb.sparse_conv_3d(pa.sparse_conv_3d(xj, w1, IG), w2, IG)

    log("[bench] timing conv-only")
    tc = {name: time_steps(torch, conv_only(dt), args.steps, args.warmup)
          for name, dt in (("bf16", torch.bfloat16), ("fp16", torch.float16), ("fp32", torch.float32))}

    log("[bench] timing dominant kernel with HIP events")
    k_ms, _ = event_avg_ms(torch, conv_a, max(10, args.steps))
    kio32 = fvdb._autograd.weight_to_kio(wa, False).transpose(1, 2).contiguous()
    k32_ms, _ = event_avg_ms(torch, lambda: fvdb.ext.implicit_gemm(x, kio32, fmap, N), max(10, args.steps))

    # algorithmic bytes of conv A (bf16): input rows + output rows + map + weights
    bytes_a = N * 64 * 2 + N * 64 * 2 + 4 * 27 * N + 27 * 64 * 64 * 2
    t_min_mfma = flops_a / (PEAK["bf16_TFLOPs"] * 1e12)
    t_min_hbm = bytes_a / (PEAK["hbm_GBs"] * 1e9)
    if t_min_mfma >= t_min_hbm:
        roof = {"bound": "mfma", "achieved": flops_a / (k_ms * 1e-3) / 1e12, "peak": PEAK["bf16_TFLOPs"],
                "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": bytes_a / (k_ms * 1e-3) / 1e9, "peak": PEAK["hbm_GBs"],
                "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None
    if os.path.exists(TRAFFIC_FILE):
        # HBM bytes per launch of this kernel from rocprofv3 FETCH_SIZE/WRITE_SIZE runs,
        # each scaled by the same counter's reading for a copy of known size
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        roof["traffic"] = tr["hbm_bytes_per_launch"]
        roof["traffic_source"] = os.path.relpath(TRAFFIC_FILE, ROOT)
    roof["kernel"] = "k_conv_tile_k<bf16, 4, 2, 64, 27, 2> (conv A, 64->64, 3x3x3)"
    roof["avg_launch_ms"] = k_ms
    roof["algorithmic_flops"] = flops_a
    roof["algorithmic_bytes"] = bytes_a

    errors, sample_n = ({}, 0) if args.no_check else check_errors(torch, fvdb, grid, x, wa, wb, 0)
    cpu = None if args.no_cpu_baseline else cpu_baseline(torch, fvdb, grid, x, wa, wb)

    value = N / t_bf16
    rec = {
        "metric": METRIC,
        "value": value,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_bf16 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": "8 sphere grids (2M points each, voxel size 1/512..1/128), conv 3x3x3 s1 64->64 "
                               "then 2x2x2 s2 64->128, kernel maps + target grid rebuilt every step",
                   "grids": grid.grid_count, "voxels": N, "voxels_per_grid": counts, "voxels_convB_out": N_b,
                   "pairs_convA": P_a, "pairs_convB": P_b, "voxel_sizes": vs},
        "tflops_step_bf16": flops / t_bf16 / 1e12,
        "fp32_step_voxels_per_s": N / t_fp32,
        "fp32_step_ms": t_fp32 * 1e3,
        "conv_only_ms": {k: v * 1e3 for k, v in tc.items()},
        "conv_only_voxels_per_s": {k: N / v for k, v in tc.items()},
        "conv_only_tflops": {k: flops / v / 1e12 for k, v in tc.items()},
        "bf16_over_fp32_conv_speedup": tc["fp32"] / tc["bf16"],
        "convA_kernel_ms": {"bf16": k_ms, "fp32": k32_ms},
        "convA_kernel_tflops": {"bf16": flops_a / (k_ms * 1e-3) / 1e12, "fp32": flops_a / (k32_ms * 1e-3) / 1e12},
        "errors_vs_fp64_oracle": errors,
        "error_sample_voxels": sample_n,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
