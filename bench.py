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
import gc
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
    from fvdb.utils.synthetic import log_uniform, sphere_points, voxelize

    vs = log_uniform(0, 1 / 512, 1 / 128, 8)
    pts = [sphere_points(seed=s + 1, n_points=2_000_000, device=device) for s in range(8)]
    # origins that put every voxel in the root tile [0, 4096)^3 (checked by voxelize)
    grid = voxelize(pts, vs)
    del pts
    g = torch.Generator(device=device).manual_seed(1234)
    N = grid.total_voxels
    x = torch.randn((N, 64), generator=g, device=device)
    wa = torch.randn((64, 64, 3, 3, 3), generator=g, device=device) * math.sqrt(2.0 / (64 * 27))
    wb = torch.randn((128, 64, 2, 2, 2), generator=g, device=device) * math.sqrt(2.0 / (64 * 8))
    return grid, vs, x, wa, wb


def run_step(fvdb, grid, x, wa, wb, side=None):
    """One step: kernel map A, conv A, target grid + kernel map B, conv B.

    With a `side` stream the target grid and map B (which do not depend on conv A)
    are built there while conv A runs on the current stream; the current stream then
    waits for them, and the map, allocated on the side stream, is recorded as used by
    the current stream so its memory is not reused before conv B has read it.  `grid`
    must be complete on both streams when the step starts.
    """
    import torch

    IG = fvdb.ConvPackBackend.IGEMM
    pa, _ = grid.sparse_conv_kernel_map(3, 1)
    pa.build_implicit_gemm()
    ya = pa.sparse_conv_3d(grid.jagged_like(x), wa, IG)
    if side is None:
        pb, coarse = grid.sparse_conv_kernel_map(2, 2)
        pb.build_implicit_gemm()
    else:
        main = torch.cuda.current_stream()
        with torch.cuda.stream(side):
            pb, coarse = grid.sparse_conv_kernel_map(2, 2)
            pb.build_implicit_gemm()
            ready = side.record_event()
        main.wait_event(ready)
        pb._map_fwd().record_stream(main)
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
    # 20 untimed steps: the caching allocator's pools keep growing over the first steps
    # (map B, allocated on the second stream, is reusable only after the main stream has
    # run conv B), and a step that grows them spends milliseconds in hipMalloc
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only loop the dominant kernel (for profiler runs)")
    ap.add_argument("--workload", default="precision",
                    choices=["precision", "single", "diff", "unet", "serve", "orders", "rows"],
                    help="precision (default, the headline line); single / diff / unet / serve: the other "
                         "workloads, one JSON line each; orders: contraction per layer in tree order vs "
                         "tap-mask order per window; rows: grid build, pooling, upsampling, injection, "
                         "GroupNorm, queries, dense I/O and save / load on the sphere batch")
    ap.add_argument("--requests", type=int, default=120, help="serve: number of requests")
    ap.add_argument("--serve-requests", type=int, default=40,
                    help="requests of the short serving stream reported in the default line (0: none)")
    ap.add_argument("--rate", type=float, default=0.0, help="serve: Poisson arrival rate (req/s); 0 = auto")
    ap.add_argument("--serve-ingest-workers", type=int, default=1, help="serve: ingest host threads (0 = inline)")
    ap.add_argument("--serve-streams", type=int, default=3, help="serve: compute streams (= batches in flight)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    import torch

    import fvdb

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.workload != "precision":
        rec = {"single": run_single, "diff": run_diff, "unet": run_unet, "serve": run_serve,
               "orders": run_orders, "rows": run_rows}[args.workload](
            torch, fvdb, dev, args)
        print(json.dumps(rec), flush=True)
        return 0
    log("[bench] building workload")
    grid, vs, x, wa, wb = build_workload(torch, fvdb, dev)
    N = grid.total_voxels
    counts = grid.num_voxels.tolist()
    log(f"[bench] {grid.grid_count} grids, {N} voxels ({counts})")

    xb, wab, wbb = x.bfloat16(), wa.bfloat16(), wb.bfloat16()
    pa, pb, ya, yb = run_step(fvdb, grid, xb, wab, wbb)
    # Park the objects of the imports and the set-up in the collector's permanent
    # generation: otherwise the first full collection (some 20 steps in) walks all of
    # them and stalls one step by ~40 ms on the host.
    gc.collect()
    gc.freeze()
    P_a = int((pa.out_in_map >= 0).sum())
    P_b = int((pb.out_in_map >= 0).sum())
    N_b = pb.target_grid.total_voxels
    flops = 2.0 * P_a * 64 * 64 + 2.0 * P_b * 64 * 128
    flops_a = 2.0 * P_a * 64 * 64

    # dominant kernel: conv A implicit GEMM (bf16), maps prebuilt
    kio = fvdb._autograd.weight_to_koi(wab, False)
    fmap = pa._map_fwd()
    conv_a = lambda: fvdb.ext.implicit_gemm(xb, kio, fmap, N)  # noqa: E731
    if args.kernel_only:
        # a streaming elementwise kernel of known size first (reads and writes COPY_BYTES
        # once; a same-dtype copy_ may run on a DMA engine, which the counters do not
        # see): its FETCH_SIZE / WRITE_SIZE readings calibrate the conv kernel's
        src = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        for i in range(args.steps):
            conv_a()
            torch.cuda.synchronize()
            log(f"[bench] kernel iter {i}")
        return 0

    side = torch.cuda.Stream(device=dev)  # default priority
    torch.cuda.synchronize()
    log("[bench] timing bf16 step")
    t_bf16 = time_steps(torch, lambda: run_step(fvdb, grid, xb, wab, wbb, side), args.steps, args.warmup)
    t_bf16_serial = time_steps(torch, lambda: run_step(fvdb, grid, xb, wab, wbb), args.steps, args.warmup)
    log("[bench] timing fp32 step")
    t_fp32 = time_steps(torch, lambda: run_step(fvdb, grid, x, wa, wb, side), args.steps, args.warmup)

    # conv-only (maps built once), per dtype
    def conv_only(dt):
        xa, w1, w2 = x.to(dt), wa.to(dt), wb.to(dt)
        IG = fvdb.ConvPackBackend.IGEMM
        xj = grid.jagged_like(xa)
        return lambda: pb.sparse_conv_3d(pa.sparse_conv_3d(xj, w1, IG), w2, IG)

    log("[bench] timing conv-only")
    # median of per-iteration HIP-event times on the launch stream (host wall clock
    # around a few launches is noisy at this scale)
    tc = {name: event_avg_ms(torch, conv_only(dt), max(10, args.steps))[0] * 1e-3
          for name, dt in (("bf16", torch.bfloat16), ("fp16", torch.float16), ("fp32", torch.float32))}

    log("[bench] timing dominant kernel with HIP events")
    k_ms, _ = event_avg_ms(torch, conv_a, max(10, args.steps))
    kio32 = fvdb._autograd.weight_to_koi(wa, False)
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
    roof["kernel"] = ("k_conv_tile_dma<bf16, LX=3, NW=8, KT=27, MI=4, NCC=1, WSKIP> (conv A, 64->64, 3x3x3: 256-row "
                      "workgroups, weights staged global->LDS, gathers 3 taps ahead, 4 waves/SIMD, wavefronts skip "
                      "the MFMAs of taps none of their rows has)")
    roof["avg_launch_ms"] = k_ms
    roof["algorithmic_flops"] = flops_a
    roof["algorithmic_bytes"] = bytes_a

    errors, sample_n = ({}, 0) if args.no_check else check_errors(torch, fvdb, grid, x, wa, wb, 0)
    cpu = None if args.no_cpu_baseline else cpu_baseline(torch, fvdb, grid, x, wa, wb)
    serve = None
    if args.serve_requests > 0:
        # the metric's serving half: a short Poisson stream through the U-Net workload's network
        log(f"[bench] serving stream of {args.serve_requests} requests")
        sv = serve_stream(torch, fvdb, dev, args.serve_requests, ingest_workers=args.serve_ingest_workers,
                          streams=args.serve_streams)
        cb = sv["bf16_continuous_batching"]
        serve = {"workload": f"{args.serve_requests} single-grid requests (log-uniform 10k-2M voxels, Poisson "
                             f"arrivals at 70% of the closed-batch service rate), 4-level U-Net 32/64/128/256",
                 "p50_ms": cb["p50_ms"], "p99_ms": cb["p99_ms"], "throughput_voxels_per_s": cb["throughput_voxels_per_s"],
                 "rate_req_per_s": cb["rate_req_per_s"], "mean_batch": cb["mean_batch"],
                 "hbm_budget_GB": cb["hbm_budget_GB"], "measured_peak_allocated_GB": cb["measured_peak_allocated_GB"],
                 "admission_steps_with_requests_waiting": cb["admission_steps_with_requests_waiting"],
                 "bf16_one_at_a_time_p50_p99_ms": [sv["bf16_one_at_a_time"]["p50_ms"], sv["bf16_one_at_a_time"]["p99_ms"]],
                 "fp32_one_at_a_time_p50_p99_ms": [sv["fp32_one_at_a_time"]["p50_ms"], sv["fp32_one_at_a_time"]["p99_ms"]]}

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
                               "then 2x2x2 s2 64->128, kernel maps + target grid rebuilt every step "
                               "(target grid and map B built on a second stream while conv A runs)",
                   "grids": grid.grid_count, "voxels": N, "voxels_per_grid": counts, "voxels_convB_out": N_b,
                   "pairs_convA": P_a, "pairs_convB": P_b, "voxel_sizes": vs,
                   "coordinates": "every voxel in the root tile [0, 4096)^3 (origin -0.3 - voxel size)"},
        "tflops_step_bf16": flops / t_bf16 / 1e12,
        "step_ms_single_stream": t_bf16_serial * 1e3,
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
        "serve": serve,
    }
    print(json.dumps(rec), flush=True)
    return 0


# ---------------------------------------------------------------------------------------------
# other workloads (one JSON line each; the default line above is the headline metric)
# ---------------------------------------------------------------------------------------------


def _median_ms(torch, fn, iters=10, warmup=2):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(iters):
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        e.synchronize()
        ts.append(s.elapsed_time(e))
    return statistics.median(ts)


def run_single(torch, fvdb, dev, args):
    """One grid: 2M points on 8 spheres at 1/256, 3x3x3 32->32 fp32; kernel map + forward."""
    import numpy as np

    import oracle
    from fvdb.utils.synthetic import sphere_points, voxelize

    pts = sphere_points(seed=0, n_points=2_000_000, device=dev)
    grid = voxelize(pts, 1 / 256)
    N = grid.total_voxels
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.randn((N, 32), generator=g, device=dev)
    w = torch.randn((32, 32, 3, 3, 3), generator=g, device=dev) * math.sqrt(2.0 / (32 * 27))
    IG = fvdb.ConvPackBackend.IGEMM

    def kmap():
        p, _ = grid.sparse_conv_kernel_map(3, 1)
        p.build_implicit_gemm()
        return p

    pack = kmap()
    t_kmap = _median_ms(torch, lambda: fvdb.ext.kernel_map(grid._impl.tensors(), grid._impl.tensors(), 1, [3, 3, 3],
                                                             [1, 1, 1], False, N), 20)
    res = {}
    for name, dt in (("fp32", torch.float32), ("bf16", torch.bfloat16)):
        xd, wd = x.to(dt), w.to(dt)
        t_fwd = _median_ms(torch, lambda: pack.sparse_conv_3d(grid.jagged_like(xd), wd, IG), 20)
        t_step = _median_ms(torch, lambda: kmap().sparse_conv_3d(grid.jagged_like(xd), wd, IG), 10)
        res[name] = {"forward_ms": t_fwd, "step_ms": t_step, "voxels_per_s": N / (t_step * 1e-3)}
    P = int((pack.out_in_map >= 0).sum())
    ijk = grid.ijk.jdata.cpu().numpy()
    km = oracle.kernel_map([ijk], [ijk], 3, 1)
    bit_exact = bool(np.array_equal(pack.out_in_map[:N].cpu().numpy(), km))
    ref = oracle.sparse_conv(x.double().cpu(), w.double().cpu(), km, N, 1)
    out = pack.sparse_conv_3d(grid.jagged_like(x), w, IG).jdata
    err = float((out.double().cpu() - ref).abs().max() / ref.abs().max())
    kmap_bytes = N * (12 + 4 * 27) + 80 * grid.total_leaf_nodes
    t0 = time.perf_counter()
    oracle.kernel_map([ijk], [ijk], 3, 1)
    t_cpu_kmap = time.perf_counter() - t0
    return {"workload": "single: 1 grid, 2M points on 8 spheres at 1/256, 3x3x3 32->32", "voxels": N, "pairs": P,
            "kmap_ms": t_kmap, "kmap_GBps": kmap_bytes / (t_kmap * 1e-3) / 1e9, "kmap_bit_exact": bit_exact,
            "fp32_max_rel_err_vs_fp64": err, "results": res,
            "cpu_kmap_ms_1thread": t_cpu_kmap * 1e3}


def run_diff(torch, fvdb, dev, args):
    """~1M voxels (planes + spheres at 1/1024), 3x3x3 128->128 bf16 forward + dgrad + wgrad."""
    from fvdb.utils.synthetic import planes_spheres_points, voxelize

    pts = planes_spheres_points(seed=7, n_points=2_400_000, n_planes=2, n_spheres=2, device=dev)
    grid = voxelize(pts, 1 / 1024)
    N = grid.total_voxels
    pack, _ = grid.sparse_conv_kernel_map(3, 1)
    pack.build_implicit_gemm(training=True)
    pack.build_gather_scatter()
    P = int(pack.neighborhood_sizes.sum())
    F = 2.0 * P * 128 * 128
    g = torch.Generator(device=dev).manual_seed(3)
    x64 = torch.randn((N, 128), generator=g, device=dev, dtype=torch.float64)
    w64 = torch.randn((128, 128, 3, 3, 3), generator=g, device=dev, dtype=torch.float64) * math.sqrt(2.0 / (128 * 27))
    R64 = torch.randn((N, 128), generator=g, device=dev, dtype=torch.float64)
    IG = fvdb.ConvPackBackend.IGEMM
    ext = fvdb.ext
    from fvdb._autograd import weight_to_kio, wgrad_chunk

    out = {}
    grads = {}
    pairs, sizes = pack._pairs()
    for name, dt in (("bf16", torch.bfloat16), ("fp32", torch.float32)):
        x, w, R = x64.to(dt), w64.to(dt), R64.to(dt)
        kio = weight_to_kio(w, False)
        a_f = kio.transpose(1, 2).contiguous()
        # row orders as the autograd path picks them (tap-mask order per window, or tree order)
        of, ob = pack._order_for(True, 128, dt, 128), pack._order_for(False, 128, dt, 128)
        fwd = lambda: ext.implicit_gemm(x, a_f, pack._map_fwd(), N, of)  # noqa: E731
        dgr = lambda: ext.implicit_gemm(R, kio, pack._map_bwd(), N, ob)  # noqa: E731
        chunk = wgrad_chunk(P, dt, 128, 128)
        koff, cstart, total = pack._wgrad_tables(chunk)
        wgr = lambda: ext.wgrad_tabled(x, R, pairs, koff, cstart, total, False, chunk)  # noqa: E731
        t = {"fwd_ms": _median_ms(torch, fwd, 10), "dgrad_ms": _median_ms(torch, dgr, 10),
             "wgrad_ms": _median_ms(torch, wgr, 10)}
        t["tflops"] = {k[:-3]: F / (v * 1e-3) / 1e12 for k, v in list(t.items())}
        t["row_order"] = "tap mask per 8192-row window" if of is not None else "tree (direct-to-LDS kernel)"
        # autograd end to end: loss = sum(out * R)
        xr = x.clone().requires_grad_(True)
        wr = w.clone().requires_grad_(True)

        def step():
            y = pack.sparse_conv_3d(grid.jagged_like(xr), wr, IG).jdata
            # loss = sum(out * R) accumulated in fp32; its gradient with respect to out is R
            # exactly, with no fp32 copies of the two [N, 128] arrays
            (y * R).sum(dtype=torch.float32).backward()

        t["train_step_ms"] = _median_ms(torch, step, 10)
        xr.grad = None
        wr.grad = None
        step()
        grads[name] = (fwd(), xr.grad.detach().clone(), wr.grad.detach().clone())
        out[name] = t
        log(f"[bench] diff {name}: {t}")

    # float64 CPU oracle (the reference gather-scatter restated) on the same rounded inputs
    import oracle

    km = pack.out_in_map[:N].cpu().numpy()
    errs = {}
    for name, dt in (("bf16", torch.bfloat16), ("fp32", torch.float32)):
        xc, wc, Rc = x64.to(dt).double().cpu(), w64.to(dt).double().cpu(), R64.to(dt).double().cpu()
        ref_out = oracle.sparse_conv(xc, wc, km, N, 1)
        gi, gw = oracle.sparse_conv_backward(xc, Rc, wc, km, (3, 3, 3))

        def rel(a, b):
            return float((a.double().cpu() - b).abs().max() / b.abs().max())

        errs[name] = {"out": rel(grads[name][0], ref_out), "dgrad": rel(grads[name][1], gi),
                      "wgrad": rel(grads[name][2], gw)}
        log(f"[bench] diff errors {name}: {errs[name]}")
    # finite differences in fp64 on a ~4k-voxel subgrid, C = 8
    ijk = grid.ijk.jdata
    lo = ijk[0]
    sel = ((ijk - lo).abs() < 40).all(1)
    sub = fvdb.GridBatch.from_ijk(ijk[sel][:4096].contiguous())
    sp, _ = sub.sparse_conv_kernel_map(3, 1)
    sp.build_implicit_gemm(training=True)
    n = sub.total_voxels
    gx = torch.Generator(device=dev).manual_seed(11)
    xs = torch.randn((n, 8), generator=gx, device=dev, dtype=torch.float64, requires_grad=True)
    ws = (0.2 * torch.randn((8, 8, 3, 3, 3), generator=gx, device=dev, dtype=torch.float64)).requires_grad_(True)
    Rs = torch.randn((n, 8), generator=gx, device=dev, dtype=torch.float64)

    def loss(a, b):
        return (sp.sparse_conv_3d(sub.jagged_like(a), b, IG).jdata * Rs).sum()

    loss(xs, ws).backward()
    h = 1e-6
    fd_err = 0.0
    for idx in [(0, 0), (n // 3, 5), (n - 1, 7)]:
        xp, xm = xs.detach().clone(), xs.detach().clone()
        xp[idx] += h
        xm[idx] -= h
        fd = (loss(xp, ws.detach()) - loss(xm, ws.detach())).item() / (2 * h)
        fd_err = max(fd_err, abs(fd - xs.grad[idx].item()) / max(1.0, abs(fd)))
    for idx in [(0, 0, 1, 1, 1), (3, 5, 0, 2, 1), (7, 7, 2, 2, 2)]:
        wp, wm = ws.detach().clone(), ws.detach().clone()
        wp[idx] += h
        wm[idx] -= h
        fd = (loss(xs.detach(), wp) - loss(xs.detach(), wm)).item() / (2 * h)
        fd_err = max(fd_err, abs(fd - ws.grad[idx].item()) / max(1.0, abs(fd)))
    return {"workload": "diff: planes+spheres at 1/1024, 3x3x3 128->128, forward + dgrad + wgrad", "voxels": N,
            "pairs": P, "flops_per_pass": F, "timings": out, "errors_vs_fp64_cpu_oracle": errs,
            "fd_check": {"subgrid_voxels": n, "channels": 8, "h": h, "max_rel_err": fd_err, "pass": fd_err < 1e-6}}


def run_rows(torch, fvdb, dev, args):
    """The operations either side of the contraction on the sphere batch (2.34M voxels, 64 bf16
    channels): grid build, 2x coarsening, max / avg pooling, nearest upsampling, injection,
    GroupNorm, voxel queries, dense read / write (one smaller grid) and save / load.  Each is the
    median of single calls through the public API (HIP events around the call, host work inside it
    included), set against its algorithmic HBM bytes (the minimum its inputs and outputs imply)."""
    import os
    import tempfile

    from fvdb.utils.synthetic import sphere_points, voxelize

    grid, vs, x, wa, wb = build_workload(torch, fvdb, dev)
    del wa, wb
    N, C, b = grid.total_voxels, 64, 2
    xb = x.to(torch.bfloat16)
    del x
    xj = grid.jagged_like(xb)
    res = {}

    def grid_bytes(g):
        return sum(t.numel() * t.element_size() for t in g._impl.tensors() if torch.is_tensor(t))

    def put(name, ms, nbytes, rows):
        gbps = nbytes / (ms * 1e-3) / 1e9
        res[name] = {"ms": ms, "algorithmic_bytes": int(nbytes), "GBps": gbps, "hbm_frac": gbps / 8000.0,
                     "rows_per_s": rows / (ms * 1e-3)}
        log(f"[bench] rows {name}: {res[name]}")

    with torch.no_grad():
        ijk = grid.ijk
        t = _median_ms(torch, lambda: fvdb.GridBatch.from_ijk(ijk, voxel_sizes=grid.voxel_sizes,
                                                                origins=grid.origins), 10)
        put("grid_from_ijk", t, N * 12 + grid_bytes(grid), N)
        coarse = grid.coarsened_grid(2)
        Nc = coarse.total_voxels
        L = int(grid._impl.keys.shape[0])
        t = _median_ms(torch, lambda: grid.coarsened_grid(2), 10)
        put("coarsened_grid_2x", t, L * (8 + 64) + grid_bytes(coarse), Nc)
        pool_bytes = N * C * b + Nc * C * b + 2 * Nc * 8 * 4  # rows in, windows out, window map written + read
        t = _median_ms(torch, lambda: grid.max_pool(2, xj, coarse_grid=coarse), 10)
        put("max_pool_2x", t, pool_bytes, Nc)
        t = _median_ms(torch, lambda: grid.avg_pool(2, xj, coarse_grid=coarse), 10)
        put("avg_pool_2x", t, pool_bytes, Nc)
        cj = coarse.jagged_like(torch.randn(Nc, C, device=dev).to(torch.bfloat16))
        t = _median_ms(torch, lambda: coarse.subdivide(2, cj, fine_grid=grid), 10)
        put("upsample_nearest_2x", t, Nc * C * b + N * C * b + N * (12 + 4), N)
        t = _median_ms(torch, lambda: grid.inject_from(grid, xj), 10)
        put("inject_from", t, 3 * N * C * b + N * (12 + 4), N)
        from fvdb._norm import group_norm

        gw = torch.ones(C, device=dev, dtype=torch.bfloat16)
        gb = torch.zeros(C, device=dev, dtype=torch.bfloat16)
        t = _median_ms(torch, lambda: group_norm(xb, grid, 32, gw, gb), 10)
        put("group_norm_32", t, 3 * N * C * b, N)
        xg = xb.detach().requires_grad_(True)
        dyg = torch.randn_like(xb)

        def gn_fwd_bwd():
            xg.grad = None
            with torch.enable_grad():
                group_norm(xg, grid, 32, gw, gb).backward(dyg)

        t = _median_ms(torch, gn_fwd_bwd, 10)
        # forward 3 passes; backward reads x and dy twice and writes dx
        put("group_norm_32_fwd_bwd", t, 8 * N * C * b, N)
        del xg, dyg
        # what torch's own streaming kernels reach over the same array: a copy (read + write)
        # and a full sum (read only), the practical ceilings for the rows above
        yb = torch.empty_like(xb)
        t = _median_ms(torch, lambda: yb.copy_(xb), 10)
        put("torch_copy_reference", t, 2 * N * C * b, N)
        t = _median_ms(torch, lambda: xb.sum(dtype=torch.float32), 10)
        put("torch_sum_reference", t, N * C * b, N)
        del yb
        t = _median_ms(torch, lambda: grid.ijk_to_index(ijk), 10)
        put("ijk_to_index", t, N * (12 + 8), N)
        # dense I/O on one grid whose box stays small (8 channels)
        gd = voxelize(sphere_points(seed=5, n_points=500_000, device=dev), 1 / 128)
        Nd, Cd = gd.total_voxels, 8
        xd = gd.jagged_like(torch.randn(Nd, Cd, device=dev).to(torch.bfloat16))
        mc = gd.ijk.jdata.min(dim=0).values
        dense = gd.write_to_dense_czyx(xd, min_coord=mc)
        t = _median_ms(torch, lambda: gd.write_to_dense_czyx(xd, min_coord=mc), 10)
        put("write_to_dense_czyx", t, dense.numel() * b + 2 * Nd * Cd * b + Nd * 12, Nd)
        back = gd.read_from_dense_czyx(dense, dense_origins=mc)
        res["dense_round_trip_equal"] = bool(torch.equal(back.jdata, xd.jdata))
        t = _median_ms(torch, lambda: gd.read_from_dense_czyx(dense, dense_origins=mc), 10)
        put("read_from_dense_czyx", t, 2 * Nd * Cd * b + Nd * 12, Nd)
        res["dense_box"] = list(dense.shape)
        del dense
        # save / load through the host (file in the temporary directory)
        fd, path = tempfile.mkstemp(suffix=".npz")
        os.close(fd)
        try:
            t0 = time.perf_counter()
            fvdb.save_gridbatch(path, grid, data=xj)
            t_save = time.perf_counter() - t0
            size = os.path.getsize(path)
            t0 = time.perf_counter()
            g2, d2, _ = fvdb.load_gridbatch(path, device=dev)
            torch.cuda.synchronize()
            t_load = time.perf_counter() - t0
            same = bool(torch.equal(g2.ijk.jdata, grid.ijk.jdata) and torch.equal(d2.jdata, xb))
        finally:
            os.remove(path)
        res["save_load"] = {"file_bytes": size, "save_ms": t_save * 1e3, "load_ms": t_load * 1e3,
                            "save_MBps": size / t_save / 1e6, "load_MBps": size / t_load / 1e6, "round_trip_equal": same}
        log(f"[bench] rows save_load: {res['save_load']}")
    return {"workload": "rows: operations either side of the contraction on the 8-grid sphere batch (64 bf16 channels)",
            "voxels": N, "coarse_voxels": Nc, "leaves": L, "dense_grid_voxels": Nd, "results": res}


def run_orders(torch, fvdb, dev, args):
    """Contraction time per layer shape in tree order and in tap-mask order per 8192-row window (what packs
    pick by themselves where ``wants_row_order`` says so), bf16 and fp32, on the precision, differentiation
    and U-Net grids; outputs are checked bit-identical between the two orders."""
    from fvdb.sparse_conv_pack_info import AUTO_ORDER_WINDOW, wants_row_order
    from fvdb.utils.synthetic import planes_spheres_points, voxelize

    ext = fvdb.ext
    rows = []

    def case(name, m, n_rows, n_in, cin, cout):
        K = m.shape[1]
        g = torch.Generator(device=dev).manual_seed(3)
        x64 = torch.randn((n_in, cin), generator=g, device=dev, dtype=torch.float64)
        a64 = torch.randn((K, cout, cin), generator=g, device=dev, dtype=torch.float64) / math.sqrt(cin * K)
        win = AUTO_ORDER_WINDOW
        order = ext.row_order(m, n_rows, win)
        rec = {"layer": name, "rows": n_rows, "taps": K, "window": win,
               "order_ms": _median_ms(torch, lambda: ext.row_order(m, n_rows, win), 10)}
        for dn, dt in (("bf16", torch.bfloat16), ("fp32", torch.float32)):
            x, a = x64.to(dt), a64.to(dt)
            ref = ext.implicit_gemm(x, a, m, n_rows)
            if not torch.equal(ext.implicit_gemm(x, a, m, n_rows, order), ref):
                raise RuntimeError(f"{name} {dn}: row order changed the result")
            rec[dn] = {"tree_ms": _median_ms(torch, lambda: ext.implicit_gemm(x, a, m, n_rows), 10),
                       "windowed_ms": _median_ms(torch, lambda: ext.implicit_gemm(x, a, m, n_rows, order), 10),
                       "picked": "windowed" if wants_row_order(K, cin, dt, cout, n_rows) else "tree"}
        log(f"[bench] orders {rec}")
        rows.append(rec)

    grid, _, _, _, _ = build_workload(torch, fvdb, dev)
    pa, _ = grid.sparse_conv_kernel_map(3, 1)
    case("sphere batch 3x3x3 64->64", pa._map_fwd(), grid.total_voxels, grid.total_voxels, 64, 64)
    del grid, pa
    pts = planes_spheres_points(seed=7, n_points=2_400_000, n_planes=2, n_spheres=2, device=dev)
    gd = voxelize(pts, 1 / 1024)
    pd, _ = gd.sparse_conv_kernel_map(3, 1)
    case("planes+spheres 3x3x3 128->128", pd._map_fwd(), gd.total_voxels, gd.total_voxels, 128, 128)
    del gd, pd
    gl = _indoor_batch(torch, fvdb, dev, [1, 2, 3, 4])
    chans = (32, 64, 128, 256)
    for l, c in enumerate(chans):
        n = gl.total_voxels
        pk, _ = gl.sparse_conv_kernel_map(3, 1)
        if l == 0 or l == 3:
            case(f"U-Net level {l} 3x3x3 {c}->{c}", pk._map_fwd(), n, n, c, c)
        else:
            case(f"U-Net level {l} decoder 3x3x3 {2 * c}->{c}", pk._map_fwd(), n, n, 2 * c, c)
        if l + 1 < len(chans):
            pdn, gc = gl.sparse_conv_kernel_map(2, 2)
            case(f"U-Net level {l} up 2x2x2 s2 transposed {chans[l + 1]}->{c}", pdn._map_bwd(), n, gc.total_voxels,
                 chans[l + 1], c)
            gl = gc
    return {"workload": "orders: contraction per layer, tree order vs tap-mask order per 8192-row window",
            "layers": rows}


def _indoor_batch(torch, fvdb, dev, seeds):
    from fvdb.utils.synthetic import indoor_ijk

    parts = [indoor_ijk(s, device=dev) for s in seeds]
    return fvdb.GridBatch.from_ijk(fvdb.JaggedTensor(parts), voxel_sizes=0.02)


def run_unet(torch, fvdb, dev, args):
    """4-level sparse U-Net (32/64/128/256), batch of 4 indoor grids, bf16 vs fp32 forward."""
    grid = _indoor_batch(torch, fvdb, dev, [1, 2, 3, 4])
    N = grid.total_voxels
    torch.manual_seed(0)
    model = fvdb.nn.SparseUNet(in_channels=32, channels=(32, 64, 128, 256)).to(dev).eval()
    g = torch.Generator(device=dev).manual_seed(5)
    x = torch.randn((N, 32), generator=g, device=dev)
    res = {}
    with torch.no_grad():
        for name, dt in (("bf16", torch.bfloat16), ("fp32", torch.float32)):
            m = model.to(dt)
            inp = fvdb.nn.VDBTensor(grid, grid.jagged_like(x.to(dt)))
            t = _median_ms(torch, lambda: m(inp), 8, 1)
            res[name] = {"forward_ms": t, "voxels_per_s": N / (t * 1e-3)}
            log(f"[bench] unet {name}: {res[name]}")
        # accuracy against the float64 network (scalar fp64 kernels) on a small room
        from fvdb.utils.synthetic import indoor_ijk

        small = fvdb.GridBatch.from_ijk(indoor_ijk(9, n_boxes=6, room=(90, 90, 40), device=dev), voxel_sizes=0.02)
        xs = torch.randn((small.total_voxels, 32), generator=g, device=dev)
        outs = {}
        for name, dt in (("bf16", torch.bfloat16), ("fp32", torch.float32), ("fp64", torch.float64)):
            m = model.to(dt)
            outs[name] = m(fvdb.nn.VDBTensor(small, small.jagged_like(xs.to(dt)))).data.jdata
        model.float()

    def rel(a, b):
        return float((a.double() - b.double()).abs().max() / b.double().abs().max())

    return {"workload": "unet: 4 levels 32/64/128/256, batch of 4 indoor grids at 2 cm", "voxels": N,
            "voxels_per_grid": grid.num_voxels.tolist(), "results": res,
            "bf16_over_fp32": res["fp32"]["forward_ms"] / res["bf16"]["forward_ms"],
            "errors_vs_fp64": {"sample_voxels": small.total_voxels, "bf16": rel(outs["bf16"], outs["fp64"]),
                               "fp32": rel(outs["fp32"], outs["fp64"])}}


def serve_stream(torch, fvdb, dev, nreq, rate=0.0, seed=42, budget_factor=2.0, ingest_workers=1, streams=3):
    """Poisson stream of single grids (log-uniform 10k-2M voxels) through the 4-level U-Net in bf16
    with continuous batching, against one-at-a-time loops (bf16 and fp32) over the same arrivals.
    Requests are ingested on `ingest_workers` host threads (0: inline in the loop) and batches run
    on `streams` compute streams (at most that many batches in flight).

    The HBM budget is `budget_factor` times the largest request's admission bound, so admission
    binds; the device memory the stream allocates beyond what was live before it is measured
    (peak allocated after ``reset_peak_memory_stats``) and reported beside the budget.  Results
    are copied to host memory as they complete.  `rate` 0 offers 70% of the engine's service
    rate measured on a closed batch of the same requests."""
    from fvdb.serving import ContinuousBatcher, Request, UNetExecutor, latency_stats
    from fvdb.utils.synthetic import log_uniform, sphere_points_for_voxels, voxelize

    sizes = [int(v) for v in log_uniform(seed, 10_000, 2_000_000, nreq)]
    reqs_ijk = []
    for i, nv in enumerate(sizes):
        pts, vs = sphere_points_for_voxels(1000 + i, nv, device=dev)
        g = voxelize(pts, vs)
        reqs_ijk.append(g.ijk.jdata.cpu().pin_memory())
        del pts, g
    counts = [t.shape[0] for t in reqs_ijk]
    torch.manual_seed(0)
    model = fvdb.nn.SparseUNet(in_channels=32, channels=(32, 64, 128, 256)).to(dev).eval()

    def one_at_a_time(arrivals, dtype):
        m = model.to(dtype)
        t0 = time.perf_counter()
        lat = []
        with torch.no_grad():
            for i, tarr in enumerate(arrivals):
                while time.perf_counter() - t0 < tarr:
                    pass
                grid = fvdb.GridBatch.from_ijk(reqs_ijk[i].to(dev, non_blocking=True))
                gen = torch.Generator(device=dev).manual_seed(100 + i)
                xx = torch.randn((grid.total_voxels, 32), generator=gen, device=dev).to(dtype)
                y = m(fvdb.nn.VDBTensor(grid, grid.jagged_like(xx))).data.jdata
                torch.empty(y.shape, dtype=y.dtype, pin_memory=True).copy_(y)  # result to host, as the server
                lat.append(time.perf_counter() - t0 - tarr)
        model.float()
        lat.sort()
        span = time.perf_counter() - t0
        k99 = min(len(lat) - 1, max(0, math.ceil(0.99 * len(lat)) - 1))
        return {"p50_ms": 1e3 * lat[(len(lat) - 1) // 2], "p99_ms": 1e3 * lat[k99],
                "throughput_voxels_per_s": sum(counts) / span, "throughput_req_per_s": nreq / span, "span_s": span}

    ex = UNetExecutor(model, dtype=torch.bfloat16, host_results=True, compute_streams=streams)
    budget = int(budget_factor * max(ex.estimate_bytes(Request(rid=i, ijk=t)) for i, t in enumerate(reqs_ijk)))
    warm = ContinuousBatcher(ex, hbm_budget_bytes=budget, max_inflight=streams, ingest_workers=ingest_workers)
    warm.serve([Request(rid=i, ijk=reqs_ijk[i], seed=100 + i) for i in range(min(8, nreq))])
    # the set-up's objects out of the cyclic collector's full scans
    gc.collect()
    gc.freeze()
    t0 = time.perf_counter()
    closed = ContinuousBatcher(ex, hbm_budget_bytes=budget, max_inflight=streams, ingest_workers=ingest_workers)
    closed.serve([Request(rid=i, ijk=reqs_ijk[i], seed=100 + i) for i in range(nreq)])
    torch.cuda.synchronize()
    service_rate = nreq / (time.perf_counter() - t0)
    closed_batch = statistics.mean(closed.batches)
    rate = rate if rate > 0 else 0.7 * service_rate
    gexp = torch.Generator().manual_seed(7)
    gaps = torch.empty(nreq, dtype=torch.float64).exponential_(rate, generator=gexp)
    arrivals = torch.cumsum(gaps, 0).tolist()
    batcher = ContinuousBatcher(ex, hbm_budget_bytes=budget, max_batch_voxels=6_000_000, max_inflight=streams,
                                ingest_workers=ingest_workers)
    waits = [0]
    admit = batcher._admit

    def counted_admit():
        admit()
        waits[0] += 1 if batcher.queue else 0
    batcher._admit = counted_admit
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated(dev)
    torch.cuda.reset_peak_memory_stats(dev)
    t_start = time.perf_counter()
    done = batcher.serve([Request(rid=i, ijk=reqs_ijk[i], seed=100 + i, arrival=arrivals[i]) for i in range(nreq)])
    torch.cuda.synchronize()
    span = time.perf_counter() - t_start
    peak = torch.cuda.max_memory_allocated(dev) - base
    ok = [r for r in done if r.status == "done"]
    stats = latency_stats(done)
    slow = sorted(ok, key=lambda r: -r.latency)[:3]
    for r in sorted(ok, key=lambda r: r.arrival):
        log(f"[bench] req #{r.rid:3d} {counts[r.rid]:8d} vox  arr {r.arrival * 1e3:7.2f}  admit +{(r.t_admit - r.arrival) * 1e3:6.2f}"
            f"  prep +{(r.t_prepared - r.arrival) * 1e3:6.2f}  launch +{(r.t_launch - r.arrival) * 1e3:6.2f}"
            f"  done +{r.latency * 1e3:6.2f} ms")
    log("[bench] slowest requests: " + ", ".join(f"#{r.rid} {counts[r.rid]} voxels arrived {r.arrival * 1e3:.1f} ms "
                                                 f"admitted +{(r.t_admit - r.arrival) * 1e3:.1f} latency "
                                                 f"{r.latency * 1e3:.1f} ms" for r in slow))
    serve = {"requests": nreq, "completed": len(ok), "rate_req_per_s": rate, "service_rate_closed": service_rate,
             "mean_batch_closed": closed_batch,
             **stats, "throughput_voxels_per_s": sum(counts[r.rid] for r in ok) / span,
             "throughput_req_per_s": len(ok) / span, "span_s": span,
             "mean_batch": statistics.mean(batcher.batches), "max_batch": max(batcher.batches),
             "hbm_budget_GB": budget / 1e9, "peak_reserved_by_admission_GB": batcher.peak_reserved / 1e9,
             "measured_peak_allocated_GB": peak / 1e9, "measured_peak_within_budget": bool(peak <= budget),
             "admission_steps_with_requests_waiting": waits[0]}
    for bt in (warm, closed, batcher):
        bt.close()
    serve["ingest_workers"], serve["compute_streams"] = ingest_workers, streams
    log(f"[bench] serve bf16 continuous batching: {serve}")
    base_bf16 = one_at_a_time(arrivals, torch.bfloat16)
    log(f"[bench] serve bf16 one at a time: {base_bf16}")
    base_fp32 = one_at_a_time(arrivals, torch.float32)
    log(f"[bench] serve fp32 one at a time: {base_fp32}")
    return {"voxels_total": sum(counts), "voxels_min": min(counts), "voxels_max": max(counts),
            "bf16_continuous_batching": serve, "bf16_one_at_a_time": base_bf16, "fp32_one_at_a_time": base_fp32}


def run_serve(torch, fvdb, dev, args):
    """The serving workload: a Poisson stream of `--requests` single grids (default 120)."""
    rec = serve_stream(torch, fvdb, dev, args.requests, args.rate, ingest_workers=args.serve_ingest_workers,
                       streams=args.serve_streams)
    return {"workload": "serve: Poisson stream, single grids log-uniform 10k-2M voxels, 4-level U-Net", **rec}


if __name__ == "__main__":
    sys.exit(main())
