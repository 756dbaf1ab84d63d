"""Continuous-batching inference server for sparse-conv networks on one GPU.

Requests are single grids (voxel coordinates plus features) that arrive over time.

* **Ingest** (a separate HIP stream, optionally on a host thread of its own): each admitted request's grid index
  and all the kernel maps its network needs (one 3x3x3 map per level, the stride-2 down map and its transpose
  between levels) are built once with the gfx950 kernels and stay resident on the device with the request. A
  grid build reads two sizes back to the host; on the ingest thread that wait does not hold up the loop.
* **Admission**: a request is admitted (FIFO) when an upper bound of its device memory (maps, features,
  activations, output, ingest buffers) fits in the remaining HBM budget. Before ingest only the input voxel count
  is known, so the bound takes every level as large as the input; once the request is ingested its level sizes
  are known and the reservation drops to the bound for those sizes (``exact_bytes``). Reservations are released
  when the request's batch completes or fails.
* **Batching**: ready requests are packed (FIFO, up to a voxel cap) into one jagged launch per layer. The
  batch's kernel maps are the per-request maps concatenated with their indices shifted by the request's row
  offset at that level, so no grid is rebuilt per batch. Each batch goes to a compute stream whose previous
  batch is done (one stream per batch in flight), so a small batch runs beside a large one instead of behind
  it, while the next requests are ingested on the ingest stream.
* **Results**: every output row of the contraction is computed by one lane in a fixed order, independent of
  the other rows of the launch (see ``fvdbk_conv_implicit_gemm`` in ``include/fvdbk.h``), and bias/ReLU/concat
  are per row. A request's output is therefore bit-identical alone and in any mix.

The scheduler (``ContinuousBatcher``) only talks to an executor through ``estimate_bytes``, ``prepare``,
``launch``, ``is_done``, ``finish`` and optionally ``exact_bytes``, so its admission and failure handling can be exercised with a CPU
executor. ``UNetExecutor`` is the GPU executor for ``fvdb.nn.SparseUNet``.
"""
from __future__ import annotations

import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, List, Optional, Sequence

import torch

from .sparse_conv_pack_info import AUTO_ORDER_WINDOW, wants_row_order

__all__ = ["Request", "ContinuousBatcher", "UNetExecutor", "UNetTopology", "latency_stats"]


@dataclass
class Request:
    rid: int
    ijk: Any                      # [N, 3] int32 (host or device)
    features: Any = None          # [N, C] or None (then generated from `seed` on the device)
    seed: int = 0
    arrival: float = 0.0          # seconds on the batcher's clock
    status: str = "new"           # new -> pending -> [ingesting ->] prepared -> running -> done | failed
    error: Optional[BaseException] = None
    result: Any = None
    reserved: int = 0
    t_admit: float = math.nan
    t_prepared: float = math.nan  # ingest finished (host side)
    t_launch: float = math.nan
    t_done: float = math.nan
    state: Any = None             # executor-private (topology, device tensors)

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


class ContinuousBatcher:
    """Admission by memory budget, FIFO batching, bounded batches in flight."""

    def __init__(self, executor, hbm_budget_bytes: int, max_batch_voxels: int = 4_000_000, max_inflight: int = 2,
                 clock: Callable[[], float] = time.perf_counter, ingest_workers: int = 0):
        self.ex = executor
        self.budget = int(hbm_budget_bytes)
        self.max_batch_voxels = int(max_batch_voxels)
        self.max_inflight = int(max_inflight)
        self.clock = clock
        self.t0 = clock()
        self.reserved = 0
        self.peak_reserved = 0
        # ingest_workers > 0: admitted requests are ingested (``prepare``) on host threads, so the
        # loop keeps completing and launching batches while an ingest waits for the device
        # (a grid build reads its sizes back); 0: ingest inline in the loop
        self._pool = (ThreadPoolExecutor(max_workers=int(ingest_workers), thread_name_prefix="fvdb-ingest")
                      if ingest_workers > 0 else None)
        self.queue: List[Request] = []      # arrived, waiting for memory
        self.ingesting: List[tuple] = []    # (future, request), admission order
        self.prepared: List[Request] = []   # ingested, waiting for a batch
        self.inflight: List[tuple] = []     # (handle, requests)
        self.done: List[Request] = []
        self.batches: List[int] = []        # batch sizes launched

    def now(self) -> float:
        return self.clock() - self.t0

    def submit(self, req: Request) -> None:
        try:
            self.ex.validate(req)
        except (ValueError, TypeError) as e:
            self._fail([req], e)
            return
        req.status = "pending"
        self.queue.append(req)

    def _release(self, reqs: Sequence[Request]) -> None:
        for r in reqs:
            self.reserved -= r.reserved
            r.reserved = 0

    def _fail(self, reqs: Sequence[Request], err: BaseException) -> None:
        self._release(reqs)
        t = self.now()
        for r in reqs:
            r.status = "failed"
            r.error = err
            r.t_done = t
            r.state = None
            self.done.append(r)

    def _admit(self) -> None:
        while self.queue:
            r = self.queue[0]
            need = self.ex.estimate_bytes(r)
            if need > self.budget:
                self.queue.pop(0)
                self._fail([r], ValueError(f"request {r.rid} needs {need} bytes, over the budget {self.budget}"))
                continue
            if self.reserved + need > self.budget:
                break  # FIFO: wait for memory
            self.queue.pop(0)
            r.reserved = need
            self.reserved += need
            self.peak_reserved = max(self.peak_reserved, self.reserved)
            r.t_admit = self.now()
            if self._pool is not None:
                r.status = "ingesting"
                self.ingesting.append((self._pool.submit(self.ex.prepare, r), r))
                continue
            try:
                self.ex.prepare(r)
            except Exception as e:  # noqa: BLE001 - a bad request fails alone
                self._fail([r], e)
                continue
            self._ingested(r)

    def _ingested(self, r: Request) -> None:
        # once ingested, the request's exact footprint is known: give back the slack of the
        # admission bound (never more than was reserved)
        exact = getattr(self.ex, "exact_bytes", None)
        if exact is not None:
            need = min(int(exact(r)), r.reserved)
            self.reserved -= r.reserved - need
            r.reserved = need
        r.status = "prepared"
        r.t_prepared = self.now()
        self.prepared.append(r)

    def _collect_ingest(self) -> None:
        # in admission order: a request waits for the ones admitted before it
        while self.ingesting and self.ingesting[0][0].done():
            fut, r = self.ingesting.pop(0)
            err = fut.exception()
            if err is not None:
                self._fail([r], err)
            else:
                self._ingested(r)

    def close(self) -> None:
        """Stops the ingest threads (after the ingests in progress)."""
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None

    def _complete(self) -> None:
        still = []
        for handle, reqs in self.inflight:
            try:
                ready = self.ex.is_done(handle)
            except Exception as e:  # noqa: BLE001
                self._fail(reqs, e)
                continue
            if not ready:
                still.append((handle, reqs))
                continue
            try:
                outs = self.ex.finish(handle)
            except Exception as e:  # noqa: BLE001
                self._fail(reqs, e)
                continue
            t = self.now()
            self._release(reqs)
            for r, o in zip(reqs, outs):
                r.result = o
                r.status = "done"
                r.t_done = t
                r.state = None
                self.done.append(r)
        self.inflight = still

    def _launch(self) -> None:
        while self.prepared and len(self.inflight) < self.max_inflight:
            batch, vox = [], 0
            while self.prepared:
                n = self.ex.voxels(self.prepared[0])
                if batch and vox + n > self.max_batch_voxels:
                    break
                batch.append(self.prepared.pop(0))
                vox += n
            t = self.now()
            for r in batch:
                r.status = "running"
                r.t_launch = t
            try:
                handle = self.ex.launch(batch)
            except Exception as e:  # noqa: BLE001 - the batch fails, the server goes on
                self._fail(batch, e)
                continue
            self.batches.append(len(batch))
            self.inflight.append((handle, batch))

    def step(self) -> None:
        self._complete()
        self._collect_ingest()
        self._admit()
        self._launch()

    def idle(self) -> bool:
        return not (self.queue or self.ingesting or self.prepared or self.inflight)

    def serve(self, requests: Sequence[Request], poll_sleep: float = 0.0) -> List[Request]:
        """Feed requests at their arrival times (seconds from now) until all are done."""
        pending = sorted(requests, key=lambda r: r.arrival)
        self.t0 = self.clock()
        i = 0
        while i < len(pending) or not self.idle():
            t = self.now()
            while i < len(pending) and pending[i].arrival <= t:
                self.submit(pending[i])
                i += 1
            self.step()
            if self.ingesting:
                # a real sleep: a thread that only yields (sleep(0)) takes the GIL straight back,
                # and the ingest thread then waits out the interpreter's switch interval (5 ms)
                # after each device call it makes
                time.sleep(max(poll_sleep, 2e-5))
            elif poll_sleep:
                time.sleep(poll_sleep)
        return sorted(self.done, key=lambda r: r.rid)


def latency_stats(reqs: Sequence[Request]) -> Dict[str, float]:
    lat = sorted(r.latency for r in reqs if r.status == "done")
    if not lat:
        return {"p50_ms": math.nan, "p99_ms": math.nan, "mean_ms": math.nan}

    def pct(p):
        k = min(len(lat) - 1, max(0, int(math.ceil(p / 100.0 * len(lat))) - 1))
        return lat[k] * 1e3

    return {"p50_ms": pct(50), "p99_ms": pct(99), "mean_ms": 1e3 * sum(lat) / len(lat)}


# -----------------------------------------------------------------------------------------------
# GPU executor for fvdb.nn.SparseUNet
# -----------------------------------------------------------------------------------------------


class UNetTopology:
    """Per-request device-resident maps of a SparseUNet with L levels.

    n[l]: voxels at level l; conv[l]: [n_l, 27] 3x3x3 map; down[l]: [n_{l+1}, 8] map of the stride-2 conv
    from level l to l+1 (entries index level l); up[l]: [n_l, 8] transposed map (entries index level l+1).
    """

    def __init__(self, grid, levels: int):
        self.n: List[int] = []
        self.conv: List[torch.Tensor] = []
        self.down: List[torch.Tensor] = []
        self.up: List[torch.Tensor] = []
        g = grid
        for l in range(levels):
            n = g.total_voxels
            self.n.append(n)
            pc, _ = g.sparse_conv_kernel_map(3, 1)
            self.conv.append(pc._map_fwd()[:n])
            if l + 1 < levels:
                pd, coarse = g.sparse_conv_kernel_map(2, 2)
                self.down.append(pd._map_fwd()[: coarse.total_voxels])
                self.up.append(pd._map_bwd()[:n])
                g = coarse

    def nbytes(self) -> int:
        return sum(t.numel() * 4 for t in self.conv + self.down + self.up)


def _shift_cat(maps: Sequence[torch.Tensor], offsets: Sequence[int]) -> torch.Tensor:
    """Concatenate maps, adding offsets[i] to the valid entries of maps[i] (one elementwise
    pass per shifted map; no host synchronisation)."""
    if len(maps) == 1:
        return maps[0]
    parts = [m if int(o) == 0 else torch.where(m >= 0, m + int(o), m) for m, o in zip(maps, offsets)]
    return torch.cat(parts, 0)


class UNetExecutor:
    """Runs a SparseUNet over batches of requests with device-resident per-request maps."""

    def __init__(self, model, dtype=torch.bfloat16, device="cuda", in_channels: Optional[int] = None,
                 host_results: bool = False, compute_streams: int = 3):
        import fvdb

        from ._autograd import weight_to_kio

        self.fvdb = fvdb
        self.device = torch.device(device)
        self.dtype = dtype
        self.host_results = bool(host_results)  # results returned in pinned host memory
        self.model = model
        self.levels = len(model.channels)
        self.cin = in_channels or model.enc[0][0].in_channels
        # one ingest stream per ingesting host thread, so concurrent ingests overlap on the device
        # and each grid build's read-back waits only for its own work
        self._ingest = threading.local()
        # a batch runs on a compute stream whose previous batch has finished, so with as many
        # streams as batches in flight a small batch does not queue behind a large one on the
        # device (every stream at the default priority)
        self.compute_streams = [torch.cuda.Stream(device=self.device) for _ in range(max(1, int(compute_streams)))]
        self._stream_done: List[Optional[torch.cuda.Event]] = [None] * len(self.compute_streams)
        self._next_stream = 0
        # results go to pinned host memory on a stream of their own, so a batch's copy overlaps
        # the next batch's compute (every stream at the default priority)
        self.copy_stream = torch.cuda.Stream(device=self.device) if self.host_results else None

        def pack(conv):
            w = conv.weight.detach().to(self.device, dtype)
            kio = weight_to_kio(w, conv.transposed)  # [K, I, O]
            b = conv.bias.detach().to(self.device, dtype) if conv.bias is not None else None
            return (kio.transpose(1, 2).contiguous(), b)

        self.enc = [[pack(c) for c in convs] for convs in model.enc]
        self.down = [pack(c) for c in model.down]
        self.up = [pack(c) for c in model.up]
        self.dec = [[pack(c) for c in convs] for convs in model.dec]
        self.head = None
        if model.head is not None:
            w = model.head.weight.detach().to(self.device, dtype)
            b = model.head.bias.detach().to(self.device, dtype) if model.head.bias is not None else None
            self.head = (w, b)
        self.chans = list(model.channels)

    # -- scheduler interface ----------------------------------------------------------------------
    def validate(self, req: Request) -> None:
        ijk = req.ijk
        if not isinstance(ijk, torch.Tensor) or ijk.dim() != 2 or ijk.shape[1] != 3:
            raise ValueError(f"request {req.rid}: ijk must be a [N, 3] tensor")
        if ijk.is_floating_point():
            raise TypeError(f"request {req.rid}: ijk must be integer")
        if ijk.shape[0] == 0:
            raise ValueError(f"request {req.rid}: empty grid")
        if req.features is not None and (req.features.shape[0] != ijk.shape[0] or req.features.shape[1] != self.cin):
            raise ValueError(f"request {req.rid}: features must have shape [N, {self.cin}]")

    def voxels(self, req: Request) -> int:
        return int(req.state["n0"]) if isinstance(req.state, dict) else int(req.ijk.shape[0])

    def _footprint(self, n_levels: Sequence[int], n0: int) -> int:
        """Device bytes a request with `n_levels[l]` voxels at level l holds from ingest to
        completion, an upper bound over the executor's allocations:

        * ingest (transient): the grid builds of two consecutive levels, <= 160 B per input
          voxel each (coordinates, leaf keys, the de-duplication set, masks, hash), and the
          generated features in fp32 before their cast;
        * per level: the 3x3x3 map ([pad128(n_l), 27] int32), the stride-2 map and its
          transpose ([pad128(n), 8] int32 each), each once more in a batch's concatenated
          copy, plus a row order (int32) per map;
        * activations in the feature dtype: the level's skip (c_l), and at most 5 c_l more
          at once (up-conv output, the 2 c_l concatenation, a convolution output);
        * the input features (cin, twice: request and batch copy) and the output (c_0).
        """
        el = torch.tensor([], dtype=self.dtype).element_size()
        pad = lambda v: (int(v) + 127) // 128 * 128  # noqa: E731
        total = (320 + 4 * self.cin) * n0 + 2 * n0 * self.cin * el + n0 * self.chans[0] * el
        for l, c in enumerate(self.chans):
            n = int(n_levels[l])
            nn_ = int(n_levels[l + 1]) if l + 1 < len(self.chans) else 0
            total += 2 * 4 * (pad(n) * 27 + pad(nn_) * 8 + pad(n) * 8) + 3 * 4 * pad(n)
            total += n * 6 * c * el
        return int(total) + (1 << 20)

    def estimate_bytes(self, req: Request) -> int:
        """Upper bound before ingest.  Only the input voxel count N is known, and a stride-2
        coarsening keeps at most as many voxels as it is given (every coarse voxel has a fine
        voxel), so every level is bounded by N."""
        n = int(req.ijk.shape[0])
        return self._footprint([n] * self.levels, n)

    def exact_bytes(self, req: Request) -> int:
        """The same bound with the level sizes of the ingested request."""
        topo = req.state["topo"]
        return self._footprint(topo.n, topo.n[0])

    def _ingest_stream(self) -> torch.cuda.Stream:
        st = getattr(self._ingest, "stream", None)
        if st is None:
            st = self._ingest.stream = torch.cuda.Stream(device=self.device)
        return st

    def prepare(self, req: Request) -> None:
        fvdb = self.fvdb
        ingest_stream = self._ingest_stream()
        with torch.cuda.stream(ingest_stream):
            ijk = req.ijk.to(self.device, torch.int32, non_blocking=True)
            grid = fvdb.GridBatch.from_ijk(ijk)
            topo = UNetTopology(grid, self.levels)
            if req.features is not None:
                x = req.features.to(self.device, self.dtype, non_blocking=True)
            else:
                gen = torch.Generator(device=self.device).manual_seed(int(req.seed))
                x = torch.randn((grid.total_voxels, self.cin), generator=gen, device=self.device).to(self.dtype)
            ev = torch.cuda.Event()
            ev.record(ingest_stream)
        req.state = {"topo": topo, "x": x, "ready": ev, "n0": grid.total_voxels}

    def _pick_stream(self) -> int:
        n = len(self.compute_streams)
        for k in range(n):
            i = (self._next_stream + k) % n
            ev = self._stream_done[i]
            if ev is None or ev.query():
                break
        else:
            i = self._next_stream
        self._next_stream = (i + 1) % n
        return i

    def launch(self, reqs: Sequence[Request]):
        ext = self.fvdb.ext
        si = self._pick_stream()
        stream = self.compute_streams[si]
        with torch.cuda.stream(stream):
            for r in reqs:
                stream.wait_event(r.state["ready"])
            topos = [r.state["topo"] for r in reqs]
            L = self.levels
            offs = [[0] for _ in range(L)]
            for t in topos:
                for l in range(L):
                    offs[l].append(offs[l][-1] + t.n[l])
            conv = [_shift_cat([t.conv[l] for t in topos], offs[l][:-1]) for l in range(L)]
            down = [_shift_cat([t.down[l] for t in topos], offs[l][:-1]) for l in range(L - 1)]
            up = [_shift_cat([t.up[l] for t in topos], offs[l + 1][:-1]) for l in range(L - 1)]
            x = torch.cat([r.state["x"] for r in reqs], 0) if len(reqs) > 1 else reqs[0].state["x"]

            orders = {}

            def layer(inp, wb, m, n_rows, cat=None):
                a, b = wb
                # rows grouped by tap mask per window where that pays (same rule as the packs)
                order = None
                cin = inp.shape[1] * (2 if cat is not None else 1)
                if n_rows > 0 and wants_row_order(m.shape[1], cin, inp.dtype, a.shape[1], n_rows):
                    if id(m) not in orders:
                        orders[id(m)] = ext.row_order(m, n_rows, AUTO_ORDER_WINDOW)
                    order = orders[id(m)]
                if cat is not None:  # input channels [inp | cat], read from both
                    return ext.implicit_gemm_cat2(inp, cat, a, m, n_rows, order, b, True)
                return ext.implicit_gemm(inp, a, m, n_rows, order, b, True)  # bias + ReLU in the epilogue

            skips = []
            for l in range(L):
                n_l = offs[l][-1]
                if l > 0:
                    x = layer(x, self.down[l - 1], down[l - 1], n_l)
                for wb in self.enc[l]:
                    x = layer(x, wb, conv[l], n_l)
                skips.append(x)
            for i, l in enumerate(range(L - 2, -1, -1)):
                n_l = offs[l][-1]
                x = layer(x, self.up[i], up[l], n_l)
                x = layer(x, self.dec[i][0], conv[l], n_l, cat=skips[l])  # the skip concatenation
                for wb in self.dec[i][1:]:
                    x = layer(x, wb, conv[l], n_l)
            if self.head is not None:
                w, b = self.head
                x = x.matmul(w.transpose(0, 1))
                if b is not None:
                    x = x + b
            done = torch.cuda.Event()
            done.record(stream)
            self._stream_done[si] = done
            # tensors made on the ingest stream are read here: tie their memory to this stream too
            for r in reqs:
                tp = r.state["topo"]
                for t in [r.state["x"]] + tp.conv + tp.down + tp.up:
                    t.record_stream(stream)
        if self.host_results:
            with torch.cuda.stream(self.copy_stream):
                self.copy_stream.wait_event(done)
                host = torch.empty(x.shape, dtype=x.dtype, pin_memory=True)
                host.copy_(x, non_blocking=True)
                x.record_stream(self.copy_stream)
                done = torch.cuda.Event()
                done.record(self.copy_stream)
            x = host
        return {"event": done, "out": x, "offs": offs[0]}

    def is_done(self, handle) -> bool:
        return handle["event"].query()

    def finish(self, handle) -> List[torch.Tensor]:
        o, out = handle["offs"], handle["out"]  # out is in pinned host memory with host_results
        return [out[o[i]:o[i + 1]] for i in range(len(o) - 1)]
