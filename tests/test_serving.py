"""Continuous-batching server: scheduler behaviour on the CPU (a fake executor forces
budget pressure, malformed requests and executor failures from the host), and on the
GPU the U-Net executor's results against the fvdb.nn module, alone and in a mix."""
import math

import pytest
import torch

from fvdb.serving import ContinuousBatcher, Request, latency_stats


class FakeExecutor:
    """CPU executor: 'runs' a batch by doubling features; can be told to fail."""

    def __init__(self, bytes_per_voxel=10, fail_batches=(), done_after=1):
        self.bpv = bytes_per_voxel
        self.fail_batches = set(fail_batches)
        self.launched = 0
        self.done_after = done_after
        self.max_seen_reserved = 0

    def validate(self, r):
        if not isinstance(r.ijk, torch.Tensor) or r.ijk.dim() != 2 or r.ijk.shape[1] != 3:
            raise ValueError("bad ijk")

    def voxels(self, r):
        return r.ijk.shape[0]

    def estimate_bytes(self, r):
        return self.bpv * r.ijk.shape[0]

    def prepare(self, r):
        if r.seed < 0:
            raise RuntimeError("ingest failure")
        r.state = {"x": torch.full((r.ijk.shape[0], 1), float(r.rid))}

    def launch(self, reqs):
        b = self.launched
        self.launched += 1
        if b in self.fail_batches:
            raise RuntimeError(f"device failure in batch {b}")
        return {"polls": 0, "outs": [2 * r.state["x"] for r in reqs]}

    def is_done(self, h):
        h["polls"] += 1
        return h["polls"] >= self.done_after

    def finish(self, h):
        return h["outs"]


def reqs(sizes, start_id=0):
    return [Request(rid=start_id + i, ijk=torch.zeros(n, 3, dtype=torch.int32), arrival=0.0) for i, n in enumerate(sizes)]


def test_budget_admission_queues_and_never_exceeds():
    ex = FakeExecutor(bytes_per_voxel=10, done_after=3)
    b = ContinuousBatcher(ex, hbm_budget_bytes=2500, max_batch_voxels=1000, max_inflight=1)
    out = b.serve(reqs([100, 100, 150, 200, 120, 90]))
    assert all(r.status == "done" for r in out)
    assert b.peak_reserved <= 2500 and b.reserved == 0
    for r in out:
        assert torch.equal(r.result, torch.full((r.ijk.shape[0], 1), 2.0 * r.rid))


def test_oversized_request_rejected_others_served():
    ex = FakeExecutor(bytes_per_voxel=10)
    b = ContinuousBatcher(ex, hbm_budget_bytes=1000, max_batch_voxels=1000)
    out = b.serve(reqs([50, 500, 60]))
    assert [r.status for r in out] == ["done", "failed", "done"]
    assert isinstance(out[1].error, ValueError)
    assert b.reserved == 0


def test_malformed_request_fails_alone():
    ex = FakeExecutor()
    b = ContinuousBatcher(ex, hbm_budget_bytes=10**6)
    rs = reqs([10, 20, 30])
    rs[1].ijk = torch.zeros(20, 2, dtype=torch.int32)
    out = b.serve(rs)
    assert out[1].status == "failed" and isinstance(out[1].error, ValueError)
    assert out[0].status == "done" and out[2].status == "done"


def test_ingest_failure_frees_reservation():
    ex = FakeExecutor(bytes_per_voxel=1)
    b = ContinuousBatcher(ex, hbm_budget_bytes=100)
    rs = reqs([40, 40, 40])
    rs[0].seed = -1
    out = b.serve(rs)
    assert out[0].status == "failed" and isinstance(out[0].error, RuntimeError)
    assert out[1].status == "done" and out[2].status == "done" and b.reserved == 0


def test_executor_failure_fails_only_its_batch():
    ex = FakeExecutor(bytes_per_voxel=1, fail_batches={0})
    b = ContinuousBatcher(ex, hbm_budget_bytes=10**6, max_batch_voxels=25, max_inflight=1)
    out = b.serve(reqs([20, 20, 20]))
    assert out[0].status == "failed" and "device failure" in str(out[0].error)
    assert out[1].status == "done" and out[2].status == "done"
    assert b.reserved == 0


@pytest.mark.parametrize("workers", [1, 2])
def test_threaded_ingest_keeps_order_budget_and_failures(workers):
    ex = FakeExecutor(bytes_per_voxel=1, done_after=2)
    b = ContinuousBatcher(ex, hbm_budget_bytes=100, max_batch_voxels=50, max_inflight=2, ingest_workers=workers)
    rs = reqs([40, 40, 40, 30, 20])
    rs[2].seed = -1  # its ingest raises on the worker thread
    out = b.serve(rs)
    b.close()
    assert [r.status for r in out] == ["done", "done", "failed", "done", "done"]
    assert isinstance(out[2].error, RuntimeError) and b.reserved == 0 and b.peak_reserved <= 100
    for r in out:
        if r.status == "done":
            assert torch.equal(r.result, torch.full((r.ijk.shape[0], 1), 2.0 * r.rid))
            assert r.t_admit <= r.t_prepared <= r.t_launch <= r.t_done
    # ingests complete in admission order: no request was prepared before an earlier admitted one
    done = [r for r in out if r.status == "done"]
    assert [r.rid for r in sorted(done, key=lambda r: r.t_prepared)] == [r.rid for r in sorted(done, key=lambda r: r.t_admit)]


def test_latency_stats():
    rs = reqs([1, 1, 1, 1])
    for i, r in enumerate(rs):
        r.status, r.arrival, r.t_done = "done", 0.0, 0.001 * (i + 1)
    s = latency_stats(rs)
    assert math.isclose(s["p50_ms"], 2.0) and math.isclose(s["p99_ms"], 4.0)


def test_exact_bytes_lowers_reservation_after_ingest():
    """An executor that knows a request's footprint after ingest gives back the slack of the
    admission bound, so a request that waited on the bound is admitted once the first is ingested."""

    class Exact(FakeExecutor):
        def exact_bytes(self, r):
            return 2 * r.ijk.shape[0]

    ex = Exact(bytes_per_voxel=10, done_after=50)
    b = ContinuousBatcher(ex, hbm_budget_bytes=1500, max_batch_voxels=10, max_inflight=1)
    for r in reqs([100, 100]):
        b.submit(r)
    b.step()
    # first admitted at 1000, lowered to 200 after ingest; the second (1000) fits beside it
    assert b.reserved == 400 and len(b.prepared) + len(b.inflight) == 2
    out = b.serve([])
    assert all(r.status == "done" for r in out) and b.reserved == 0


def _requests(fvdb, sizes, seed0):
    from fvdb.utils.synthetic import sphere_points_for_voxels, voxelize

    rs = []
    for i, nv in enumerate(sizes):
        pts, vs = sphere_points_for_voxels(seed0 + i, nv, device="cuda")
        g = voxelize(pts, vs)
        rs.append(Request(rid=i, ijk=g.ijk.jdata.cpu(), seed=100 + i))
    return rs


@pytest.mark.gpu
def test_gpu_engine_matches_module_alone_and_in_mix():
    """The serving workload's network (4 levels, 32/64/128/256) on requests of 20k to ~1.2M
    voxels: each result is bit-identical alone and in the mix, and equal to the fvdb.nn module's."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb
    from fvdb.serving import UNetExecutor

    torch.manual_seed(0)
    model = fvdb.nn.SparseUNet(in_channels=32, channels=(32, 64, 128, 256)).cuda().eval()
    rs = _requests(fvdb, [40_000, 1_200_000, 20_000], 10)
    assert rs[1].ijk.shape[0] >= 1_000_000
    print(f"[serve] request voxels {[r.ijk.shape[0] for r in rs]}", flush=True)
    for dtype in (torch.bfloat16, torch.float32):
        ex = UNetExecutor(model, dtype=dtype)
        alone = []
        for r in rs:
            b = ContinuousBatcher(ex, hbm_budget_bytes=64 << 30)
            alone.append(b.serve([Request(rid=r.rid, ijk=r.ijk, seed=r.seed)])[0])
        b = ContinuousBatcher(ex, hbm_budget_bytes=64 << 30, max_batch_voxels=10**8)
        mixed = b.serve([Request(rid=r.rid, ijk=r.ijk, seed=r.seed) for r in rs])
        assert b.batches[0] == 3, b.batches
        for a, m in zip(alone, mixed):
            assert a.status == "done" and m.status == "done"
            assert torch.equal(a.result, m.result), f"request {a.rid} differs alone vs in the mix"
        # the serving configuration: ingest on a host thread, batches on three compute streams at
        # once (staggered arrivals put the requests in separate batches that overlap on the device)
        b = ContinuousBatcher(ex, hbm_budget_bytes=64 << 30, max_inflight=3, ingest_workers=1)
        conc = b.serve([Request(rid=r.rid, ijk=r.ijk, seed=r.seed, arrival=2e-4 * r.rid) for r in rs])
        b.close()
        print(f"[serve] {dtype} concurrent batches {b.batches}", flush=True)
        for a, c in zip(alone, conc):
            assert c.status == "done" and torch.equal(a.result, c.result), f"request {a.rid} differs concurrently"
        # the engine runs the same kernels and maps as the fvdb.nn module
        mdl = model.to(dtype)
        with torch.no_grad():
            for r, a in zip(rs, alone):
                grid = fvdb.GridBatch.from_ijk(r.ijk.cuda())
                gen = torch.Generator(device="cuda").manual_seed(r.seed)
                x = torch.randn((grid.total_voxels, 32), generator=gen, device="cuda").to(dtype)
                y = mdl(fvdb.nn.VDBTensor(grid, grid.jagged_like(x)))
                assert torch.equal(y.data.jdata, a.result)
        model = model.float()


@pytest.mark.gpu
def test_gpu_admission_bounds_measured_memory():
    """A stream under a budget of 1.2 times the largest request's admission bound: the budget
    binds (requests wait for memory), every request completes, and the device memory the stream
    allocates beyond what was live before it never exceeds the budget."""
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb
    from fvdb.serving import UNetExecutor

    torch.manual_seed(0)
    model = fvdb.nn.SparseUNet(in_channels=32, channels=(32, 64, 128, 256)).cuda().eval()
    rs = _requests(fvdb, [300_000, 900_000, 60_000, 600_000, 150_000, 800_000, 30_000, 400_000], 40)
    ex = UNetExecutor(model, dtype=torch.bfloat16, host_results=True)
    budget = int(1.2 * max(ex.estimate_bytes(r) for r in rs))
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    b = ContinuousBatcher(ex, hbm_budget_bytes=budget, max_batch_voxels=4_000_000, max_inflight=3, ingest_workers=1)
    waited = []
    orig = b._admit

    def admit():
        orig()
        waited.append(len(b.queue))
    b._admit = admit
    out = b.serve([Request(rid=r.rid, ijk=r.ijk, seed=r.seed) for r in rs])
    b.close()
    torch.cuda.synchronize()
    peak = torch.cuda.max_memory_allocated() - base
    print(f"[serve budget] budget {budget / 1e9:.2f} GB, peak reserved by admission {b.peak_reserved / 1e9:.2f} GB, "
          f"measured peak {peak / 1e9:.2f} GB, batches {b.batches}", flush=True)
    assert all(r.status == "done" for r in out)
    assert max(waited) > 0, "the budget never held a request back"
    assert b.peak_reserved <= budget and b.reserved == 0
    assert peak <= budget
