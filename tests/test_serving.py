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


def test_latency_stats():
    rs = reqs([1, 1, 1, 1])
    for i, r in enumerate(rs):
        r.status, r.arrival, r.t_done = "done", 0.0, 0.001 * (i + 1)
    s = latency_stats(rs)
    assert math.isclose(s["p50_ms"], 2.0) and math.isclose(s["p99_ms"], 4.0)


@pytest.mark.gpu
def test_gpu_engine_matches_module_alone_and_in_mix():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import fvdb
    from fvdb.serving import UNetExecutor
    from fvdb.utils.synthetic import sphere_points

    torch.manual_seed(0)
    model = fvdb.nn.SparseUNet(in_channels=32, channels=(32, 64, 128)).cuda().eval()
    rs = []
    for i, (n, vs) in enumerate([(40_000, 1 / 64), (120_000, 1 / 128), (20_000, 1 / 48)]):
        pts = sphere_points(seed=10 + i, n_points=n, device="cuda")
        g = fvdb.GridBatch.from_points(pts, voxel_sizes=vs)
        rs.append(Request(rid=i, ijk=g.ijk.jdata.cpu(), seed=100 + i))
    for dtype in (torch.bfloat16, torch.float32):
        ex = UNetExecutor(model, dtype=dtype)
        alone = []
        for r in rs:
            b = ContinuousBatcher(ex, hbm_budget_bytes=8 << 30)
            alone.append(b.serve([Request(rid=r.rid, ijk=r.ijk, seed=r.seed)])[0])
        b = ContinuousBatcher(ex, hbm_budget_bytes=8 << 30, max_batch_voxels=10**8)
        mixed = b.serve([Request(rid=r.rid, ijk=r.ijk, seed=r.seed) for r in rs])
        assert b.batches[0] == 3, b.batches
        for a, m in zip(alone, mixed):
            assert a.status == "done" and m.status == "done"
            assert torch.equal(a.result, m.result), f"request {a.rid} differs alone vs in the mix"
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
