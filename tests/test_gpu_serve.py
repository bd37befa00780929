"""Serving loop on the GPU: every request completes, results are identical
(bitwise) alone and in the mix, and match the oracle."""
import numpy as np
import pytest
import torch

from oracle import adi, compact

pytestmark = pytest.mark.gpu


def _small_requests(count=24, rate=20000.0, seed=5, upload_share=0.5):
    from ntamd.serve import make_requests
    return make_requests(count, rate, seed=seed, deriv_sizes=(32, 96), adi_sides=(64, 200),
                         adi_steps=(1, 4), upload_share=upload_share)


def test_serve_continuous_completes_and_matches_alone(dev):
    from ntamd.serve import Server
    reqs = _small_requests()
    # fields above 300k points take the fused per-request launch, the rest the
    # persistent batch: both routes are in the mix
    srv = Server(keep_results=True, hbm_budget=1 << 30, small_points=300_000)
    stats = srv.run(reqs, mode="continuous")
    assert stats["requests"] == len(reqs)
    assert 0 < stats["uploads"] < len(reqs)                  # both input sources are in the mix
    assert all(np.isfinite(q.latency) and q.latency >= 0 for q in reqs)
    assert stats["batches"] >= 1
    assert any(q.kind == "deriv" and not srv._batchable(q) for q in reqs)
    mixed = {q.rid: q.result.clone() for q in reqs}
    for q in reqs:
        alone = Server(keep_results=True, hbm_budget=1 << 30, small_points=300_000)
        alone.run([q], mode="continuous")
        assert torch.equal(q.result, mixed[q.rid]), (q.kind, q.rid)


def test_serve_results_match_oracle(dev):
    from ntamd.serve import Server, trig_field, adi_field
    reqs = _small_requests(count=10, seed=6)
    srv = Server(keep_results=True, hbm_budget=1 << 30)
    srv.run(reqs, mode="continuous")
    assert any(q.upload for q in reqs) and not all(q.upload for q in reqs)
    for q in reqs:
        got = q.result.cpu().numpy()
        if q.kind == "deriv":
            f = (srv.host_pool.view(q.shape).numpy() if q.upload
                 else trig_field(q.shape, q.seed, "cuda").cpu().numpy())
            want = compact.derivative(f, 1.0 / (q.shape[q.axis] - 1), q.axis)
        else:
            u0 = (srv.host_pool.view((q.side, q.side)).numpy() if q.upload
                  else adi_field(q.side, q.seed, "cuda").cpu().numpy())
            want = adi.pr_run(u0, q.r, q.steps)
        tol = 1e-12 if q.precision == "fp64" else 1e-10
        assert np.abs(got - want).max() / np.abs(want).max() < tol, (q.kind, q.precision)


def test_serve_admission_waits_for_memory(dev):
    from ntamd.serve import Server
    reqs = _small_requests(count=12, seed=7)
    biggest = max(q.nbytes() for q in reqs)
    srv = Server(keep_results=False, hbm_budget=biggest)      # room for about one request at a time
    stats = srv.run(reqs, mode="continuous")
    assert stats["requests"] == len(reqs)
    assert all(np.isfinite(q.done) for q in reqs)


def test_serve_sequential_baseline(dev):
    from ntamd.serve import Server
    reqs = _small_requests(count=8, seed=8)
    stats = Server(hbm_budget=1 << 30).run(reqs, mode="sequential")
    assert stats["requests"] == 8 and stats["batches"] == 0


def test_serve_warmup_then_run(dev):
    from ntamd.serve import Server
    srv = Server(hbm_budget=1 << 30)
    srv.warmup()
    reqs = _small_requests(count=6, seed=9)
    stats = srv.run(reqs, mode="continuous")
    assert stats["requests"] == 6 and all(np.isfinite(q.done) for q in reqs)


def test_image_pool_reuses_slots(dev):
    """More batches than slots through one ImagePool, on two streams: every
    batch equals a MixedBatch of the same jobs, bit for bit."""
    from ntamd.batch import ImagePool, MixedBatch
    from ntamd.workloads import compact_field_device
    pool = ImagePool(slots=2, slot_bytes=1 << 16)
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    runs = []
    for k in range(7):
        shape = (8 + k, 20, 64 + 16 * k)
        f = compact_field_device(shape, seed=k, device=dev)
        jobs = [dict(f=f, out=torch.full_like(f, float("nan")), axis=a, dx=0.01 * (k + 1)) for a in (2, 1)]
        with torch.cuda.stream(streams[k % 2]):
            keep, ev = pool.run(jobs)
        runs.append((jobs, keep, ev))
    torch.cuda.synchronize()
    for jobs, _, _ in runs:
        ref = [dict(j, out=torch.empty_like(j["out"])) for j in jobs]
        MixedBatch(ref).run()
        torch.cuda.synchronize()
        for a, b in zip(jobs, ref):
            assert torch.equal(a["out"], b["out"])


def test_serve_uploads_only(dev):
    """Every request's input crosses from pinned host memory; results match the
    oracle on that data, and the sequential baseline (separate fused launches)
    agrees with the continuous server on the FP64 requests."""
    from ntamd.serve import Server
    reqs = _small_requests(count=10, seed=12, upload_share=1.0)
    srv = Server(keep_results=True, hbm_budget=1 << 30)
    stats = srv.run(reqs, mode="continuous")
    assert stats["uploads"] == len(reqs)
    cont = {q.rid: q.result.clone() for q in reqs}
    for q in reqs:
        if q.kind == "deriv":
            f = srv.host_pool.view(q.shape).numpy()
            want = compact.derivative(f, 1.0 / (q.shape[q.axis] - 1), q.axis)
            tol = 1e-12 if q.precision == "fp64" else 1e-10
            assert np.abs(cont[q.rid].cpu().numpy() - want).max() / np.abs(want).max() < tol
    seq = Server(keep_results=True, hbm_budget=1 << 30)
    seq.run(reqs, mode="sequential")
    for q in reqs:
        if q.precision == "fp64":          # the baseline runs everything in FP64
            a, b = q.result.cpu().numpy(), cont[q.rid].cpu().numpy()
            assert np.abs(a - b).max() / np.abs(b).max() < 1e-12
