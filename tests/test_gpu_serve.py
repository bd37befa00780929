"""Serving loop on the GPU: every request completes, results are identical
(bitwise) alone and in the mix, and match the oracle."""
import numpy as np
import pytest
import torch

from oracle import adi, compact

pytestmark = pytest.mark.gpu


def _small_requests(count=24, rate=20000.0, seed=5):
    from ntamd.serve import make_requests
    return make_requests(count, rate, seed=seed, deriv_sizes=(32, 96), adi_sides=(64, 200),
                         adi_steps=(1, 4))


def test_serve_continuous_completes_and_matches_alone(dev):
    from ntamd.serve import Server
    reqs = _small_requests()
    # fields above 300k points take the fused per-request launch, the rest the
    # persistent batch: both routes are in the mix
    srv = Server(keep_results=True, hbm_budget=1 << 30, small_points=300_000)
    stats = srv.run(reqs, mode="continuous")
    assert stats["requests"] == len(reqs)
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
    for q in reqs:
        got = q.result.cpu().numpy()
        if q.kind == "deriv":
            f = trig_field(q.shape, q.seed, "cuda").cpu().numpy()
            want = compact.derivative(f, 1.0 / (q.shape[q.axis] - 1), q.axis)
        else:
            u0 = adi_field(q.side, q.seed, "cuda").cpu().numpy()
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
