"""GPU parity of the mixed-size persistent batch: per-field results against the
oracle, and each field identical (bitwise) alone and in the mix."""
import numpy as np
import pytest
import torch

from oracle import compact, cr

pytestmark = pytest.mark.gpu


def rel_err(x, ref):
    return np.abs(np.asarray(x) - np.asarray(ref)).max() / np.abs(np.asarray(ref)).max()


def _jobs(dev, seed=1, count=12):
    rng = np.random.default_rng(seed)
    jobs, host = [], []
    sizes = [5, 17, 32, 64, 96, 100, 128, 160, 257, 512]   # 5 only across lines
    for i in range(count):
        shape = [int(rng.choice(sizes)) for _ in range(3)]
        axis = int(rng.integers(0, 3))
        if shape[axis] < 16:
            shape[axis] = 64
        # keep every field small: at most ~4 M points
        while np.prod(shape) > 4_000_000:
            k = int(np.argmax([s if a != axis else 0 for a, s in enumerate(shape)]))
            shape[k] = max(4, shape[k] // 2)
        f = rng.standard_normal(shape)
        f32 = bool(i % 4 == 3)
        prec = "fp32_refine" if i % 3 == 2 else "fp64"
        ft = torch.tensor(f.astype(np.float32) if f32 else f, device=dev)
        out = torch.empty(shape, dtype=torch.float64, device=dev)
        dx = 0.1 + 0.01 * i
        jobs.append(dict(f=ft, out=out, axis=axis, dx=dx, precision=prec, sweeps=2))
        host.append((f.astype(np.float32).astype(np.float64) if f32 else f, axis, dx, prec))
    return jobs, host


def test_mixed_batch_matches_oracle(dev):
    from ntamd.batch import MixedBatch
    jobs, host = _jobs(dev)
    b = MixedBatch(jobs)
    b.run()
    torch.cuda.synchronize()
    for j, (f, axis, dx, prec) in zip(jobs, host):
        want = compact.derivative(f, dx, axis)
        tol = 1e-12 if prec == "fp64" else 1e-10
        assert rel_err(j["out"].cpu().numpy(), want) < tol, (f.shape, axis, prec)


def test_mixed_batch_alone_equals_in_mix(dev):
    from ntamd.batch import MixedBatch
    jobs, _ = _jobs(dev, seed=2, count=10)
    MixedBatch(jobs).run()
    torch.cuda.synchronize()
    mixed = [j["out"].clone() for j in jobs]
    for j, m in zip(jobs, mixed):
        j["out"].zero_()
        MixedBatch([j], grid=7).run()
        torch.cuda.synchronize()
        assert torch.equal(j["out"], m)


def test_batch_in_place_solves_and_rerun(dev):
    from ntamd.batch import MixedBatch
    co = (2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7)
    d = np.random.default_rng(3).standard_normal((3, 40, 300))
    x = torch.tensor(d, device=dev)
    b = MixedBatch([dict(out=x, axis=2, solve=True, coeffs=co)])
    b.run()
    assert rel_err(x.cpu().numpy(), cr.cr_solve(co, d)) < 1e-12
    x.copy_(torch.tensor(d, device=dev))
    b.run()                                  # the counter is reset on every run
    assert rel_err(x.cpu().numpy(), cr.cr_solve(co, d)) < 1e-12


def test_batch_rejects_bad_jobs(dev):
    from ntamd.batch import MixedBatch
    from ntamd import NtInvalidArgument, NtError
    f = torch.zeros((4, 4, 600), dtype=torch.float64, device=dev)
    with pytest.raises(NtError):
        MixedBatch([dict(f=f, out=torch.empty_like(f), axis=2, dx=0.1)])   # n = 600 > 512
    g = torch.zeros((4, 8, 600), dtype=torch.float64, device=dev)
    with pytest.raises(NtError):
        MixedBatch([dict(f=g, out=torch.empty_like(g), axis=1, dx=0.1)])   # n = 8 < 16
    with pytest.raises(NtInvalidArgument):
        MixedBatch([dict(f=f, out=f, axis=1, dx=0.1)])                      # aliasing
