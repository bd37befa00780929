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


def test_batch_every_job_class(dev):
    """One launch holding a job of every class the batch kernel specialises:
    x, y and z lines; 1, 2, 4 and 8 rows per lane, each with a power-of-two
    length (compile-time level structure) and a general one; FP64 and FP32 +
    refinement; in-place solves, compact derivatives of FP64 fields and of
    FP32-stored fields (FP32 rows at one row per lane included).  Strided
    fields with 16-line-aligned slabs take the 16-byte tile path, the others
    (37 lines) the element path.  Every job equals the oracle."""
    from ntamd.batch import MixedBatch
    co = (2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7)
    lengths = [16, 37, 64, 100, 128, 200, 256, 300, 512]     # R = 1, 1, 1(full), 2, 2(full), 4, 4(full), 8, 8(full)
    jobs, want = [], []
    rng = np.random.default_rng(21)
    for axis in (2, 1, 0):
        for n in lengths:
            for prec in ("fp64", "fp32_refine"):
                for kind in ("solve", "cd64", "cd32"):
                    other = 37 if (n + len(jobs)) % 2 else 32
                    shape = {2: (2, other, n), 1: (2, n, other), 0: (n, 1, other)}[axis]
                    f = rng.standard_normal(shape)
                    if kind == "cd32":
                        f = f.astype(np.float32).astype(np.float64)
                    h = 0.05 + 0.001 * len(jobs)
                    if kind == "solve":
                        x = torch.tensor(f, device=dev)
                        jobs.append(dict(out=x, axis=axis, solve=True, coeffs=co, precision=prec, sweeps=2))
                        ref = compact.last_to_axis(cr.banded_solve(co, compact.axis_to_last(f, axis)), axis)
                    else:
                        ft = torch.tensor(f.astype(np.float32) if kind == "cd32" else f, device=dev)
                        jobs.append(dict(f=ft, out=torch.full(shape, float("nan"), dtype=torch.float64, device=dev),
                                         axis=axis, dx=h, precision=prec, sweeps=2))
                        ref = compact.derivative_banded(f, h, axis)
                    want.append((ref, prec, (axis, n, prec, kind)))
    MixedBatch(jobs).run()
    torch.cuda.synchronize()
    for j, (ref, prec, tag) in zip(jobs, want):
        tol = 1e-12 if prec == "fp64" else 1e-10
        assert rel_err(j["out"].cpu().numpy(), ref) < tol, tag
