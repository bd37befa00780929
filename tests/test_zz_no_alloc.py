"""The allocation-free mixed batch (nt_batch_run_noalloc) against the existing
entry point (nt_batch_pack + nt_batch_launch, through MixedBatch).

Both run the same kernel on the same job table, so the results must be equal
bit for bit.  Two calls in a row, the second with fewer and smaller fields,
check that the static device array's counter is cleared and that the bytes a
larger earlier image left behind are not read.  Sizes beyond the array are
refused with an error before anything is enqueued."""
import numpy as np
import pytest
import torch

from oracle import compact

pytestmark = pytest.mark.gpu


def _jobs(dev, specs, seed):
    from ntamd.workloads import compact_field_device
    jobs = []
    for k, (shape, axis, precision, f32) in enumerate(specs):
        f = compact_field_device(shape, seed=seed + k, device=dev,
                                 dtype=torch.float32 if f32 else torch.float64)
        jobs.append(dict(f=f, out=torch.full(shape, float("nan"), dtype=torch.float64, device=dev),
                         axis=axis, dx=1.0 / (shape[axis] - 1), precision=precision, sweeps=1))
    return jobs


def _fresh_outputs(jobs):
    return [dict(j, out=torch.full_like(j["out"], float("nan"))) for j in jobs]


CALL1 = [((64, 96, 128), 2, "fp64", False), ((160, 64, 96), 0, "fp32_refine", False),
         ((96, 288, 64), 1, "fp64", True), ((256, 40, 130), 2, "fp32_refine", True),
         ((33, 500, 48), 1, "fp32_refine", False), ((512, 24, 40), 0, "fp64", False),
         ((70, 70, 70), 1, "fp64", False), ((20, 30, 512), 2, "fp64", False)]
CALL2 = [((48, 40, 64), 0, "fp64", False), ((16, 17, 100), 2, "fp32_refine", True)]


def test_noalloc_equals_batch_launch_on_two_calls(dev):
    from ntamd.batch import MixedBatch, run_noalloc
    for call, specs in enumerate((CALL1, CALL2)):
        print(f"[no_alloc] call {call}: {len(specs)} jobs", flush=True)
        jobs = _jobs(dev, specs, seed=500 + 50 * call)
        MixedBatch(jobs).run()
        mine = _fresh_outputs(jobs)
        run_noalloc(mine)
        torch.cuda.synchronize()
        for k, (a, b) in enumerate(zip(jobs, mine)):
            assert not torch.isnan(b["out"]).any(), (call, k)
            assert torch.equal(a["out"], b["out"]), (call, k)
        # and the oracle on a few lines of each field
        for k, j in enumerate(mine):
            axis = j["axis"]
            f = j["f"].double().cpu().numpy()
            want = compact.derivative_banded(f, j["dx"], axis)
            got = j["out"].cpu().numpy()
            tol = 1e-12 if j["precision"] == "fp64" else 1e-10
            assert np.abs(got - want).max() / np.abs(want).max() < tol, (call, k)


def test_noalloc_repeated_calls_on_one_stream(dev):
    """Calls queued back to back on one stream without a host wait between them
    are ordered on the static array: each sees its own image and counter."""
    from ntamd.batch import MixedBatch, run_noalloc
    a = _jobs(dev, CALL1[:4], seed=900)
    b = _jobs(dev, CALL2, seed=950)
    MixedBatch(a).run()
    MixedBatch(b).run()
    ma, mb = _fresh_outputs(a), _fresh_outputs(b)
    run_noalloc(ma)
    run_noalloc(mb)
    run_noalloc(ma)
    torch.cuda.synchronize()
    for x, y in zip(a + b, ma + mb):
        assert torch.equal(x["out"], y["out"])


def test_noalloc_refuses_images_beyond_the_array(dev):
    from ntamd import _lib
    from ntamd.batch import run_noalloc
    lib = _lib.load()
    f = torch.zeros((2, 2, 64), dtype=torch.float64, device=dev)
    o = torch.zeros_like(f)
    cap = _lib.NT_BATCH_NOALLOC_IMAGE_MAX
    njobs = 1
    while lib.nt_batch_image_bytes(njobs, 1) <= cap:
        njobs *= 2
    jobs = [dict(f=f, out=o, axis=2, dx=0.1)] * njobs
    with pytest.raises(_lib.NtInvalidArgument, match="static scratch"):
        run_noalloc(jobs)
    # nothing was enqueued and the array still works
    small = _jobs(dev, CALL2, seed=990)
    run_noalloc(small)
    torch.cuda.synchronize()
    assert not any(torch.isnan(j["out"]).any() for j in small)
