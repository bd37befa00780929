"""Every launch-shape variant of the tuning table gives the oracle's answer.

The variants differ only in how lines are mapped onto wavefronts and staged
(wave-private LDS transpose, direct chunk loads, workgroup tiles, strided tiles
of 16 or 32 lines), so each must match the SciPy banded oracle to the FP64 (or
FP32 + refinement) tolerance, and the variants must agree with each other to a
few roundings."""
import numpy as np
import pytest
import torch

from oracle import compact

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32_refine": 1e-10}


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("n", [64, 100, 256, 512, 768, 1024, 2048])
@pytest.mark.parametrize("axis", [2, 1])
def test_every_variant_matches_oracle(dev, precision, n, axis):
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.autotune import _set, current
    from ntamd.workloads import compact_field
    layout = 0 if axis == 2 else 1
    shape = (3, 40, n) if axis == 2 else (3, n, 40)
    f = compact_field(shape, seed=n)
    h = 1.0 / (n - 1)
    want = compact.derivative_banded(f, h, axis)
    fd = torch.tensor(f, device=dev)
    plan = Plan(n, COMPACT_COEFFS)
    keep, nv = current(layout, n, precision)
    outs = []
    try:
        for v in range(nv):
            _set(layout, precision, n, v)
            out = torch.full_like(fd, float("nan"))
            core.compact_derivative(plan, fd, out, h, axis, precision)
            got = out.cpu().numpy()
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < TOL[precision], (v, err)
            outs.append(got)
    finally:
        _set(layout, precision, n, keep)
    scale = np.abs(want).max()
    for v in range(1, nv):
        assert np.abs(outs[v] - outs[0]).max() / scale < 1e-13, v
