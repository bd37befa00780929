"""The HIP path against the committed golden fixtures."""
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_gpu_cr_lines_golden(dev):
    from ntamd import Plan, core
    g = load("cr_lines.npz")
    for i in range(int(g["ncases"])):
        co, d, x = g[f"c{i}_coeffs"], g[f"c{i}_d"], g[f"c{i}_x"]
        t = torch.tensor(d, device=dev)
        core.solve(Plan(d.shape[1], co), t)
        err = np.abs(t.cpu().numpy() - x).max() / np.abs(x).max()
        assert err < 1e-12, (str(g[f"c{i}_name"]), d.shape[1], err)


@pytest.mark.parametrize("axis", [2, 1, 0])
def test_gpu_compact64_golden(dev, axis):
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.workloads import compact_field
    g = load("compact64.npz")
    f = torch.tensor(compact_field(64, seed=0), device=dev)
    out = torch.empty_like(f)
    core.compact_derivative(Plan(64, COMPACT_COEFFS), f, out, float(g["h"]), axis)
    d = out.cpu().numpy()
    scale = np.abs(g[f"d_axis{axis}"]).max()
    for k, (i, j) in enumerate(g["idx"]):
        sl = [int(i), int(j)]
        sl.insert(axis, slice(None))
        np.testing.assert_allclose(d[tuple(sl)], g[f"d_axis{axis}"][k], rtol=0, atol=1e-12 * scale)
    assert (d * d).sum() == pytest.approx(float(g[f"sumsq_axis{axis}"]), rel=1e-12)


def test_gpu_adi_golden(dev):
    from adi.solver import ADISolver
    g = load("adi66.npz")
    u = torch.tensor(g["u0"], device=dev)
    ADISolver(66, float(g["r"])).step(u, int(g["steps"]))
    assert np.abs(u.cpu().numpy() - g["u"]).max() < 1e-13
