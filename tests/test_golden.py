"""The oracle (and the workload generator) against the committed golden fixtures."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import adi, compact, cr, partition


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def test_cr_lines_golden():
    g = load("cr_lines.npz")
    for i in range(int(g["ncases"])):
        co, d, x = g[f"c{i}_coeffs"], g[f"c{i}_d"], g[f"c{i}_x"]
        got = cr.cr_solve(co, d)
        assert np.abs(got - x).max() / np.abs(x).max() < 1e-12, str(g[f"c{i}_name"])


def test_compact64_generator_and_oracle_golden():
    from ntamd.workloads import compact_field
    g = load("compact64.npz")
    f = compact_field(64, seed=0)
    h = float(g["h"])
    for axis in (2, 1, 0):
        d = compact.derivative(f, h, axis)
        for k, (i, j) in enumerate(g["idx"]):
            sl = [int(i), int(j)]
            sl.insert(axis, slice(None))
            np.testing.assert_array_equal(f[tuple(sl)], g[f"f_axis{axis}"][k])
            np.testing.assert_allclose(d[tuple(sl)], g[f"d_axis{axis}"][k], rtol=0,
                                       atol=1e-12 * np.abs(g[f"d_axis{axis}"]).max())
        assert d.sum() == pytest.approx(float(g[f"sum_axis{axis}"]), rel=1e-9, abs=1e-9)
        assert (d * d).sum() == pytest.approx(float(g[f"sumsq_axis{axis}"]), rel=1e-12)


def test_adi_golden():
    from ntamd.workloads import gaussian_ic
    g = load("adi66.npz")
    np.testing.assert_array_equal(gaussian_ic(66, seed=1), g["u0"])
    np.testing.assert_array_equal(adi.gaussian_ic(66, seed=1), g["u0"])
    u = adi.pr_run(g["u0"], float(g["r"]), int(g["steps"]))
    assert np.abs(u - g["u"]).max() < 1e-13


def test_partition_golden():
    g = load("partition2048.npz")
    x = partition.partitioned_solve(tuple(g["coeffs"]), g["d"], int(g["P"]))
    assert np.abs(x - g["x"]).max() / np.abs(g["x"]).max() < 1e-12
