"""GPU parity of the one-launch partitioned solve against the monolithic oracle solve."""
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN
from oracle import adi, compact, cr, partition

pytestmark = pytest.mark.gpu

COMPACT = (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0)


def rel_err(x, ref):
    return np.abs(np.asarray(x) - np.asarray(ref)).max() / np.abs(np.asarray(ref)).max()


@pytest.mark.parametrize("coeffs", [COMPACT, adi.pr_coeffs(0.5), (2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7)])
@pytest.mark.parametrize("n,P", [(2048, 16), (512, 4), (96, 3), (64, 1), (100, 2), (1024, 8),
                                 (2000, 16), (512, 2), (1536, 8), (2048, 4), (1000, 2), (2048, 2),
                                 (1800, 3), (2048, 1), (64, 16), (34, 2), (1812, 12)])
def test_partitioned_equals_monolithic(dev, coeffs, n, P):
    """Subdomains of m = n / P points on segments of 16, 32 or 64 lanes (four, two
    or one per wavefront; 4..2048 points, R = 1..32), idle segments included."""
    from ntamd.partition import PartitionPlan
    d = np.random.default_rng(n + P).standard_normal((33, n))
    x = torch.tensor(d, device=dev)
    PartitionPlan(n, P, coeffs).solve(x)
    assert rel_err(x.cpu().numpy(), cr.banded_solve(coeffs, d)) < 1e-12
    # and the oracle's own partitioned restatement agrees
    assert rel_err(x.cpu().numpy(), partition.partitioned_solve(coeffs, d, P)) < 1e-12


def test_partition_golden(dev):
    from ntamd.partition import PartitionPlan
    g = np.load(os.path.join(GOLDEN, "partition2048.npz"), allow_pickle=False)
    x = torch.tensor(g["d"], device=dev)
    PartitionPlan(2048, int(g["P"]), tuple(g["coeffs"])).solve(x)
    assert rel_err(x.cpu().numpy(), g["x"]) < 1e-12


def test_partitioned_compact_known_answer(dev):
    """run-parallel-compact.py:89-90: f = x*y*z gives d/dx = y*z, here with the
    x-lines cut into 4 subdomains on one GPU."""
    from ntamd.partition import PartitionPlan
    N = 256
    xs = np.linspace(0, 1, N)
    ys = np.linspace(0, 1, 16)
    Z, Y, X = np.meshgrid(ys, ys, xs, indexing="ij")
    f = torch.tensor(X * Y * Z, device=dev)
    out = torch.empty_like(f)
    PartitionPlan(N, 4, COMPACT).compact_derivative_x(f, out, xs[1])
    np.testing.assert_allclose(out.cpu().numpy(), Y * Z, rtol=1e-7, atol=1e-12)


def test_partitioned_compact_matches_oracle(dev):
    from ntamd.partition import PartitionPlan
    f = np.random.default_rng(1).standard_normal((8, 8, 512))
    out = torch.empty((8, 8, 512), dtype=torch.float64, device=dev)
    PartitionPlan(512, 8, COMPACT).compact_derivative_x(torch.tensor(f, device=dev), out, 0.01)
    assert rel_err(out.cpu().numpy(), compact.derivative(f, 0.01, 2)) < 1e-12


def test_partitioned_adi_half_step(dev):
    from ntamd.partition import PartitionPlan
    N, r = 2048, 0.5
    u0 = adi.gaussian_ic(N, seed=3)
    u = torch.tensor(u0, device=dev)
    work = u.clone()
    PartitionPlan(N, 16, adi.pr_coeffs(r)).adi_half_x(u, work, r)
    want = u0.copy()
    d = u0[1:-1].copy()
    d[:, 1:-1] = 0.5 * r * (u0[:-2, 1:-1] + u0[2:, 1:-1]) + (1 - r) * u0[1:-1, 1:-1]
    want[1:-1] = cr.banded_solve(adi.pr_coeffs(r), d)
    assert rel_err(work.cpu().numpy(), want) < 1e-12
