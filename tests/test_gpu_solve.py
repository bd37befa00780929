"""GPU parity of the near-Toeplitz solves against the CPU oracle.

Tolerances (relative, ||x - x_ref||_inf / ||x_ref||_inf): 1e-12 for the FP64
solve, 1e-10 for FP32 + one FP64-residual refinement sweep (the accuracy each
mode is specified to reach).
"""
import numpy as np
import pytest
import torch

from oracle import adi, cr

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
REF_TOL = 1e-10

COEFFS = {
    "compact": (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0),
    "adi": adi.pr_coeffs(0.5),
    "adi_r2": adi.pr_coeffs(2.0),
    "general": (2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7),
}
SIZES = [2, 4, 5, 8, 13, 16, 17, 31, 32, 33, 63, 64, 65, 96, 100, 127, 128, 160, 255, 256,
         257, 300, 512, 513, 777, 1000, 1024, 1025, 1500, 2048]


def rel_err(x, ref):
    x = np.asarray(x)
    ref = np.asarray(ref)
    return np.abs(x - ref).max() / np.abs(ref).max()


def _skip_singular(name, n):
    return name == "compact" and n == 3


@pytest.mark.parametrize("name", sorted(COEFFS))
def test_solve_fp64_all_sizes(dev, name):
    from ntamd import Plan, core
    rng = np.random.default_rng(1)
    for n in SIZES:
        nrhs = 37
        d = rng.standard_normal((nrhs, n))
        want = cr.cr_solve(COEFFS[name], d)
        x = torch.tensor(d, device=dev)
        core.solve(Plan(n, COEFFS[name]), x)
        got = x.cpu().numpy()
        assert rel_err(got, want) < FP64_TOL, (name, n, rel_err(got, want))


@pytest.mark.parametrize("n", [8, 64, 96, 512, 1000, 2048])
@pytest.mark.parametrize("sweeps,tol", [(1, REF_TOL), (2, 1e-13)])
def test_solve_fp32_refine(dev, n, sweeps, tol):
    from ntamd import Plan, core
    d = np.random.default_rng(n).standard_normal((41, n))
    co = COEFFS["compact"]
    want = cr.banded_solve(co, d)
    x = torch.tensor(d, device=dev)
    core.solve(Plan(n, co), x, precision="fp32_refine", sweeps=sweeps)
    assert rel_err(x.cpu().numpy(), want) < tol


def test_solve_fp32_no_sweep_is_single_precision(dev):
    from ntamd import Plan, core
    d = np.random.default_rng(3).standard_normal((16, 512))
    co = COEFFS["compact"]
    want = cr.banded_solve(co, d)
    x = torch.tensor(d, device=dev)
    core.solve(Plan(512, co), x, precision="fp32_refine", sweeps=0)
    e = rel_err(x.cpu().numpy(), want)
    assert 1e-9 < e < 1e-5          # plain FP32 accuracy: refinement is what buys 1e-10


def test_neato_dropin_reference_semantics(dev):
    """NearToeplitzSolver(n, nrhs, coeffs).solve(x_d): in place, from x_d's first
    element regardless of its shape (run-adi.py:65 passes d_x_gpu[1, 0])."""
    from neato import NearToeplitzSolver
    N = 64
    co = adi.reference_coeffs(1.0 / (N - 1), 0.1)
    d = np.random.default_rng(7).standard_normal((N, N))
    x = torch.tensor(d, device=dev)
    solver = NearToeplitzSolver(N, N - 2, co)
    solver.solve(x[1, 0])                          # rows 1..N-2 only
    got = x.cpu().numpy()
    want = d.copy()
    want[1:-1] = cr.cr_solve(co, d[1:-1])
    np.testing.assert_array_equal(got[0], d[0])
    np.testing.assert_array_equal(got[-1], d[-1])
    assert rel_err(got[1:-1], want[1:-1]) < 1e-10   # cond ~4e4 matrix (see test_oracle)
    # over-long requests are refused before launch
    with pytest.raises(ValueError):
        NearToeplitzSolver(N, N, co).solve(x[1, 0])


def test_solve_zero_and_ragged_batches(dev):
    from ntamd import Plan, core
    co = COEFFS["general"]
    p = Plan(100, co)
    x = torch.zeros((0, 100), dtype=torch.float64, device=dev)
    core.solve(p, x)                               # no systems: no launch, no error
    for nrhs in (1, 3, 5, 4097):                   # ragged last tile
        d = np.random.default_rng(nrhs).standard_normal((nrhs, 100))
        x = torch.tensor(d, device=dev)
        core.solve(p, x)
        assert rel_err(x.cpu().numpy(), cr.cr_solve(co, d)) < FP64_TOL


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("shape", [(16, 24, 40), (33, 17, 65), (64, 64, 64), (8, 100, 8)])
def test_solve_axis_strided_lines(dev, axis, shape):
    from ntamd import Plan, core
    co = COEFFS["general"]
    d = np.random.default_rng(sum(shape) + axis).standard_normal(shape)
    n = shape[axis]
    x = torch.tensor(d, device=dev)
    core.solve_axis(Plan(n, co), x, axis)
    want = np.moveaxis(cr.cr_solve(co, np.moveaxis(d, axis, -1)), -1, axis)
    assert rel_err(x.cpu().numpy(), want) < FP64_TOL


def test_solve_axis_long_strided_lines(dev):
    from ntamd import Plan, core
    co = COEFFS["adi"]
    for n in (700, 1024, 2048):
        d = np.random.default_rng(n).standard_normal((n, 40))
        x = torch.tensor(d[None], device=dev)                 # (1, n, 40): y-lines of length n
        core.solve_axis(Plan(n, co), x, 1)
        want = cr.cr_solve(co, d.T).T
        assert rel_err(x[0].cpu().numpy(), want) < FP64_TOL


def test_solve_linearity_large_batch(dev):
    """Size-independent property at a realistic batch: A^{-1}(a u + b v) = a A^{-1}u + b A^{-1}v."""
    from ntamd import Plan, core
    n, nrhs = 512, 65536
    p = Plan(n, COEFFS["compact"])
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    u = torch.randn((nrhs, n), generator=g, dtype=torch.float64, device=dev)
    v = torch.randn((nrhs, n), generator=g, dtype=torch.float64, device=dev)
    w = 0.75 * u - 1.25 * v
    for t in (u, v, w):
        core.solve(p, t)
    diff = (w - (0.75 * u - 1.25 * v)).abs().max().item()
    assert diff < 1e-12 * w.abs().max().item()
    # spot-check lines against the oracle
    idx = [0, 1, 4095, 32768, nrhs - 1]
    uu = torch.randn((nrhs, n), generator=g, dtype=torch.float64, device=dev)
    src = uu[idx].cpu().numpy()
    core.solve(p, uu)
    assert rel_err(uu[idx].cpu().numpy(), cr.cr_solve(COEFFS["compact"], src)) < FP64_TOL
