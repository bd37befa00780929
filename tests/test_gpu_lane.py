"""The block-partitioned lane solve (strided variant 5) against the oracle.

Strided lines of 49 to 512 points whose neighbouring lines are neighbours in
memory (y- and z-lines of C-ordered fields) are solved with one lane per line:
sixteen wavefronts each take one block of every line (m = ceil(n / 16) rows for
the first n - 16 (m - 1) blocks, m - 1 for the others), solve their block by a
Thomas sweep, and the blocks are joined through the tridiagonal reduced system
of the reference's partition method
(parallelcompact/reduced.py:163-178).  The results must match the SciPy banded
oracle to the FP64 tolerance (1e-12 relative) for the plain solve
(neato.py:168-197 on permuted lines) and the fused compact derivative
(compact/rhs.py:21-36 + the solve), on the shapes below:
* z-lines of a (n, 3, 100) field: 300 lines, four groups of 64 and a ragged
  group of 44;
* y-lines of a (2, n, 70) field: two slabs of 70 lines, each a group of 64 and
  a ragged group of 6.
Each is also compared with the cyclic-reduction tile kernel (variant 2) to a
few roundings.
"""
import numpy as np
import pytest
import torch

from oracle import adi, compact, cr

pytestmark = pytest.mark.gpu

COEFFS = {
    "compact": (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0),
    "adi": adi.pr_coeffs(0.5),
    "adi_r2": adi.pr_coeffs(2.0),
    "general": (2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7),
}
SHAPES = {0: lambda n: (n, 3, 100), 1: lambda n: (2, n, 70)}


def rel_err(x, ref):
    return np.abs(np.asarray(x) - ref).max() / np.abs(ref).max()


def _with_variant(n, v, fn):
    from ntamd.autotune import _set, current
    keep, _ = current(1, n, "fp64")
    _set(1, "fp64", n, v)
    try:
        return fn()
    finally:
        _set(1, "fp64", n, keep)


@pytest.mark.parametrize("name", sorted(COEFFS))
@pytest.mark.parametrize("n", [49, 100, 128, 300, 497, 512])
@pytest.mark.parametrize("axis", [0, 1])
def test_lane_solve_axis_in_place(dev, name, n, axis):
    from ntamd import Plan, core
    co = COEFFS[name]
    shape = SHAPES[axis](n)
    d = np.random.default_rng(n + 7 * axis).standard_normal(shape)
    plan = Plan(n, co)

    def run(v):
        x = torch.tensor(d, device=dev)
        _with_variant(n, v, lambda: core.solve_axis(plan, x, axis))
        return x.cpu().numpy()

    got = run(5)
    want = np.moveaxis(cr.cr_solve(co, np.moveaxis(d, axis, -1)), -1, axis)
    assert rel_err(got, want) < 1e-12
    assert np.abs(got - run(2)).max() / np.abs(want).max() < 1e-13


@pytest.mark.parametrize("n", [52, 64, 96, 100, 128, 200, 256, 300, 384, 480, 511, 512])
@pytest.mark.parametrize("axis", [0, 1])
def test_lane_compact_derivative(dev, n, axis):
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.workloads import compact_field
    shape = SHAPES[axis](n)
    f = compact_field(shape, seed=3 * n + axis)
    h = 1.0 / (n - 1)
    want = compact.derivative_banded(f, h, axis)
    fd = torch.tensor(f, device=dev)
    plan = Plan(n, COMPACT_COEFFS)

    def run(v):
        out = torch.full_like(fd, float("nan"))
        _with_variant(n, v, lambda: core.compact_derivative(plan, fd, out, h, axis, "fp64"))
        return out.cpu().numpy()

    got = run(5)
    assert not np.isnan(got).any()
    assert rel_err(got, want) < 1e-12
    assert np.abs(got - run(2)).max() / np.abs(want).max() < 1e-13


def test_lane_many_groups_z_lines(dev):
    """8320 z-lines of 256 points (130 groups of 64): every line against the
    oracle, and the input left untouched."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.workloads import compact_field
    shape = (256, 64, 130)
    f = compact_field(shape, seed=11)
    h = 1.0 / 255
    fd = torch.tensor(f, device=dev)
    out = torch.full_like(fd, float("nan"))
    _with_variant(256, 5, lambda: core.compact_derivative(Plan(256, COMPACT_COEFFS), fd, out, h, 0, "fp64"))
    assert rel_err(out.cpu().numpy(), compact.derivative_banded(f, h, 0)) < 1e-12
    assert np.array_equal(fd.cpu().numpy(), f)


@pytest.mark.parametrize("n", [57, 256, 512])
@pytest.mark.parametrize("shape_of", [lambda n: (n, 1, 1), lambda n: (700, n, 3), lambda n: (n, 1, 65)])
def test_lane_narrow_slabs(dev, n, shape_of):
    """Groups with one line (61..63 lanes idle), 700 slabs of 3 y-lines each,
    and 65 lines (a group of 64 and a group of 1): the in-place solve equals
    the oracle and the idle lanes store nothing (the field around the lines
    is untouched because every element belongs to some line)."""
    from ntamd import Plan, core
    co = COEFFS["general"]
    shape = shape_of(n)
    axis = 1 if shape[1] == n and shape[0] != n else 0
    d = np.random.default_rng(n + len(shape)).standard_normal(shape)
    x = torch.tensor(d, device=dev)
    _with_variant(n, 5, lambda: core.solve_axis(Plan(n, co), x, axis))
    want = np.moveaxis(cr.cr_solve(co, np.moveaxis(d, axis, -1)), -1, axis)
    assert rel_err(x.cpu().numpy(), want) < 1e-12


@pytest.mark.parametrize("sweeps", [1, 2])
@pytest.mark.parametrize("n", [52, 64, 96, 100, 128, 200, 256, 300, 480, 497, 512])
@pytest.mark.parametrize("axis", [0, 1])
def test_lane_fp32_storage_fp32_refine(dev, n, axis, sweeps):
    """FP32-stored fields with the FP32 solve and FP64-residual sweeps: with one
    sweep the FP32 lane solve runs (full and padded blocks; 497 points, 32 rows
    with padding, and two sweeps take the previous FP32 tile kernels).  Both
    equal the oracle's derivative of the same FP32 values to the refinement
    tolerance (1e-10) and each other to 1e-12; the input is left untouched."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.autotune import _set, current
    from ntamd.workloads import compact_field
    shape = SHAPES[axis](n)
    f32 = compact_field(shape, seed=5 * n + axis).astype(np.float32)
    h = 1.0 / (n - 1)
    want = compact.derivative_banded(f32.astype(np.float64), h, axis)
    fd = torch.tensor(f32, device=dev)
    plan = Plan(n, COMPACT_COEFFS)

    def run(v):
        keep, _ = current(1, n, "fp32_refine")
        _set(1, "fp32_refine", n, v)
        try:
            out = torch.full(shape, float("nan"), dtype=torch.float64, device=dev)
            core.compact_derivative(plan, fd, out, h, axis, "fp32_refine", sweeps)
            return out.cpu().numpy()
        finally:
            _set(1, "fp32_refine", n, keep)

    got = run(5)
    assert not np.isnan(got).any()
    assert rel_err(got, want) < 1e-10
    assert np.abs(got - run(2)).max() / np.abs(want).max() < 1e-12
    assert np.array_equal(fd.cpu().numpy(), f32)
