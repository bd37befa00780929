"""Compact (Pade) 4th-order first derivative on 3-D grids, restated on the CPU.

TEST INFRASTRUCTURE ONLY (see the package docstring).

The derivative solves ``1/4 f'_{i-1} + f'_i + 1/4 f'_{i+1} = 3/(4h) (f_{i+1} - f_{i-1})``
in the interior with the 3rd-order one-sided closures ``f'_0 + 2 f'_1 =
(-5 f_0 + 4 f_1 + f_2)/(2h)`` and its mirror at the last point: matrix 7-tuple
``(1, 2, 1/4, 1, 1/4, 2, 1)`` (reference ``experiments/run-compact.py:48``) and
RHS stencil of ``compact/rhs.py:21-36``.
"""
import numpy as np

from .cr import cr_solve, banded_solve

__all__ = ["COMPACT_COEFFS", "compact_rhs", "permute", "derivative",
           "derivative_banded", "axis_to_last", "last_to_axis"]

COMPACT_COEFFS = (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0)


def compact_rhs(f, dx):
    """RHS of the compact scheme along the last axis (``compact/rhs.py:21-36``)."""
    f = np.asarray(f, dtype=np.float64)
    r = np.empty_like(f)
    r[..., 1:-1] = (3.0 / (4 * dx)) * (f[..., 2:] - f[..., :-2])
    r[..., 0] = (1.0 / (2 * dx)) * (-5 * f[..., 0] + 4 * f[..., 1] + f[..., 2])
    r[..., -1] = -(1.0 / (2 * dx)) * (-5 * f[..., -1] + 4 * f[..., -2] + f[..., -3])
    return r


def permute(a, permutation):
    """Axis permutation of a 3-D array as ``compact/permute.py:15-19,45-46``.

    The reference kernel reads input strides reordered by ``permutation`` and
    writes a C-contiguous output, i.e. ``out = a.transpose(permutation)``.
    """
    return np.ascontiguousarray(np.transpose(np.asarray(a), permutation))


# the permutations the reference uses to make y and z lines contiguous
# (experiments/run-compact.py:70,73,83,86)
_TO_LAST = {2: None, 1: (0, 2, 1), 0: (1, 2, 0)}
_FROM_LAST = {2: None, 1: (0, 2, 1), 0: (2, 0, 1)}


def axis_to_last(f, axis):
    p = _TO_LAST[axis]
    return np.asarray(f) if p is None else permute(f, p)


def last_to_axis(g, axis):
    p = _FROM_LAST[axis]
    return np.asarray(g) if p is None else permute(g, p)


def derivative(f, dx, axis=2, solver=cr_solve):
    """Compact derivative of a (nz, ny, nx) array along ``axis`` (2 = x, 1 = y, 0 = z).

    Mirrors the reference call sequence (``experiments/run-compact.py:57-90``):
    permute the axis last, RHS, near-Toeplitz solve, permute back.
    """
    g = axis_to_last(f, axis)
    x = solver(COMPACT_COEFFS, compact_rhs(g, dx))
    return last_to_axis(x, axis)


def derivative_banded(f, dx, axis=2):
    """The same derivative with the SciPy banded solve (the CPU baseline path)."""
    return derivative(f, dx, axis, solver=banded_solve)
