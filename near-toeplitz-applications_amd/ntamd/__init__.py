"""MI355X-native near-Toeplitz tridiagonal solver library (HIP kernels for gfx950).

The reference-compatible modules (``neato``, ``compact``, ``adi``,
``parallelcompact``) live beside this package; this package holds the ctypes
boundary (``_lib``), the validating tensor layer (``core``) and the fused
entry points built on it.
"""
from .core import (Plan, solve, solve_axis, compact_derivative, compact_rhs,  # noqa: F401
                   permute, adi_rhs, transpose, adi_prepare, adi_steps, precision_code,
                   NtError, NtInvalidArgument, NtLibraryError, NtSingular, MAX_N)
from ._lib import LIB_PATH, load  # noqa: F401

COMPACT_COEFFS = (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0)


def adi_coeffs(r):
    """Peaceman-Rachford line matrix (-r/2, 1+r, -r/2) with identity (Dirichlet) ends."""
    r = float(r)
    return (1.0, 0.0, -0.5 * r, 1.0 + r, -0.5 * r, 0.0, 1.0)
