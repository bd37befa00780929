"""Drop-in for ``parallelcompact/rhs.py:8-73``: ``compute_RHS(f_local, x, dx, shape, mx, npx)``.

``x`` (nz, ny, nx) receives the 4th-order compact right-hand side along x,
``(3/(4 dx)) (f[i+1] - f[i-1])``, evaluated on the width-1 ghosted subdomain
array ``f_local`` (nz+2, ny+2, nx+2); the one-sided closure
``(1/(2 dx)) (-5 f0 + 4 f1 + f2)`` replaces the first point only on rank
``mx == 0`` and its mirror the last point only on rank ``mx == npx - 1``
(``rhs.py:42-52``).  One launch (``nt_part_rhs_ghosted``), asynchronous on the
current stream.  Extents need not be multiples of 8 (the reference's grid of
8x8x8 blocks needs them), but nx >= 3.
"""
from ntamd import core
from ntamd._lib import NtInvalidArgument, NT_EINVAL

__all__ = ["compute_RHS"]


def compute_RHS(f_local, x, dx, shape, mx, npx):
    if tuple(int(s) for s in shape) != tuple(x.shape):
        raise NtInvalidArgument(NT_EINVAL, f"shape {tuple(shape)} != x.shape {tuple(x.shape)}")
    core.part_rhs_ghosted(f_local, x, dx, mx, npx)
