"""Drop-in for ``parallelcompact/sum.py:7-51``: ``sum_solutions(x_UH, x_LH, x_R, alpha, beta, shape)``.

In place, ``x_R[z, y, x] += alpha[z, y] x_UH[x] + beta[z, y] x_LH[x]``
(``sum.py:30``) for one subdomain: ``x_R`` (nz, ny, nx), ``x_UH``/``x_LH``
(nx,), ``alpha``/``beta`` (nz, ny) -- for example ``alpha[p]`` of the arrays
``reduced.get_params`` returns.  One launch (``nt_part_sum``).
"""
from ntamd import core
from ntamd._lib import NtInvalidArgument, NT_EINVAL

__all__ = ["sum_solutions"]


def sum_solutions(x_UH, x_LH, x_R, alpha, beta, shape):
    if tuple(int(s) for s in shape) != tuple(x_R.shape):
        raise NtInvalidArgument(NT_EINVAL, f"shape {tuple(shape)} != x_R.shape {tuple(x_R.shape)}")
    core.part_sum(x_UH, x_LH, x_R, alpha, beta)
