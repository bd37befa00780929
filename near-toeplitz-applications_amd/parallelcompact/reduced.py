"""Drop-in for ``parallelcompact/reduced.py:135-196``: ``get_params(line_da, xU_d, xL_d, xR_d)``.

The reference gathers the end values of every rank's secondary solutions and
the negated faces of its subdomain solution ``xR`` to the line's root rank,
builds the reduced system of 2P unknowns (``reduced.py:163-178``), solves it
per line with a serial Thomas sweep (``reduced.py:11-49``) and scatters
``(alpha, beta)`` back.  On one device the P subdomains of a line are already
together, so the whole exchange is one launch (``nt_part_params``):

* ``line_da``: the line's decomposition (``da.DA.get_line_DA``, or any object
  with ``size`` = P);
* ``xU_d``, ``xL_d``: (P, nx) secondary solutions, row p for subdomain p
  (``misc.solve_secondary_systems(nx, p, P, coeffs)``);
* ``xR_d``: (P, nz, ny, nx) subdomain solutions, or (..., P, nz, ny, nx) for
  several lines of subdomains (as ``da.DA`` stacks them); one launch per line.

Returns ``alpha_d, beta_d`` shaped like ``xR_d`` without its last axis:
``alpha_d[..., p, :, :]`` is subdomain p's ``alpha``.
"""
import torch

from ntamd import core
from ntamd._lib import NtInvalidArgument, NT_EINVAL

__all__ = ["get_params"]


def get_params(line_da, xU_d, xL_d, xR_d):
    P = int(line_da.size)
    if xR_d.dim() < 4 or int(xR_d.shape[-4]) != P:
        raise NtInvalidArgument(NT_EINVAL, f"xR_d must be (..., P={P}, nz, ny, nx), got {tuple(xR_d.shape)}")
    alpha = torch.empty(xR_d.shape[:-1], dtype=torch.float64, device=xR_d.device)
    beta = torch.empty_like(alpha)
    lines = xR_d.reshape(-1, *xR_d.shape[-4:])
    al = alpha.view(-1, *alpha.shape[-3:])
    be = beta.view(-1, *beta.shape[-3:])
    for g in range(lines.shape[0]):
        core.part_params(xU_d, xL_d, lines[g], al[g], be[g])
    return alpha, beta
