"""Drop-in for the reference's ``adi/rhs.py``: ``compute_rhs(u, d, dx, dt, N)``.

Reference: ``adi/rhs.py:35-55``: on the interior of the N x N array,
``d[j,i] = bottom*u[j-1,i] + center*u[j,i] + top*u[j+1,i]`` with
``bottom = top = -1/dx^2`` and ``center = 2(-1/dt + 1/dx^2)``; boundary entries
of ``d`` are left untouched.
"""
from ntamd import core

__all__ = ["compute_rhs"]


def compute_rhs(u, d, dx, dt, N):
    core.adi_rhs(u, d, dx, dt, N)
