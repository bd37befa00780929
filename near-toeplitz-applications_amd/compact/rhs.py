"""Drop-in for the reference's ``compact/rhs.py``: ``compute_rhs(f, rhs, dx)``.

Reference: ``compact/rhs.py:8-53``.  Writes the 4th-order compact RHS along the
last (x) axis of the (nz, ny, nx) float64 field ``f`` into ``rhs``.  Unlike the
reference it accepts any extents (not multiples of 8) and checks shapes.
"""
from ntamd import core

__all__ = ["compute_rhs"]


def compute_rhs(f, rhs, dx):
    core.compact_rhs(f, rhs, dx)
