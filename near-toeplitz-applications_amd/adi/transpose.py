"""Drop-in for the reference's ``adi/transpose.py``: ``transpose(a, b, N)`` sets b = a^T.

Reference: ``adi/transpose.py:40-50`` (N x N float64; any N here, not only
multiples of 32).
"""
from ntamd import core

__all__ = ["transpose"]


def transpose(a, b, N):
    core.transpose(a, b, N)
