"""Drop-in for the reference's ``compact/permute.py``: ``permute(a_d, b_d, permutation)``.

Reference: ``compact/permute.py:32-52``: ``b_d = a_d.transpose(permutation)``
for a C-contiguous 3-D float64 array.  Any shape is accepted (the reference
needs cubes with extents divisible by 8).
"""
from ntamd import core

__all__ = ["permute"]


def permute(a_d, b_d, permutation):
    core.permute(a_d, b_d, permutation)
