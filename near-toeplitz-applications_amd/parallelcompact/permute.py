"""Drop-in for ``parallelcompact/permute.py:1-52`` (a copy of ``compact/permute.py``):
``permute(a_d, b_d, permutation)`` writes ``a_d.transpose(permutation)`` into ``b_d``."""
from compact.permute import permute

__all__ = ["permute"]
