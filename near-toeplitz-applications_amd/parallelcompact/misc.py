"""Drop-in for ``parallelcompact/misc.py`` (reference lines 5-46) on one GPU.

``solve_secondary_systems(nx, rank, size, coeffs)`` returns the device arrays
(xU, xL) of subdomain ``rank`` of ``size``: solutions of the subdomain block
with right-hand sides ``-a[0] e_0`` and ``-c[-1] e_last``.  They come from the
partition plan's host FP64 precompute (the reference solves them with SciPy
on every call) and are copied to the current device, as the reference's
``gpuarray.to_gpu`` does.
"""
import torch

from ntamd.partition import PartitionPlan

__all__ = ["solve_secondary_systems"]


def solve_secondary_systems(nx, rank, size, coeffs):
    plan = PartitionPlan(int(nx) * int(size), int(size), coeffs)
    xu, xl = plan.secondary(int(rank))
    dev = torch.device("cuda")
    return torch.tensor(xu, device=dev), torch.tensor(xl, device=dev)
