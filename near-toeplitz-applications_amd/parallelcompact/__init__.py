"""Drop-ins for the reference's ``parallelcompact`` package on one device.

The reference runs one MPI rank per GPU (``experiments/run-parallel-compact.py``);
here every subdomain of the decomposition lives on the same device, stacked
block-major, and the steps a rank runs (``rhs.compute_RHS``,
``misc.solve_secondary_systems``, ``reduced.get_params``,
``sum.sum_solutions``, ``permute.permute``) keep the reference's names and
arguments.  ``da.DA`` replaces the MPI domain decomposition.
"""
