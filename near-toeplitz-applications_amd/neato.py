"""Drop-in for the reference's ``neato.py``: ``NearToeplitzSolver(n, nrhs, coeffs).solve(x_d)``.

Reference: ``neato.py:129-197``.  ``coeffs = (b1, c1, ai, bi, ci, an, bn)``;
``solve`` works in place on ``nrhs`` systems of length ``n`` stored contiguously
from ``x_d``'s first element, whatever ``x_d``'s shape (``run-adi.py:65`` passes
the element view ``d_x_gpu[1, 0]``).  ``x_d`` is a float64 torch tensor on the
GPU.  Differences from the reference, all additive: any ``2 <= n <= 2048``
(not only powers of two), the remaining storage is checked to hold the
systems, and ``precision="fp32_refine"`` selects the FP32 solve with FP64
residual refinement.
"""
from ntamd import core

__all__ = ["NearToeplitzSolver"]


class NearToeplitzSolver:

    def __init__(self, n, nrhs, coeffs, precision="fp64", sweeps=1):
        """
        n: size of each tridiagonal system
        nrhs: number of right-hand sides solved per call
        coeffs: (b1, c1, ai, bi, ci, an, bn)
        """
        self.n = int(n)
        self.nrhs = int(nrhs)
        self.coeffs = tuple(float(c) for c in coeffs)
        if self.nrhs < 0:
            raise ValueError("nrhs must be >= 0")
        if not 2 <= self.n <= core.MAX_N:
            raise ValueError(f"n={self.n} outside the supported range [2, {core.MAX_N}]")
        self.precision = precision
        self.sweeps = int(sweeps)
        core.precision_code(precision)
        self.plan = core.Plan(self.n, self.coeffs)

    def solve(self, x_d):
        """Solve the systems stored from ``x_d`` in place (asynchronous)."""
        core.solve(self.plan, x_d, nrhs=self.nrhs, precision=self.precision, sweeps=self.sweeps)
