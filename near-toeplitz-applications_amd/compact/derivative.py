"""Fused compact derivative: the reference's d/dx, d/dy, d/dz call sequences in one launch.

Reference sequences (``experiments/run-compact.py:57-90``):
  d/dx: compute_rhs -> solve                                (2 HBM passes)
  d/dy: permute(0,2,1) -> compute_rhs -> solve -> permute    (4 passes)
  d/dz: permute(1,2,0) -> compute_rhs -> solve -> permute    (4 passes)
Here each is one kernel that reads f once and writes the derivative once, for
lines along any axis, FP64 or FP32-stored fields, FP64 or FP32+refinement solve.
"""
from ntamd import core, COMPACT_COEFFS

__all__ = ["CompactDerivative", "derivative"]


class CompactDerivative:
    """Derivatives of (nz, ny, nx) fields with grid spacing ``(dz, dy, dx)``."""

    def __init__(self, shape, spacing, precision="fp64", sweeps=1):
        self.shape = tuple(int(s) for s in shape)
        if len(self.shape) != 3:
            raise ValueError("shape must be (nz, ny, nx)")
        if isinstance(spacing, (int, float)):
            spacing = (float(spacing),) * 3
        self.spacing = tuple(float(h) for h in spacing)
        self.precision = precision
        self.sweeps = int(sweeps)
        core.precision_code(precision)
        self._plans = {}

    def plan(self, axis):
        n = self.shape[axis]
        p = self._plans.get(n)
        if p is None:
            p = self._plans[n] = core.Plan(n, COMPACT_COEFFS)
        return p

    def __call__(self, f, out, axis):
        """out = d f / d(axis); axis 0 = z, 1 = y, 2 = x.  Asynchronous."""
        if tuple(f.shape) != self.shape:
            raise ValueError(f"f has shape {tuple(f.shape)}, expected {self.shape}")
        return core.compact_derivative(self.plan(axis), f, out, self.spacing[axis], axis,
                                       self.precision, self.sweeps)


def derivative(f, out, dx, axis=2, precision="fp64", sweeps=1):
    """One-shot fused compact derivative (builds the plan on the fly)."""
    n = int(f.shape[axis])
    plan = core.Plan(n, COMPACT_COEFFS)
    return core.compact_derivative(plan, f, out, dx, axis, precision, sweeps)
