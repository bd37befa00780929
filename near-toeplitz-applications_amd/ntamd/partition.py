"""One-GPU partitioned solve (the reference's distributed partition method on one device).

Reference pieces and their counterparts here:
  parallelcompact/misc.py:21-46    solve_secondary_systems  -> plan data (host FP64, once)
  parallelcompact/rhs.py:8-73      compute_RHS (ghosted)    -> RHS in line coordinates, in-kernel
  neato.py solve per rank                                   -> per-subdomain register CR
  parallelcompact/reduced.py:135-196 get_params             -> reduced 2P system, in-kernel
  parallelcompact/sum.py:7-51      sum_solutions            -> correction, in-kernel
One launch handles a batch of lines; no host round trip, no per-call allocation.
"""
import ctypes

import numpy as np

from . import _lib
from .core import device_ptr, stream_handle, storage_elems_from

__all__ = ["PartitionPlan"]

PRODUCER_SOLVE = 0
PRODUCER_COMPACT = 1
PRODUCER_ADI = 2


class PartitionPlan:
    """Lines of n = P*m points cut into P <= 16 subdomains, matrix ``coeffs``."""

    def __init__(self, n, P, coeffs):
        self.n, self.P = int(n), int(P)
        self.m = self.n // self.P
        self.coeffs = tuple(float(c) for c in coeffs)
        lib = _lib.load()
        h = ctypes.c_void_p()
        arr = (ctypes.c_double * 7)(*self.coeffs)
        _lib.check(lib.nt_part_plan_create(ctypes.c_int64(self.n), self.P, arr, ctypes.byref(h)))
        self._h, self._lib = h, lib

    def secondary(self, p):
        """(xU, xL) of subdomain p: solutions of the subdomain block with RHS
        -a e_0 and -c e_{m-1} (parallelcompact/misc.py:21-46)."""
        xu = (ctypes.c_double * self.m)()
        xl = (ctypes.c_double * self.m)()
        _lib.check(self._lib.nt_part_plan_secondary(self._h, int(p), xu, xl, self.m))
        return np.frombuffer(xu).copy(), np.frombuffer(xl).copy()

    def _run(self, producer, src, dst, lines, p1=0.0):
        _lib.call("nt_part_solve_lines", self._h, int(producer), src, dst, ctypes.c_int64(lines),
                  float(p1), 0.0, stream_handle())

    def solve(self, x, lines=None):
        """Solve ``lines`` contiguous systems of length n in place (default: all of x)."""
        ptr = device_ptr(x, "x", contiguous=lines is None)
        if lines is None:
            if x.numel() % self.n:
                raise _lib.NtInvalidArgument(_lib.NT_EINVAL, "x must hold whole lines of length n")
            lines = x.numel() // self.n
        if lines * self.n > storage_elems_from(x):
            raise _lib.NtInvalidArgument(_lib.NT_EINVAL, "x's storage is too small for the lines")
        self._run(PRODUCER_SOLVE, ptr, ptr, lines)
        return x

    def compact_derivative_x(self, f, out, dx):
        """Compact d/dx along the contiguous (last) axis; f, out same shape."""
        if f.shape != out.shape or f.shape[-1] != self.n:
            raise _lib.NtInvalidArgument(_lib.NT_EINVAL, "f/out shape mismatch or last axis != n")
        self._run(PRODUCER_COMPACT, device_ptr(f, "f"), device_ptr(out, "out"),
                  f.numel() // self.n, dx)
        return out

    def adi_half_x(self, u, work, r):
        """x-implicit PR half step u -> work (rows 1..n-2) of an n x n field."""
        if tuple(u.shape) != (self.n, self.n) or tuple(work.shape) != (self.n, self.n):
            raise _lib.NtInvalidArgument(_lib.NT_EINVAL, "u and work must be n x n")
        self._run(PRODUCER_ADI, device_ptr(u, "u"), device_ptr(work, "work"), self.n, r)
        return work

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.nt_part_plan_destroy(h)
            except Exception:
                pass
            self._h = None
