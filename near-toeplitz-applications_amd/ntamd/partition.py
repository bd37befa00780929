"""One-GPU partitioned solve (the reference's distributed partition method on one device).

Reference pieces and their counterparts here:
  parallelcompact/misc.py:21-46    solve_secondary_systems  -> plan data (host FP64, once)
  parallelcompact/rhs.py:8-73      compute_RHS (ghosted)    -> RHS in line coordinates, in-kernel
  neato.py solve per rank                                   -> per-subdomain register CR
  parallelcompact/reduced.py:135-196 get_params             -> reduced 2P system, in-kernel
  parallelcompact/sum.py:7-51      sum_solutions            -> correction, in-kernel
One launch handles a batch of lines; no host round trip, no per-call allocation.
Contiguous x-lines take one line per workgroup; y- and z-lines (and the ADI
y-implicit half step) are staged in tiles of adjacent lines, so the
reference's permute before and after the partitioned solve
(experiments/run-parallel-compact.py:92-164) never happens.
"""
import ctypes

import numpy as np

from . import _lib
from .core import device_ptr, stream_handle, storage_elems_from

__all__ = ["PartitionPlan", "PartitionedADISolver"]

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

    def _axis(self, producer, src, dst, shape, axis, p1=0.0):
        nz, ny, nx = (int(v) for v in shape)
        _lib.call("nt_part_solve_axis", self._h, int(producer), src, dst, ctypes.c_int64(nz),
                  ctypes.c_int64(ny), ctypes.c_int64(nx), int(axis), float(p1), stream_handle())

    def solve_axis(self, x, axis):
        """Solve every line along ``axis`` (0 z, 1 y, 2 x) of a 3-D field in place."""
        if x.dim() != 3 or x.shape[axis] != self.n:
            raise _lib.NtInvalidArgument(_lib.NT_EINVAL, "x must be 3-D with extent n along axis")
        p = device_ptr(x, "x")
        self._axis(PRODUCER_SOLVE, p, p, x.shape, axis)
        return x

    def compact_derivative(self, f, out, dx, axis=2):
        """Compact derivative along ``axis`` (0 z, 1 y, 2 x) of a 3-D field; f, out same shape."""
        if f.dim() != 3 or tuple(f.shape) != tuple(out.shape) or f.shape[axis] != self.n:
            raise _lib.NtInvalidArgument(_lib.NT_EINVAL, "f/out must be 3-D, same shape, extent n along axis")
        self._axis(PRODUCER_COMPACT, device_ptr(f, "f"), device_ptr(out, "out"), f.shape, axis, dx)
        return out

    def adi_half_y(self, work, u, r):
        """y-implicit PR half step work -> u (columns 1..n-2) of an n x n field."""
        if tuple(u.shape) != (self.n, self.n) or tuple(work.shape) != (self.n, self.n):
            raise _lib.NtInvalidArgument(_lib.NT_EINVAL, "u and work must be n x n")
        self._axis(PRODUCER_ADI, device_ptr(work, "work"), device_ptr(u, "u"), (1, self.n, self.n), 1, r)
        return u

    def adi_half_x(self, u, work, r, tiles=True):
        """x-implicit PR half step u -> work (rows 1..n-2) of an n x n field:
        in tiles of rows (``tiles=True``) or one row per workgroup."""
        if tuple(u.shape) != (self.n, self.n) or tuple(work.shape) != (self.n, self.n):
            raise _lib.NtInvalidArgument(_lib.NT_EINVAL, "u and work must be n x n")
        if tiles:
            self._axis(PRODUCER_ADI, device_ptr(u, "u"), device_ptr(work, "work"), (1, self.n, self.n), 2, r)
        else:
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


class PartitionedADISolver:
    """Peaceman-Rachford ADI steps with both half steps partitioned: every line
    of the N x N grid is cut into P subdomains, solved as a batch, and joined
    through the reduced 2P-unknown system (the one-GPU share of the reference's
    distributed solve).  A step is two launches, the x-implicit half step
    u -> work (rows) and the y-implicit one work -> u (columns); ``capture``
    records k steps into a hipGraph (``torch.cuda.CUDAGraph``)."""

    def __init__(self, N, r, P=16):
        from . import adi_coeffs
        self.N, self.r, self.P = int(N), float(r), int(P)
        self.plan = PartitionPlan(self.N, self.P, adi_coeffs(self.r))
        self.work = None
        self._graph = None

    def _work_for(self, u):
        import torch
        if self.work is None or self.work.device != u.device:
            self.work = torch.empty((self.N, self.N), dtype=torch.float64, device=u.device)
        return self.work

    def prepare(self, u):
        """Copy u (with its Dirichlet boundary) into the work buffer."""
        from .core import adi_prepare
        adi_prepare(u, self._work_for(u))

    def _steps(self, u, nsteps):
        for _ in range(int(nsteps)):
            self.plan.adi_half_x(u, self.work, self.r)
            self.plan.adi_half_y(self.work, u, self.r)

    def step(self, u, nsteps=1, prepare=True):
        """Advance u (N x N float64, boundary = Dirichlet data) by nsteps, in place."""
        if prepare:
            self.prepare(u)
        self._steps(u, nsteps)
        return u

    def capture(self, u, nsteps):
        """Record nsteps steps on u into a graph (u and the work buffer stay bound)."""
        import torch
        self.prepare(u)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self._steps(u, nsteps)
        torch.cuda.current_stream().wait_stream(s)
        self._graph, self._graph_u = g, u
        return g

    def replay(self):
        if self._graph is None:
            raise RuntimeError("capture() first")
        self._graph.replay()
        return self._graph_u
