"""Fused Peaceman-Rachford ADI stepping with optional hipGraph capture.

Reference loop (``experiments/run-adi.py:60-71``), six launches per step:
compute_rhs -> solve rows -> transpose, twice.  Here a step is two launches: an
x-implicit half step (explicit stencil along y fused into the load, solve of
the interior rows) writing a work buffer, and a y-implicit half step solving
the interior columns in place of a transpose.  Lines are (-r/2, 1+r, -r/2)
with identity (Dirichlet) end rows, r = dt/h^2 (the reference's rows times
-dt/2).  ``capture(k)`` records k steps once into a hipGraph (via
``torch.cuda.CUDAGraph``, which is a hipGraph on ROCm) and ``replay()``
relaunches them with one call.
"""
import torch

from ntamd import core, adi_coeffs

__all__ = ["ADISolver"]


class ADISolver:

    def __init__(self, N, r, precision="fp64", sweeps=1):
        self.N = int(N)
        self.r = float(r)
        self.precision = precision
        self.sweeps = int(sweeps)
        self.plan = core.Plan(self.N, adi_coeffs(self.r))
        self.work = None
        self._graph = None
        self._graph_u = None

    def _work_for(self, u):
        if self.work is None or self.work.device != u.device:
            self.work = torch.empty((self.N, self.N), dtype=torch.float64, device=u.device)
        return self.work

    def prepare(self, u):
        """Copy u (including its Dirichlet boundary) into the work buffer."""
        core.adi_prepare(u, self._work_for(u))

    def step(self, u, nsteps=1, prepare=True):
        """Advance u (N x N float64, boundary = Dirichlet data) by nsteps, in place."""
        if prepare:
            self.prepare(u)
        core.adi_steps(self.plan, u, self.work, self.r, nsteps, self.precision, self.sweeps)
        return u

    def capture(self, u, nsteps):
        """Record nsteps steps on u into a graph (u and the work buffer stay bound)."""
        self.prepare(u)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                core.adi_steps(self.plan, u, self.work, self.r, nsteps, self.precision, self.sweeps)
        torch.cuda.current_stream().wait_stream(s)
        self._graph, self._graph_u, self._graph_steps = g, u, nsteps
        return g

    def replay(self):
        if self._graph is None:
            raise RuntimeError("capture() first")
        self._graph.replay()
        return self._graph_u
