"""Peaceman-Rachford ADI heat step, restated on the CPU.

TEST INFRASTRUCTURE ONLY (see the package docstring).

Two forms of the same method:

* ``reference_run`` follows ``experiments/run-adi.py:35-71`` literally: the
  unscaled line matrix ``(1, 0, 1/h^2, -2(1/dt + 1/h^2), 1/h^2, 0, 1)``, the
  explicit stencil ``bottom = top = -1/h^2``, ``center = 2(1/h^2 - 1/dt)`` on the
  interior only (``adi/rhs.py:21-24,46-49``), solves of rows 1..N-2, and a
  transpose after each half step.  Dirichlet values ride in the never-written
  entries of the two RHS buffers (``run-adi.py:43-47``).
* ``pr_step`` is the same step written with r = dt/h^2 and no transposes:
  lines ``(-r/2, 1+r, -r/2)`` with identity boundary rows, explicit weights
  ``(r/2, 1-r, r/2)`` (the reference's rows multiplied by -dt/2).
"""
import numpy as np

from .cr import cr_solve

__all__ = ["reference_coeffs", "pr_coeffs", "adi_rhs", "transpose",
           "reference_run", "pr_step", "pr_run", "eigen_amplification",
           "gaussian_ic"]


def reference_coeffs(dx, dt):
    """Line matrix of ``experiments/run-adi.py:50-53``."""
    return (1.0, 0.0, 1.0 / (dx * dx), -2.0 * (1.0 / dt + 1.0 / (dx * dx)),
            1.0 / (dx * dx), 0.0, 1.0)


def pr_coeffs(r):
    """The same matrix scaled by -dt/2: ``(-r/2, 1+r, -r/2)``, identity ends."""
    return (1.0, 0.0, -0.5 * r, 1.0 + r, -0.5 * r, 0.0, 1.0)


def adi_rhs(u, d, dx, dt):
    """``adi/rhs.py:9-55``: d[j,i] = bottom u[j-1,i] + center u[j,i] + top u[j+1,i]
    on the interior; boundary entries of ``d`` are left as they are."""
    bottom = top = -1.0 / (dx * dx)
    center = 2.0 * (-1.0 / dt + 1.0 / (dx * dx))
    d[1:-1, 1:-1] = (bottom * u[:-2, 1:-1] + center * u[1:-1, 1:-1]
                     + top * u[2:, 1:-1])
    return d


def transpose(a):
    """``adi/transpose.py``: b = a^T."""
    return np.ascontiguousarray(np.asarray(a).T)


def reference_run(u0, dx, dt, nsteps, solver=cr_solve):
    """``experiments/run-adi.py:43-71``: returns u after ``nsteps`` full steps."""
    u = np.array(u0, dtype=np.float64, copy=True)
    d_x = u.copy()
    d_y = u.T.copy()
    co = reference_coeffs(dx, dt)
    for _ in range(nsteps):
        adi_rhs(u, d_x, dx, dt)
        d_x[1:-1] = solver(co, d_x[1:-1])
        u = transpose(d_x)
        adi_rhs(u, d_y, dx, dt)
        d_y[1:-1] = solver(co, d_y[1:-1])
        u = transpose(d_y)
    return u


def _half_step(u, r, solver):
    """Implicit along the last axis, explicit along the first; rows 0 and N-1 fixed."""
    v = u.copy()
    d = u[1:-1].copy()
    d[:, 1:-1] = 0.5 * r * (u[:-2, 1:-1] + u[2:, 1:-1]) + (1.0 - r) * u[1:-1, 1:-1]
    v[1:-1] = solver(pr_coeffs(r), d)
    return v


def pr_step(u, r, solver=cr_solve):
    """One full PR step (x-implicit then y-implicit) on a square grid."""
    v = _half_step(np.asarray(u, dtype=np.float64), r, solver)
    w = _half_step(v.T, r, solver).T
    return np.ascontiguousarray(w)


def pr_run(u0, r, nsteps, solver=cr_solve):
    u = np.array(u0, dtype=np.float64, copy=True)
    for _ in range(nsteps):
        u = pr_step(u, r, solver)
    return u


def eigen_amplification(N, k, l, r):
    """Closed-form factor of one PR step on the mode sin(k pi x) sin(l pi y),
    x_i = i h, h = 1/(N-1), zero Dirichlet boundaries."""
    h = 1.0 / (N - 1)
    sk = np.sin(k * np.pi * h / 2) ** 2
    sl = np.sin(l * np.pi * h / 2) ** 2
    return ((1 - 2 * r * sl) * (1 - 2 * r * sk)) / ((1 + 2 * r * sk) * (1 + 2 * r * sl))


def gaussian_ic(N, seed=0, noise=1e-3):
    """Gaussian exp(-((x-.5)^2+(y-.5)^2)/0.01) plus seeded U[0, noise) on [0,1]^2,
    zero Dirichlet boundary (boundary values set to 0)."""
    x = np.linspace(0.0, 1.0, N)
    y, x2 = np.meshgrid(x, x, indexing="ij")
    u = np.exp(-((x2 - 0.5) ** 2 + (y - 0.5) ** 2) / 0.01)
    u += np.random.default_rng(seed).uniform(0.0, noise, size=(N, N))
    u[0, :] = u[-1, :] = 0.0
    u[:, 0] = u[:, -1] = 0.0
    return u
