"""The partition (superposition) method for a line cut into P subdomains, on the CPU.

TEST INFRASTRUCTURE ONLY (see the package docstring).

Subdomain p of a line holds m = n/P consecutive unknowns.  Its local solution
is ``x = xR + alpha xU + beta xL`` where ``xR`` solves the subdomain block with
the local RHS, ``xU``/``xL`` solve it with ``-a e_0`` / ``-c e_{m-1}``
(``parallelcompact/misc.py:21-46``), ``alpha_p`` is the last value of
subdomain p-1 and ``beta_p`` the first value of subdomain p+1.  The 2P unknowns
``(alpha_0, beta_0, alpha_1, ...)`` satisfy the reduced tridiagonal system of
``parallelcompact/reduced.py:163-178``, solved by the serial Thomas sweep of
``reduced.py:11-49``; the correction is ``parallelcompact/sum.py:30``.
"""
import numpy as np

from .cr import cr_solve, tridiagonal_solve

__all__ = ["subdomain_coeffs", "solve_secondary_systems", "reduced_system",
           "thomas", "partitioned_solve", "ghosted_rhs"]


def subdomain_coeffs(coeffs, p, P):
    """7-tuple of subdomain p: interior subdomains are the Toeplitz principal
    block; the first keeps the physical first row, the last the physical last
    row (``experiments/run-parallel-compact.py:63-67`` for the compact matrix)."""
    b1, c1, ai, bi, ci, an, bn = [float(v) for v in coeffs]
    fb, fc = (b1, c1) if p == 0 else (bi, ci)
    la, lb = (an, bn) if p == P - 1 else (ai, bi)
    return (fb, fc, ai, bi, ci, la, lb)


def solve_secondary_systems(nx, rank, size, coeffs):
    """``parallelcompact/misc.py:21-46``: homogeneous solutions (xU, xL)."""
    b1, c1, ai, bi, ci, an, bn = [float(v) for v in coeffs]
    a = np.full(nx, ai); b = np.full(nx, bi); c = np.full(nx, ci)
    if rank == 0:
        a[0], b[0], c[0] = 0.0, b1, c1
    if rank == size - 1:
        a[-1], b[-1], c[-1] = an, bn, 0.0
    rU = np.zeros(nx); rL = np.zeros(nx)
    rU[0] = -a[0]
    rL[-1] = -c[-1]
    return tridiagonal_solve(a, b, c, rU), tridiagonal_solve(a, b, c, rL)


def reduced_system(xU_line, xL_line):
    """Diagonals of the 2P reduced system (``parallelcompact/reduced.py:163-178``).

    ``xU_line``/``xL_line`` hold ``[x[0], x[-1]]`` of every subdomain in order."""
    P2 = len(xU_line)
    a = np.zeros(P2); b = np.zeros(P2); c = np.zeros(P2)
    a[0::2] = -1.0
    a[1::2] = xU_line[1::2]
    b[0::2] = xU_line[0::2]
    b[1::2] = xL_line[1::2]
    c[0::2] = xL_line[0::2]
    c[1::2] = -1.0
    a[0] = c[0] = 0.0
    b[0] = 1.0
    a[-1] = c[-1] = 0.0
    b[-1] = 1.0
    a[1] = 0.0
    c[-2] = 0.0
    return a, b, c


def thomas(a, b, c, d):
    """Serial Thomas sweep over the first axis of ``d`` (``reduced.py:11-49``)."""
    d = np.array(d, dtype=np.float64, copy=True)
    n = len(b)
    c2 = np.zeros(n)
    c2[0] = c[0] / b[0]
    d[0] = d[0] / b[0]
    for i in range(1, n):
        bmac = b[i] - a[i] * c2[i - 1]
        c2[i] = c[i] / bmac
        d[i] = (d[i] - a[i] * d[i - 1]) / bmac
    for i in range(n - 2, -1, -1):
        d[i] = d[i] - c2[i] * d[i + 1]
    return d


def partitioned_solve(coeffs, d, P, solver=cr_solve):
    """Solve every trailing line of ``d`` (length n = P*m) by the partition method."""
    d = np.asarray(d, dtype=np.float64)
    n = d.shape[-1]
    if n % P:
        raise ValueError("n must be a multiple of P")
    m = n // P
    lines = d.reshape(-1, P, m)
    xR = np.empty_like(lines)
    xU = np.empty((P, m)); xL = np.empty((P, m))
    for p in range(P):
        co = subdomain_coeffs(coeffs, p, P)
        xR[:, p, :] = solver(co, lines[:, p, :])
        xU[p], xL[p] = solve_secondary_systems(m, p, P, coeffs)
    xU_line = np.stack([xU[:, 0], xU[:, -1]], axis=1).ravel()
    xL_line = np.stack([xL[:, 0], xL[:, -1]], axis=1).ravel()
    a, b, c = reduced_system(xU_line, xL_line)
    faces = np.empty((2 * P, lines.shape[0]))
    faces[0::2] = -xR[:, :, 0].T
    faces[1::2] = -xR[:, :, -1].T
    faces[0] = 0.0           # no neighbour before the first subdomain
    faces[-1] = 0.0          # none after the last
    sol = thomas(a, b, c, faces)
    alpha, beta = sol[0::2].T, sol[1::2].T          # (lines, P)
    x = xR + alpha[:, :, None] * xU[None] + beta[:, :, None] * xL[None]
    return x.reshape(d.shape)


def ghosted_rhs(f_local, dx, mx, npx):
    """``parallelcompact/rhs.py:33-52`` on a width-1 ghosted (nz+2, ny+2, nx+2) array."""
    f = np.asarray(f_local, dtype=np.float64)
    c = f[1:-1, 1:-1, :]
    r = (3.0 / (4 * dx)) * (c[..., 2:] - c[..., :-2])
    if mx == 0:
        r[..., 0] = (1.0 / (2 * dx)) * (-5 * c[..., 1] + 4 * c[..., 2] + c[..., 3])
    if mx == npx - 1:
        r[..., -1] = -(1.0 / (2 * dx)) * (-5 * c[..., -2] + 4 * c[..., -3] + c[..., -4])
    return r
