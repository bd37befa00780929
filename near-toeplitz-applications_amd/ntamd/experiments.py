"""The reference's experiment drivers, on one GPU, through the drop-in modules.

``convergence_study`` follows ``experiments/run-compact-errors.py:14-58``:
* f = sin(x)/x^3 on [2, 6]^3 with N points per axis;
* d/dx from ``compact.rhs.compute_rhs`` + ``NearToeplitzSolver.solve``;
* the absolute error against (x cos x - 3 sin x)/x^4 along the line
  (z, y) = (0, 0), at the mid point, at the boundary point, and as a mean;
* the ratio of each error to the one at the next size.  The 4th-order
  interior scheme gives mid-point ratios near 16, and the 3rd-order boundary
  closure ratios near 8.

``partition_profile`` follows the d/dx part of
``experiments/run-parallel-compact-profile.py:51-104``:
* P subdomains of N^3 points along x, f = x y z on [0, 1]^3;
* the five phases of a rank (global-to-local halo fill, ghosted RHS, primary
  subdomain solve, reduced parameters, sum), timed with the reference's
  sync-and-time per phase;
* the fused one-launch partitioned derivative and the monolithic fused
  derivative of the same field, timed beside them;
* a check that the result equals y z, as the reference asserts.
"""
import math
import time

import numpy as np
import torch

__all__ = ["convergence_study", "partition_profile"]


def convergence_study(sizes=(32, 64, 128, 256), device="cuda"):
    """Errors and successive error ratios of the compact d/dx; returns a dict."""
    from compact.rhs import compute_rhs
    from neato import NearToeplitzSolver
    rows = []
    for N in sizes:
        x1, xn = 2.0, 6.0
        dx = (xn - x1) / (N - 1)
        g = np.linspace(x1, xn, N)
        z, y, x = np.meshgrid(g, g, g, indexing="ij")
        f = np.sin(x) / x ** 3
        f_d = torch.tensor(f, device=device)
        dfdx_d = torch.zeros_like(f_d)
        solver = NearToeplitzSolver(N, N * N, (1.0, 2.0, 1.0 / 4, 1.0, 1.0 / 4, 2.0, 1.0))
        compute_rhs(f_d, dfdx_d, dx)
        solver.solve(dfdx_d)
        dfdx = dfdx_d.cpu().numpy()
        true = (x * np.cos(x) - 3 * np.sin(x)) / x ** 4
        e = np.abs(true[0, 0, :] - dfdx[0, 0, :])
        rows.append({"N": N, "err_mid": float(e[N // 2]), "err_boundary": float(e[0]), "err_mean": float(e.mean())})
    ratios = []
    for a, b in zip(rows[:-1], rows[1:]):
        ratios.append({"from": a["N"], "to": b["N"], "mid": a["err_mid"] / b["err_mid"],
                       "boundary": a["err_boundary"] / b["err_boundary"], "mean": a["err_mean"] / b["err_mean"]})
    return {"errors": rows, "ratios": ratios}


class _Timer:
    """Synchronize, then wall-clock one phase (the reference's Timer.tic/toc
    around each phase, run-parallel-compact-profile.py:20-35)."""

    def tic(self):
        torch.cuda.synchronize()
        self.t = time.perf_counter()

    def toc(self):
        torch.cuda.synchronize()
        return time.perf_counter() - self.t


def partition_profile(N=128, P=4, nsteps=20, device="cuda"):
    """Per-phase milliseconds of the partitioned d/dx on P x-subdomains of N^3."""
    from neato import NearToeplitzSolver
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.partition import PartitionPlan
    from parallelcompact.da import DA, DA_arange, DA_gather_blocks
    from parallelcompact.misc import solve_secondary_systems
    from parallelcompact.reduced import get_params
    from parallelcompact.rhs import compute_RHS
    from parallelcompact.sum import sum_solutions

    da = DA((N, N, N), (1, 1, P), 1, device=device)
    x, y, z = DA_arange(da, (0, 1), (0, 1), (0, 1))
    f_d = torch.tensor(x * y * z, device=device)                 # (1, 1, P, N, N, N)
    dx = float(x[0, 0, 0, 0, 0, 1] - x[0, 0, 0, 0, 0, 0])
    x_d = da.create_global_vector()
    f_local_d = da.create_local_vector()
    line_da = da.get_line_DA(0)

    # per-rank matrices (run-parallel-compact-profile.py:59-65) and secondary solutions
    solvers, xU, xL = [], [], []
    for rank in range(P):
        coeffs = [1.0, 1.0 / 4, 1.0 / 4, 1.0, 1.0 / 4, 1.0 / 4, 1.0]
        if rank == 0:
            coeffs[1] = 2.0
        if rank == P - 1:
            coeffs[-2] = 2.0
        solvers.append(NearToeplitzSolver(N, N * N, coeffs))
        u, l_ = solve_secondary_systems(N, rank, P, coeffs)
        xU.append(u)
        xL.append(l_)
    xU_d, xL_d = torch.stack(xU), torch.stack(xL)

    def step(t=None, acc=None):
        def phase(name, fn):
            if t is not None:
                t.tic()
            fn()
            if t is not None:
                acc[name] += t.toc()
        phase("global_to_local", lambda: da.global_to_local(f_d, f_local_d))
        phase("rhs", lambda: [compute_RHS(f_local_d[0, 0, p], x_d[0, 0, p], dx, (N, N, N), p, P) for p in range(P)])
        phase("primary", lambda: [solvers[p].solve(x_d[0, 0, p]) for p in range(P)])
        ab = {}
        phase("params", lambda: ab.update(zip(("alpha", "beta"), get_params(line_da, xU_d, xL_d, x_d[0, 0]))))
        phase("sum", lambda: [sum_solutions(xU_d[p], xL_d[p], x_d[0, 0, p], ab["alpha"][p], ab["beta"][p],
                                            (N, N, N)) for p in range(P)])

    step()                                                           # warm-up, as the reference does
    acc = {k: 0.0 for k in ("global_to_local", "rhs", "primary", "params", "sum")}
    timer = _Timer()
    for _ in range(nsteps):
        step(timer, acc)
    phases_ms = {k: 1e3 * v / nsteps for k, v in acc.items()}
    phases_ms["total"] = sum(phases_ms.values())

    # the result is y z (the scheme is exact for f linear in x)
    got = np.empty((N, N, P * N))
    DA_gather_blocks(da, x_d, got)
    want = np.empty_like(got)
    DA_gather_blocks(da, torch.tensor(y * z), want)
    err = float(np.abs(got - want).max() / np.abs(want).max())

    # fused alternatives on the assembled (N, N, P N) field
    f_glob = torch.empty((N, N, P * N), dtype=torch.float64, device=device)
    DA_gather_blocks(da, f_d, f_glob)
    out = torch.empty_like(f_glob)
    pp = PartitionPlan(P * N, P, COMPACT_COEFFS)
    mono = Plan(P * N, COMPACT_COEFFS) if P * N <= 2048 else None
    fused = {}
    for name, fn in (("partitioned_one_launch", lambda: pp.compact_derivative_x(f_glob, out, dx)),
                     ("monolithic_one_launch",
                      (lambda: core.compact_derivative(mono, f_glob, out, dx, 2)) if mono else None)):
        if fn is None:
            continue
        fn()
        timer.tic()
        for _ in range(nsteps):
            fn()
        fused[name] = 1e3 * timer.toc() / nsteps
        fused[name + "_max_rel_err"] = float(np.abs(out.cpu().numpy() - want).max() / np.abs(want).max())
    return {"N": N, "P": P, "steps": nsteps, "phases_ms": phases_ms, "max_rel_err_vs_yz": err,
            "fused_ms": fused, "points": N ** 3 * P, "grid": (N, N, P * N), "dx": dx,
            "sqrt_eps_check": err < math.sqrt(np.finfo(np.float64).eps)}
