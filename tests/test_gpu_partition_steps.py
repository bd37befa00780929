"""The partition method one step at a time (parallelcompact drop-ins) on the GPU.

Each step is checked against the oracle's restatement of the reference step
(ghosted RHS rhs.py:33-52, reduced system + Thomas reduced.py:11-49,163-178,
correction sum.py:30); the whole per-rank sequence of
run-parallel-compact.py:72-76 is checked against the monolithic SciPy solve of
the assembled line (the partition method is exact, so they agree to rounding);
and the reference's two experiment drivers are reproduced."""
import numpy as np
import pytest
import torch

from oracle import compact, partition

pytestmark = pytest.mark.gpu

COMPACT = (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0)


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


@pytest.mark.parametrize("mx,npx", [(0, 1), (0, 3), (1, 3), (2, 3)])
@pytest.mark.parametrize("shape", [(4, 5, 6), (3, 7, 33), (8, 8, 64)])
def test_ghosted_rhs_matches_oracle(dev, shape, mx, npx):
    from parallelcompact.rhs import compute_RHS
    rng = np.random.default_rng(sum(shape) + 7 * mx)
    nz, ny, nx = shape
    fl = rng.standard_normal((nz + 2, ny + 2, nx + 2))
    x = torch.empty(shape, dtype=torch.float64, device=dev)
    compute_RHS(torch.tensor(fl, device=dev), x, 0.01, shape, mx, npx)
    assert rel(x.cpu().numpy(), partition.ghosted_rhs(fl, 0.01, mx, npx)) < 1e-15


def test_sum_solutions_matches_oracle(dev):
    from parallelcompact.sum import sum_solutions
    rng = np.random.default_rng(3)
    nz, ny, nx = 5, 6, 70
    xr = rng.standard_normal((nz, ny, nx))
    xu, xl = rng.standard_normal(nx), rng.standard_normal(nx)
    al, be = rng.standard_normal((nz, ny)), rng.standard_normal((nz, ny))
    t = lambda a: torch.tensor(a, device=dev)  # noqa: E731
    xr_d = t(xr)
    sum_solutions(t(xu), t(xl), xr_d, t(al), t(be), (nz, ny, nx))
    want = xr + al[:, :, None] * xu[None, None, :] + be[:, :, None] * xl[None, None, :]
    assert rel(xr_d.cpu().numpy(), want) < 1e-15


@pytest.mark.parametrize("P", [1, 2, 5, 16])
def test_get_params_matches_oracle(dev, P):
    from parallelcompact.reduced import get_params
    from parallelcompact.da import DA
    rng = np.random.default_rng(P)
    nz, ny, m = 3, 4, 9
    xu = rng.uniform(-0.5, 0.5, (P, m))
    xl = rng.uniform(-0.5, 0.5, (P, m))
    xr = rng.standard_normal((P, nz, ny, m))
    line_da = DA((nz, ny, m), (1, 1, P), 1, device=dev)
    t = lambda a: torch.tensor(a, device=dev)  # noqa: E731
    alpha, beta = get_params(line_da, t(xu), t(xl), t(xr))
    xU_line = np.stack([xu[:, 0], xu[:, -1]], axis=1).ravel()
    xL_line = np.stack([xl[:, 0], xl[:, -1]], axis=1).ravel()
    a, b, c = partition.reduced_system(xU_line, xL_line)
    faces = np.empty((2 * P, nz * ny))
    faces[0::2] = -xr[:, :, :, 0].reshape(P, -1)
    faces[1::2] = -xr[:, :, :, -1].reshape(P, -1)
    faces[0] = 0.0
    faces[-1] = 0.0
    sol = partition.thomas(a, b, c, faces)
    np.testing.assert_allclose(alpha.cpu().numpy().reshape(P, -1), sol[0::2], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(beta.cpu().numpy().reshape(P, -1), sol[1::2], rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("N,P", [(16, 1), (24, 3), (32, 4), (16, 16)])
def test_per_rank_sequence_equals_monolithic_solve(dev, N, P):
    """global_to_local -> compute_RHS -> solve -> get_params -> sum_solutions on
    every subdomain (run-parallel-compact.py:72-76) == the compact derivative of
    the assembled (N, N, P N) field by the SciPy banded oracle."""
    from neato import NearToeplitzSolver
    from parallelcompact.da import DA, DA_gather_blocks, DA_scatter_blocks
    from parallelcompact.misc import solve_secondary_systems
    from parallelcompact.reduced import get_params
    from parallelcompact.rhs import compute_RHS
    from parallelcompact.sum import sum_solutions
    from ntamd.workloads import compact_field
    da = DA((N, N, N), (1, 1, P), 1, device=dev)
    F = compact_field((N, N, P * N), seed=N + P)
    dx = 1.0 / (P * N - 1)
    f_d = da.create_global_vector()
    DA_scatter_blocks(da, torch.tensor(F, device=dev), f_d)
    f_local = da.create_local_vector()
    da.global_to_local(f_d, f_local)
    x_d = da.create_global_vector()
    xU, xL = [], []
    for p in range(P):
        co = [1.0, 0.25, 0.25, 1.0, 0.25, 0.25, 1.0]
        if p == 0:
            co[1] = 2.0
        if p == P - 1:
            co[-2] = 2.0
        compute_RHS(f_local[0, 0, p], x_d[0, 0, p], dx, (N, N, N), p, P)
        NearToeplitzSolver(N, N * N, co).solve(x_d[0, 0, p])
        u, l_ = solve_secondary_systems(N, p, P, co)
        xU.append(u)
        xL.append(l_)
    xU, xL = torch.stack(xU), torch.stack(xL)
    alpha, beta = get_params(da.get_line_DA(0), xU, xL, x_d[0, 0])
    for p in range(P):
        sum_solutions(xU[p], xL[p], x_d[0, 0, p], alpha[p], beta[p], (N, N, N))
    got = np.empty((N, N, P * N))
    DA_gather_blocks(da, x_d, got)
    assert rel(got, compact.derivative_banded(F, dx, 2)) < 1e-12


def test_convergence_study_reproduces_orders(dev):
    """run-compact-errors.py: 4th-order interior (mid ratio ~16), 3rd-order
    closure (boundary ratio -> 8); ratios equal the oracle's for the same grids."""
    from ntamd.experiments import convergence_study
    sizes = (32, 64, 128)
    res = convergence_study(sizes)
    for r, (n0, n1) in zip(res["ratios"], zip(sizes[:-1], sizes[1:])):
        want = []
        for N in (n0, n1):
            g = np.linspace(2.0, 6.0, N)[None, None, :]
            d = compact.derivative_banded(np.sin(g) / g ** 3, 4.0 / (N - 1), 2)[0, 0]
            e = np.abs((g * np.cos(g) - 3 * np.sin(g)) / g ** 4 - d)[0, 0]
            want.append((e[N // 2], e[0], e.mean()))
        assert abs(r["mid"] - want[0][0] / want[1][0]) < 1e-3 * r["mid"]
        assert abs(r["boundary"] - want[0][1] / want[1][1]) < 1e-3 * r["boundary"]
        assert 15.0 < r["mid"] < 17.0 and 6.5 < r["boundary"] < 8.5


def test_partition_profile_small(dev):
    from ntamd.experiments import partition_profile
    res = partition_profile(N=32, P=4, nsteps=2)
    assert res["max_rel_err_vs_yz"] < 1e-12
    assert res["fused_ms"]["partitioned_one_launch_max_rel_err"] < 1e-12
    assert res["fused_ms"]["monolithic_one_launch_max_rel_err"] < 1e-12
    assert all(v >= 0 for v in res["phases_ms"].values())
