"""GPU parity of the partition method along y and z, and of the partitioned ADI
steps with both half steps partitioned.

* The fused strided form (nt_part_solve_axis: tiles of adjacent lines, no
  permute) against the monolithic SciPy banded oracle of the same lines.
* The reference's own y and z sequence on the drop-in modules, one device for
  all subdomains: permute -> global_to_local -> compute_RHS -> solve ->
  get_params -> sum_solutions -> permute back
  (experiments/run-parallel-compact.py:92-164).
* Partitioned Peaceman-Rachford steps against the oracle's step, and their
  hipGraph replay against eager stepping.
"""
import numpy as np
import pytest
import torch

from oracle import adi, compact, cr

pytestmark = pytest.mark.gpu

COMPACT = (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0)
GENERAL = (2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7)


def rel_err(x, ref):
    return np.abs(np.asarray(x) - np.asarray(ref)).max() / np.abs(np.asarray(ref)).max()


# (n, P): subdomains on 16-, 32- and 64-lane segments, R = 1 .. 32, 16-, 8- and
# 4-line tiles (the last for 2048-point lines of 256-point subdomains)
SHAPES_NP = [(512, 4), (2048, 16), (96, 3), (64, 16), (1024, 8), (2048, 8), (300, 2), (2048, 1),
             (1812, 12), (34, 2)]


@pytest.mark.parametrize("n,P", SHAPES_NP)
@pytest.mark.parametrize("axis", [2, 1, 0])
def test_partitioned_compact_tiles_equal_oracle(dev, n, P, axis):
    """Compact derivative in tiles of lines along x (contiguous rows), y
    (several slabs) and z (one slab), the lines cut into P subdomains; 21, 63
    or 37 lines leave ragged tiles."""
    from ntamd.partition import PartitionPlan
    from ntamd.workloads import compact_field
    shape = {2: (3, 21, n), 1: (3, n, 21), 0: (n, 1, 37)}[axis]
    f = compact_field(shape, seed=n + P + axis)
    h = 1.0 / (n - 1)
    fd = torch.tensor(f, device=dev)
    out = torch.full_like(fd, float("nan"))
    PartitionPlan(n, P, COMPACT).compact_derivative(fd, out, h, axis)
    assert rel_err(out.cpu().numpy(), compact.derivative_banded(f, h, axis)) < 1e-12


@pytest.mark.parametrize("n,P", [(512, 4), (2048, 16), (96, 3), (64, 16), (1024, 8), (2048, 8), (34, 2)])
@pytest.mark.parametrize("axis", [2, 1, 0])
def test_partitioned_compact_paired_equals_oracle(dev, n, P, axis):
    """16-byte pairs: adjacent points of x-lines, adjacent y- or z-lines (even
    row strides; 22 and 46 lines leave ragged, even last tiles). The same fields
    in views 8 bytes into their buffers take the unpaired path."""
    from ntamd.partition import PartitionPlan
    from ntamd.workloads import compact_field
    shape = {2: (3, 7, n), 1: (3, n, 22), 0: (n, 2, 23)}[axis]
    f = compact_field(shape, seed=3 * n + P + axis)
    h = 1.0 / (n - 1)
    plan = PartitionPlan(n, P, COMPACT)
    want = compact.derivative_banded(f, h, axis)
    fd = torch.tensor(f, device=dev)
    out = torch.full_like(fd, float("nan"))
    plan.compact_derivative(fd, out, h, axis)
    assert rel_err(out.cpu().numpy(), want) < 1e-12
    # 8-byte offset views: the unpaired path
    fb = torch.zeros(fd.numel() + 1, dtype=torch.float64, device=dev)
    ob = torch.full_like(fb, float("nan"))
    fv = fb[1:].view(shape)
    ov = ob[1:].view(shape)
    fv.copy_(fd)
    plan.compact_derivative(fv, ov, h, axis)
    assert rel_err(ov.cpu().numpy(), want) < 1e-12


@pytest.mark.parametrize("n,P", [(512, 4), (2048, 16), (100, 2), (1024, 8)])
@pytest.mark.parametrize("axis", [2, 1, 0])
@pytest.mark.parametrize("coeffs", [COMPACT, GENERAL])
def test_partitioned_solve_axis_in_place(dev, n, P, axis, coeffs):
    from ntamd.partition import PartitionPlan
    shape = {2: (2, 17, n), 1: (2, n, 19), 0: (n, 3, 11)}[axis]
    d = np.random.default_rng(n * P + axis).standard_normal(shape)
    x = torch.tensor(d, device=dev)
    PartitionPlan(n, P, coeffs).solve_axis(x, axis)
    want = compact.last_to_axis(cr.banded_solve(coeffs, compact.axis_to_last(d, axis)), axis)
    assert rel_err(x.cpu().numpy(), want) < 1e-12


def test_partitioned_compact_strided_known_answer(dev):
    """run-parallel-compact.py:127,164: d(xyz)/dy = xz and d(xyz)/dz = xy with
    the y- and z-lines cut into 4 subdomains."""
    from ntamd.partition import PartitionPlan
    N = 128
    g = np.linspace(0, 1, N)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    f = torch.tensor(X * Y * Z, device=dev)
    plan = PartitionPlan(N, 4, COMPACT)
    for axis, true in ((1, X * Z), (0, X * Y)):
        out = torch.empty_like(f)
        plan.compact_derivative(f, out, g[1], axis)
        np.testing.assert_allclose(out.cpu().numpy(), true, rtol=1e-7, atol=1e-12)


def test_partitioned_512cubed_ddz_sampled(dev):
    """The 512^3 field along z in 4 subdomains of 128 (262,144 lines): 256 lines
    drawn at random equal the oracle."""
    from ntamd.partition import PartitionPlan
    from ntamd.workloads import compact_field_device
    N = 512
    h = 1.0 / (N - 1)
    fd = compact_field_device(N, seed=0, device=dev)
    out = torch.full_like(fd, float("nan"))
    PartitionPlan(N, 4, COMPACT).compact_derivative(fd, out, h, 0)
    rng = np.random.default_rng(9)
    ys = torch.tensor(rng.integers(0, N, 256), device=dev)
    xs = torch.tensor(rng.integers(0, N, 256), device=dev)
    cols = fd[:, ys, xs].t().contiguous().cpu().numpy()          # (256, N) lines
    got = out[:, ys, xs].t().contiguous().cpu().numpy()
    assert rel_err(got, compact.derivative_banded(cols[None], h, 2)[0]) < 1e-12
    assert not torch.isnan(out).any()
    del fd, out


def _dropin_axis_sequence(dev, N, P, direction):
    """The reference's y (direction 1) or z (direction 2) derivative on P
    subdomains along that axis, every step through the drop-in modules."""
    from neato import NearToeplitzSolver
    from compact.permute import permute
    from parallelcompact.da import DA, DA_gather_blocks, DA_scatter_blocks
    from parallelcompact.misc import solve_secondary_systems
    from parallelcompact.reduced import get_params
    from parallelcompact.rhs import compute_RHS
    from parallelcompact.sum import sum_solutions
    from ntamd.workloads import compact_field
    procs = {1: (1, P, 1), 2: (P, 1, 1)}[direction]
    to_last = {1: (0, 2, 1), 2: (1, 2, 0)}[direction]
    back = {1: (0, 2, 1), 2: (2, 0, 1)}[direction]
    axis = {1: 1, 2: 0}[direction]
    da = DA((N, N, N), procs, 1, device=dev)
    shape = tuple(N * p for p in procs)
    F = compact_field(shape, seed=10 * N + P + direction)
    h = 1.0 / (P * N - 1)
    f_d = da.create_global_vector()
    DA_scatter_blocks(da, torch.tensor(F, device=dev), f_d)
    line_da = da.get_line_DA(direction)
    # permute(f_d, x_d, ...) on every rank's block: the line axis last
    fp = line_da.create_global_vector()
    for p in range(P):
        idx = (0, p, 0) if direction == 1 else (p, 0, 0)
        permute(f_d[idx], fp[0, 0, p], to_last)
    f_local = line_da.create_local_vector()
    line_da.global_to_local(fp, f_local)
    x_d = line_da.create_global_vector()
    xU, xL = [], []
    for p in range(P):
        co = [1.0, 0.25, 0.25, 1.0, 0.25, 0.25, 1.0]
        if p == 0:
            co[1] = 2.0
        if p == P - 1:
            co[-2] = 2.0
        compute_RHS(f_local[0, 0, p], x_d[0, 0, p], h, (N, N, N), p, P)
        NearToeplitzSolver(N, N * N, co).solve(x_d[0, 0, p])
        u, l_ = solve_secondary_systems(N, p, P, co)
        xU.append(u)
        xL.append(l_)
    xU, xL = torch.stack(xU), torch.stack(xL)
    alpha, beta = get_params(line_da, xU, xL, x_d[0, 0])
    for p in range(P):
        sum_solutions(xU[p], xL[p], x_d[0, 0, p], alpha[p], beta[p], (N, N, N))
    y_d = da.create_global_vector()
    for p in range(P):
        idx = (0, p, 0) if direction == 1 else (p, 0, 0)
        permute(x_d[0, 0, p], y_d[idx], back)
    got = np.empty(shape)
    DA_gather_blocks(da, y_d, got)
    return got, compact.derivative_banded(F, h, axis)


@pytest.mark.parametrize("N,P", [(16, 1), (24, 3), (32, 4), (16, 8)])
@pytest.mark.parametrize("direction", [1, 2])
def test_dropin_reference_sequence_y_and_z(dev, N, P, direction):
    got, want = _dropin_axis_sequence(dev, N, P, direction)
    assert rel_err(got, want) < 1e-12


@pytest.mark.parametrize("N,P,r", [(2048, 16, 0.5), (512, 4, 1.3), (130, 2, 0.2), (1024, 8, 2.0), (99, 3, 0.8)])
def test_partitioned_adi_steps_equal_oracle(dev, N, P, r):
    """Both half steps partitioned (x-implicit rows, y-implicit columns), two
    steps from the Gaussian initial condition, against the oracle's PR steps.
    Even N moves the columns in 16-byte pairs, odd N (99) one at a time."""
    from ntamd.partition import PartitionedADISolver
    u0 = adi.gaussian_ic(N, seed=N)
    u = torch.tensor(u0, device=dev)
    PartitionedADISolver(N, r, P).step(u, 2)
    assert rel_err(u.cpu().numpy(), adi.pr_run(u0, r, 2)) < 1e-12


def test_partitioned_adi_y_half_step_equals_oracle(dev):
    """The partitioned y-implicit half step alone (2048^2, 16 x 128), from the
    work buffer the partitioned x-implicit half step left, against the
    oracle's half step along the columns."""
    from ntamd import adi_coeffs, core
    from ntamd.partition import PartitionPlan
    from oracle.adi import _half_step
    N, r = 2048, 0.5
    u0 = torch.tensor(adi.gaussian_ic(N, seed=3), device=dev)
    w = torch.empty_like(u0)
    core.adi_prepare(u0, w)
    plan = PartitionPlan(N, 16, adi_coeffs(r))
    plan.adi_half_x(u0, w, r)
    u = u0.clone()
    plan.adi_half_y(w, u, r)
    want = _half_step(w.cpu().numpy().T.copy(), r, cr.banded_solve).T
    assert rel_err(u.cpu().numpy(), want) < 1e-12


@pytest.mark.parametrize("N,P", [(2048, 16), (512, 4), (130, 2)])
def test_partitioned_adi_x_half_step_tiles_equal_one_line_kernel(dev, N, P):
    """The x-implicit half step in tiles of rows against the one-row-per-
    workgroup kernel and the oracle's half step."""
    from ntamd import adi_coeffs, core
    from ntamd.partition import PartitionPlan
    from oracle.adi import _half_step
    r = 0.7
    u0 = adi.gaussian_ic(N, seed=N + 1)
    u = torch.tensor(u0, device=dev)
    w1, w2 = torch.empty_like(u), torch.empty_like(u)
    core.adi_prepare(u, w1)
    core.adi_prepare(u, w2)
    plan = PartitionPlan(N, P, adi_coeffs(r))
    plan.adi_half_x(u, w1, r, tiles=True)
    plan.adi_half_x(u, w2, r, tiles=False)
    want = _half_step(u0, r, cr.banded_solve)
    assert rel_err(w1.cpu().numpy(), want) < 1e-12
    assert rel_err(w2.cpu().numpy(), want) < 1e-12


def test_partitioned_adi_graph_replay_equals_eager(dev):
    """20 partitioned steps captured once and replayed: the same bits as eager
    stepping, on every replay."""
    from ntamd.partition import PartitionedADISolver
    N, r, k = 512, 0.5, 20
    u0 = adi.gaussian_ic(N, seed=5)
    eager = torch.tensor(u0, device=dev)
    PartitionedADISolver(N, r, 4).step(eager, k)
    s = PartitionedADISolver(N, r, 4)
    ug = torch.tensor(u0, device=dev)
    s.capture(ug, k)
    for _ in range(3):
        ug.copy_(torch.tensor(u0, device=dev))
        s.prepare(ug)
        torch.cuda.synchronize()
        s.replay()
        torch.cuda.synchronize()
        assert torch.equal(ug, eager)
