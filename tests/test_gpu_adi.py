"""GPU parity of the Peaceman-Rachford ADI paths against the CPU oracle: the
reference's six-launch loop on the drop-in modules, the fused two-launch step,
and its hipGraph replay."""
import numpy as np
import pytest
import torch

from oracle import adi

pytestmark = pytest.mark.gpu


def rel_err(x, ref):
    return np.abs(np.asarray(x) - np.asarray(ref)).max() / np.abs(np.asarray(ref)).max()


def test_reference_loop_on_dropins(dev):
    """experiments/run-adi.py:43-71 verbatim (non-zero Dirichlet data, dt = 0.1)."""
    from neato import NearToeplitzSolver
    from adi.rhs import compute_rhs
    from adi.transpose import transpose
    N = 96
    dx = 1.0 / (N - 1)
    dt = 0.1
    u = np.full((N, N), 0.5)
    u[:, 0] = 1.0
    u_gpu = torch.tensor(u, device=dev)
    d_x_gpu = torch.tensor(u.copy(), device=dev)
    d_y_gpu = torch.tensor(u.T.copy(), device=dev)
    co = (1., 0., 1. / (dx * dx), -2. * (1. / dt + 1. / (dx * dx)), 1. / (dx * dx), 0., 1.)
    solver_x = NearToeplitzSolver(N, N - 2, co)
    solver_y = NearToeplitzSolver(N, N - 2, co)
    nsteps = 6
    for _ in range(nsteps):
        compute_rhs(u_gpu, d_x_gpu, dx, dt, N)
        solver_x.solve(d_x_gpu[1, 0])
        transpose(d_x_gpu, u_gpu, N)
        compute_rhs(u_gpu, d_y_gpu, dx, dt, N)
        solver_y.solve(d_y_gpu[1, 0])
        transpose(d_y_gpu, u_gpu, N)
    want = adi.reference_run(u, dx, dt, nsteps)
    assert rel_err(u_gpu.cpu().numpy(), want) < 1e-10     # cond ~4e4 unscaled rows


@pytest.mark.parametrize("N,r,nsteps", [(34, 0.5, 3), (256, 0.5, 5), (300, 1.7, 4),
                                        (1024, 0.1, 2), (2048, 0.5, 2)])
def test_fused_steps_match_oracle(dev, N, r, nsteps):
    from adi.solver import ADISolver
    u0 = adi.gaussian_ic(N, seed=1)
    u = torch.tensor(u0, device=dev)
    ADISolver(N, r).step(u, nsteps)
    want = adi.pr_run(u0, r, nsteps)
    assert rel_err(u.cpu().numpy(), want) < 1e-12


@pytest.mark.parametrize("N", [1024, 2048, 1800])
def test_fused_steps_fp32_refine(dev, N):
    """FP32 solves plus one FP64-residual sweep in both half steps (at 2048 the
    y half step takes the 16-byte-row tile kernel, at 1800 the general one)
    equal the oracle's FP64 Peaceman-Rachford steps to the refinement tolerance."""
    from adi.solver import ADISolver
    u0 = adi.gaussian_ic(N, seed=3)
    u = torch.tensor(u0, device=dev)
    ADISolver(N, 0.5, precision="fp32_refine", sweeps=1).step(u, 2)
    want = adi.pr_run(u0, 0.5, 2)
    assert rel_err(u.cpu().numpy(), want) < 1e-10


def test_graph_replay_matches_eager_2048(dev):
    """The 2048^2 steps (coalesced-stencil x half step, 16-byte-row y half step)
    replayed from a hipGraph equal eager stepping bit for bit."""
    from adi.solver import ADISolver
    N = 2048
    u0 = adi.gaussian_ic(N, seed=4)
    eager = torch.tensor(u0, device=dev)
    ADISolver(N, 0.5).step(eager, 3)
    ug = torch.tensor(u0, device=dev)
    g = ADISolver(N, 0.5)
    g.capture(ug, 3)
    ug.copy_(torch.tensor(u0, device=dev))
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(ug, eager)


def test_fused_equals_reference_loop_form(dev):
    """Fused r-form and the reference's unscaled form agree (same method)."""
    from adi.solver import ADISolver
    N = 128
    h = 1.0 / (N - 1)
    r = 0.5
    u0 = np.full((N, N), 0.5)
    u0[:, 0] = 1.0
    u = torch.tensor(u0, device=dev)
    ADISolver(N, r).step(u, 3)
    want = adi.reference_run(u0, h, r * h * h, 3)
    assert rel_err(u.cpu().numpy(), want) < 1e-11


@pytest.mark.parametrize("k,l,r", [(3, 5, 0.5), (7, 2, 2.0)])
def test_eigenmode_decay(dev, k, l, r):
    from adi.solver import ADISolver
    N = 512
    xs = np.linspace(0, 1, N)
    Y, X = np.meshgrid(xs, xs, indexing="ij")
    u0 = np.sin(k * np.pi * X) * np.sin(l * np.pi * Y)
    u = torch.tensor(u0, device=dev)
    ADISolver(N, r).step(u, 10)
    g = adi.eigen_amplification(N, k, l, r)
    assert np.abs(u.cpu().numpy() - g ** 10 * u0).max() < 1e-12


def test_graph_replay_matches_eager(dev):
    from adi.solver import ADISolver
    N = 512
    u0 = adi.gaussian_ic(N, seed=2)
    eager = torch.tensor(u0, device=dev)
    s = ADISolver(N, 0.5)
    s.step(eager, 8)
    ug = torch.tensor(u0, device=dev)
    g = ADISolver(N, 0.5)
    g.capture(ug, 4)
    ug.copy_(torch.tensor(u0, device=dev))
    for rep in range(2):
        g.replay()
        torch.cuda.synchronize()
        _ = (ug * 2).sum().item()       # eager work between replays
    assert torch.equal(ug, eager)
    # a third replay continues from the current state, as eager stepping does
    s.step(eager, 4, prepare=False)
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(ug, eager)
