"""GPU parity of the compact derivative: the fused kernel and the reference's
unfused call sequence (run-compact.py:57-90) against the CPU oracle."""
import numpy as np
import pytest
import torch

from oracle import compact, cr

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
REF_TOL = 1e-10


def rel_err(x, ref):
    return np.abs(np.asarray(x) - np.asarray(ref)).max() / np.abs(np.asarray(ref)).max()


def _field(shape, seed=0):
    from ntamd.workloads import compact_field
    return compact_field(shape, seed=seed)


@pytest.mark.parametrize("axis", [2, 1, 0])
def test_config1_64cubed_fused_vs_banded(dev, axis):
    """64^3 sin*sin*sin + N(0,1e-3), FP64; the oracle is the SciPy banded solve."""
    from ntamd import COMPACT_COEFFS, Plan, core
    N = 64
    f = _field(N)
    h = 1.0 / (N - 1)
    want = compact.derivative_banded(f, h, axis)
    fd = torch.tensor(f, device=dev)
    out = torch.empty_like(fd)
    core.compact_derivative(Plan(N, COMPACT_COEFFS), fd, out, h, axis)
    assert rel_err(out.cpu().numpy(), want) < FP64_TOL


@pytest.mark.parametrize("shape", [(16, 24, 40), (33, 17, 65), (20, 96, 160), (5, 6, 513),
                                   (512, 3, 4), (7, 300, 9)])
def test_fused_all_axes_non_cubic(dev, shape):
    from ntamd import COMPACT_COEFFS, Plan, core
    f = np.random.default_rng(sum(shape)).standard_normal(shape)
    fd = torch.tensor(f, device=dev)
    for axis in range(3):
        n = shape[axis]
        if n < 4:
            continue
        h = 0.1 + 0.01 * axis
        want = compact.derivative(f, h, axis)
        out = torch.empty_like(fd)
        core.compact_derivative(Plan(n, COMPACT_COEFFS), fd, out, h, axis)
        assert rel_err(out.cpu().numpy(), want) < FP64_TOL, (shape, axis)


@pytest.mark.parametrize("axis", [2, 1, 0])
@pytest.mark.parametrize("sweeps", [1, 2])
def test_fused_fp32_refine(dev, axis, sweeps):
    from ntamd import COMPACT_COEFFS, Plan, core
    N = 96
    f = _field((N, N, N), seed=3)
    h = 1.0 / (N - 1)
    want = compact.derivative(f, h, axis)
    fd = torch.tensor(f, device=dev)
    out = torch.empty_like(fd)
    core.compact_derivative(Plan(N, COMPACT_COEFFS), fd, out, h, axis, "fp32_refine", sweeps)
    assert rel_err(out.cpu().numpy(), want) < REF_TOL


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
def test_fused_fp32_storage(dev, precision):
    """FP32-stored field, FP64 result: the oracle differentiates the same FP32 values."""
    from ntamd import COMPACT_COEFFS, Plan, core
    N = 64
    f32 = _field(N, seed=4).astype(np.float32)
    h = 1.0 / (N - 1)
    fd = torch.tensor(f32, device=dev)
    for axis in range(3):
        want = compact.derivative(f32.astype(np.float64), h, axis)
        out = torch.empty(f32.shape, dtype=torch.float64, device=dev)
        core.compact_derivative(Plan(N, COMPACT_COEFFS), fd, out, h, axis, precision, 2)
        assert rel_err(out.cpu().numpy(), want) < (FP64_TOL if precision == "fp64" else REF_TOL)


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("shape,axis", [
    ((8, 1024, 128), 1),      # 1024-point y-lines, 32 whole 32-line tiles (fast path, R = 16)
    ((1024, 8, 128), 0),      # 1024-point z-lines, rows 4 KiB apart (fast path)
    ((72, 512, 128), 1),      # 288 tiles: more tiles than workgroups, so the tile loop turns (R = 8)
    ((40, 1024, 256), 1),     # 320 tiles at R = 16
    ((8, 1024, 40), 1),       # 40 lines per slab: a ragged last tile, the general kernel
])
def test_fp32_storage_strided_tiles(dev, shape, axis, precision):
    """FP32-stored fields along strided 512- and 1024-point lines (32-line FP32
    tiles): the oracle differentiates the same FP32 values in FP64."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.workloads import compact_field
    f32 = compact_field(shape, seed=9).astype(np.float32)
    n = shape[axis]
    h = 1.0 / (n - 1)
    want = compact.derivative_banded(f32.astype(np.float64), h, axis)
    fd = torch.tensor(f32, device=dev)
    out = torch.full(shape, float("nan"), dtype=torch.float64, device=dev)
    core.compact_derivative(Plan(n, COMPACT_COEFFS), fd, out, h, axis, precision, 1)
    tol = FP64_TOL if precision == "fp64" else REF_TOL
    assert rel_err(out.cpu().numpy(), want) < tol


@pytest.mark.parametrize("variant", [None, 2])
@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("shape", [(512, 32, 8192), (512, 64, 8192)])
def test_fp32_storage_z_lines_far_rows(dev, shape, precision, variant):
    """FP32-stored z-lines with rows 1 MiB and 2 MiB apart, 262,144 and 524,288
    lines, with the default launch shape (the FP64 solve: one lane per line)
    and with the 32-line FP32 tile kernels (variant 2: rows 1 MiB apart take
    the fast tile path, 2 MiB apart the general one; FP32 + refinement takes
    them by default): 256 lines drawn at random are checked against the oracle
    (lines are independent)."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.autotune import _set, current
    from ntamd.workloads import compact_field_device
    n = shape[0]
    h = 1.0 / (n - 1)
    fd = compact_field_device(shape, seed=11, dtype=torch.float32, device=dev)
    out = torch.full(shape, float("nan"), dtype=torch.float64, device=dev)
    keep, _ = current(1, n, precision)
    _set(1, precision, n, keep if variant is None else variant)
    try:
        core.compact_derivative(Plan(n, COMPACT_COEFFS), fd, out, h, 0, precision, 1)
    finally:
        _set(1, precision, n, keep)
    rng = np.random.default_rng(3)
    ys = torch.tensor(rng.integers(0, shape[1], 256), device=dev)
    xs = torch.tensor(rng.integers(0, shape[2], 256), device=dev)
    cols = fd[:, ys, xs].double().cpu().numpy()[:, :, None]        # (n, 256, 1)
    want = compact.derivative_banded(cols, h, 0)
    got = out[:, ys, xs].cpu().numpy()[:, :, None]
    tol = FP64_TOL if precision == "fp64" else REF_TOL
    assert rel_err(got, want) < tol
    del fd, out


@pytest.fixture(scope="module")
def field_1024_fp32(dev):
    from ntamd.workloads import compact_field_device
    f = compact_field_device(1024, seed=0, dtype=torch.float32, device=dev)
    yield f
    del f
    torch.cuda.empty_cache()


@pytest.mark.parametrize("precision,sweeps", [("fp64", 1), ("fp32_refine", 1), ("fp32_refine", 2)])
@pytest.mark.parametrize("axis", [2, 1, 0])
def test_1024cubed_fp32_storage_sampled_lines(dev, field_1024_fp32, axis, precision, sweeps):
    """The full 1024^3 FP32-stored field (4 GiB, 1,048,576 lines per direction):
    256 lines drawn at random along each axis equal the oracle's derivative of
    the same FP32 values in FP64, and no point of the result is left unwritten."""
    from ntamd import COMPACT_COEFFS, Plan, core
    N = 1024
    h = 1.0 / (N - 1)
    fd = field_1024_fp32
    out = torch.full(fd.shape, float("nan"), dtype=torch.float64, device=dev)
    core.compact_derivative(Plan(N, COMPACT_COEFFS), fd, out, h, axis, precision, sweeps)
    assert not torch.isnan(out).any()
    rng = np.random.default_rng(40 + axis)
    i = torch.tensor(rng.integers(0, N, 256), device=dev)
    j = torch.tensor(rng.integers(0, N, 256), device=dev)
    idx = [i, j]
    idx.insert(axis, slice(None))
    # the gathered lines are (256, N) for axes 2 and 1 (separated index arrays
    # put their dimension first) and (N, 256) for axis 0: make them rows
    src = fd[tuple(idx)].double().cpu().numpy()
    got = out[tuple(idx)].cpu().numpy()
    if src.shape == (N, 256):
        src, got = src.T.copy(), got.T.copy()
    assert src.shape == (256, N)
    want = compact.derivative_banded(src[None], h, 2)[0]
    tol = FP64_TOL if precision == "fp64" else REF_TOL
    assert rel_err(got, want) < tol
    del out


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("shape", [(1024, 1024, 768), (1024, 999, 801)])
def test_z_lines_rows_6mib_apart(dev, shape, precision):
    """FP64-stored z-lines with rows at least 6 MiB apart, walked in contiguous
    tile ranges (the FP64 solve on the general tile kernel; FP32 + refinement
    takes the whole-tile fast path for whole tiles): 786,432 lines in whole
    tiles and 800,199 lines with a ragged last tile.  256 lines drawn at random
    equal the oracle."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.workloads import compact_field_device
    n = shape[0]
    h = 1.0 / (n - 1)
    fd = compact_field_device(shape, seed=13, device=dev)
    out = torch.full(shape, float("nan"), dtype=torch.float64, device=dev)
    core.compact_derivative(Plan(n, COMPACT_COEFFS), fd, out, h, 0, precision, 1)
    rng = np.random.default_rng(5)
    ys = torch.tensor(rng.integers(0, shape[1], 256), device=dev)
    xs = torch.tensor(rng.integers(0, shape[2], 256), device=dev)
    cols = fd[:, ys, xs].cpu().numpy()[:, :, None]
    want = compact.derivative_banded(cols, h, 0)
    got = out[:, ys, xs].cpu().numpy()[:, :, None]
    tol = FP64_TOL if precision == "fp64" else REF_TOL
    assert rel_err(got, want) < tol
    assert not torch.isnan(out[:, -1, -1]).any()     # the last (ragged) tile's last line
    del fd, out


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("shape,axis", [
    ((8, 1024, 64), 1),       # 1024-point y-lines in 32 whole 16-line tiles (fast path, R = 16)
    ((40, 1024, 128), 1),     # 320 tiles: more tiles than workgroups, so the tile loop turns
    ((1024, 4, 64), 0),       # z-lines with rows 2 KiB apart (fast path)
])
def test_fp64_1024_point_strided_fast_path(dev, shape, axis, precision):
    """FP64-stored 1024-point strided lines on the whole-tile fast path at 16 rows
    per lane, in FP64 and with FP32 + one refinement sweep, equal the oracle."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.workloads import compact_field
    f = compact_field(shape, seed=17)
    n = shape[axis]
    h = 1.0 / (n - 1)
    want = compact.derivative_banded(f, h, axis)
    fd = torch.tensor(f, device=dev)
    out = torch.full_like(fd, float("nan"))
    core.compact_derivative(Plan(n, COMPACT_COEFFS), fd, out, h, axis, precision, 1)
    tol = FP64_TOL if precision == "fp64" else REF_TOL
    assert rel_err(out.cpu().numpy(), want) < tol


def test_reference_call_sequence_unfused(dev):
    """run-compact.py:57-90 verbatim on the drop-in modules: permute -> compute_rhs ->
    NearToeplitzSolver.solve -> permute, for x, y and z."""
    from neato import NearToeplitzSolver
    from compact.rhs import compute_rhs
    from compact.permute import permute
    N = 48
    L = 2 * np.pi
    dx = L / (N - 1)
    xs = np.linspace(0, L, N)
    z, y, x = np.meshgrid(xs, xs, xs, indexing="ij")
    f = np.sin(x) + 2 * np.sin(y) + 3 * np.sin(z)
    f_d = torch.tensor(f, device=dev)
    dfdx_d = torch.zeros_like(f_d)
    dfdy_d = torch.zeros_like(f_d)
    dfdz_d = torch.zeros_like(f_d)
    tmp_d = torch.zeros_like(f_d)
    solver = NearToeplitzSolver(N, N * N, (1., 2., 1. / 4, 1., 1. / 4, 2., 1.))
    compute_rhs(f_d, dfdx_d, dx)
    solver.solve(dfdx_d)
    permute(f_d, dfdy_d, (0, 2, 1))
    compute_rhs(dfdy_d, tmp_d, dx)
    solver.solve(tmp_d)
    permute(tmp_d, dfdy_d, (0, 2, 1))
    permute(f_d, dfdz_d, (1, 2, 0))
    compute_rhs(dfdz_d, tmp_d, dx)
    solver.solve(tmp_d)
    permute(tmp_d, dfdz_d, (2, 0, 1))
    for got, axis, true in [(dfdx_d, 2, np.cos(x)), (dfdy_d, 1, 2 * np.cos(y)),
                            (dfdz_d, 0, 3 * np.cos(z))]:
        got = got.cpu().numpy()
        assert rel_err(got, compact.derivative(f, dx, axis)) < FP64_TOL
        np.testing.assert_allclose(got, true, rtol=1e-1, atol=1e-2)   # run-compact.py:99-102


def test_dropin_rhs_and_permute_exact(dev):
    from compact.rhs import compute_rhs
    from compact.permute import permute
    f = np.random.default_rng(8).standard_normal((5, 7, 11))
    fd = torch.tensor(f, device=dev)
    r = torch.empty_like(fd)
    compute_rhs(fd, r, 0.3)
    np.testing.assert_allclose(r.cpu().numpy(), compact.compact_rhs(f, 0.3), rtol=1e-15, atol=1e-15)
    for p in [(0, 2, 1), (1, 2, 0), (2, 0, 1), (1, 0, 2), (2, 1, 0), (0, 1, 2)]:
        out = torch.empty(tuple(f.shape[i] for i in p), dtype=torch.float64, device=dev)
        permute(fd, out, p)
        np.testing.assert_array_equal(out.cpu().numpy(), compact.permute(f, p))


def test_known_answer_xyz_all_axes(dev):
    """Reference KAT (run-parallel-compact.py:89-90 etc.): d(xyz)/dx = yz exactly."""
    from ntamd import COMPACT_COEFFS, Plan, core
    N = 40
    xs = np.linspace(0, 1, N)
    Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
    fd = torch.tensor(X * Y * Z, device=dev)
    h = xs[1]
    for axis, true in [(2, Y * Z), (1, X * Z), (0, X * Y)]:
        out = torch.empty_like(fd)
        core.compact_derivative(Plan(N, COMPACT_COEFFS), fd, out, h, axis)
        np.testing.assert_allclose(out.cpu().numpy(), true, rtol=1e-7, atol=1e-12)


@pytest.mark.parametrize("axis", [2, 1, 0])
def test_512cubed_properties(dev, axis):
    """Full-size (512^3) checks: exactness on a cubic polynomial (the scheme is
    exact for cubics) and sampled lines of the seeded field against the oracle."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.workloads import compact_field_device
    N = 512
    h = 1.0 / (N - 1)
    t = torch.linspace(0, 1, N, dtype=torch.float64, device=dev)
    shape = [1, 1, 1]
    shape[axis] = N
    line = t.reshape(shape)
    other = [a for a in range(3) if a != axis][0]
    oshape = [1, 1, 1]
    oshape[other] = N
    # a cubic along the line plus a term that is constant along it
    f = (line ** 3 - 0.5 * line ** 2 + 0.25 * line + 7.0 * t.reshape(oshape)).expand(N, N, N)
    f = f.contiguous()
    out = torch.empty_like(f)
    plan = Plan(N, COMPACT_COEFFS)
    core.compact_derivative(plan, f, out, h, axis)
    true = (3 * line ** 2 - line + 0.25).expand(N, N, N)
    err = (out - true).abs().max().item()
    assert err < 1e-9, err
    del f, true
    g = compact_field_device(N, seed=0, device=dev)
    core.compact_derivative(plan, g, out, h, axis)
    rng = np.random.default_rng(axis)
    for _ in range(8):
        i, j = rng.integers(0, N, 2)
        sl = [i, j]
        sl.insert(axis, slice(None))
        src = g[tuple(sl)].cpu().numpy()
        want = cr.banded_solve(compact.COMPACT_COEFFS, compact.compact_rhs(src[None], h))[0]
        assert rel_err(out[tuple(sl)].cpu().numpy(), want) < FP64_TOL


@pytest.mark.parametrize("variant", [None, 2])
@pytest.mark.parametrize("shape", [(512, 64, 128), (512, 33, 250)])
def test_z_lines_far_rows(dev, shape, variant):
    """z-lines whose rows are 64 KiB or more apart: 8192 lines and 8250 lines,
    with the default launch shape (one lane per line, 128 groups of 64 lines
    and a ragged group of 58) and with the whole-tile kernel (variant 2: whole
    aligned tiles take its fast path, the ragged last tile the general kernel).
    All equal the oracle."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.autotune import _set, current
    from ntamd.workloads import compact_field
    f = compact_field(shape, seed=7)
    n = shape[0]
    h = 1.0 / (n - 1)
    want = compact.derivative_banded(f, h, 0)
    fd = torch.tensor(f, device=dev)
    out = torch.full_like(fd, float("nan"))
    keep, _ = current(1, n, "fp64")
    _set(1, "fp64", n, keep if variant is None else variant)
    try:
        core.compact_derivative(Plan(n, COMPACT_COEFFS), fd, out, h, 0, "fp64")
    finally:
        _set(1, "fp64", n, keep)
    assert np.abs(out.cpu().numpy() - want).max() / np.abs(want).max() < 1e-12
