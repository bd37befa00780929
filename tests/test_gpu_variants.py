"""Every launch-shape variant of the tuning table gives the oracle's answer.

The variants differ only in how lines are mapped onto wavefronts and staged
(wave-private LDS transpose, direct chunk loads, workgroup tiles, strided tiles
of 16 or 32 lines), so each must match the SciPy banded oracle to the FP64 (or
FP32 + refinement) tolerance, and the variants must agree with each other to a
few roundings."""
import numpy as np
import pytest
import torch

from oracle import compact

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32_refine": 1e-10}


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("n", [64, 100, 256, 512, 768, 1024, 2048])
@pytest.mark.parametrize("axis", [2, 1, 0])
def test_every_variant_matches_oracle(dev, precision, n, axis):
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.autotune import _set, current
    from ntamd.workloads import compact_field
    layout = 0 if axis == 2 else 1
    shape = {2: (3, 40, n), 1: (3, n, 40), 0: (n, 2, 36)}[axis]
    f = compact_field(shape, seed=n)
    h = 1.0 / (n - 1)
    want = compact.derivative_banded(f, h, axis)
    fd = torch.tensor(f, device=dev)
    plan = Plan(n, COMPACT_COEFFS)
    keep, nv = current(layout, n, precision)
    outs = []
    try:
        for v in range(nv):
            _set(layout, precision, n, v)
            out = torch.full_like(fd, float("nan"))
            core.compact_derivative(plan, fd, out, h, axis, precision)
            got = out.cpu().numpy()
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < TOL[precision], (v, err)
            outs.append(got)
    finally:
        _set(layout, precision, n, keep)
    scale = np.abs(want).max()
    for v in range(1, nv):
        assert np.abs(outs[v] - outs[0]).max() / scale < 1e-13, v


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("n", [64, 100, 256, 512, 768, 1024, 2048])
@pytest.mark.parametrize("axis", [1, 0])
def test_every_strided_variant_fp32_storage(dev, precision, n, axis):
    """FP32-stored fields along strided lines (the 32-line tile kernel for FP32
    rows among them; ragged last tile with 40 lines): every variant equals the
    oracle's derivative of the same FP32 values."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.autotune import _set, current
    from ntamd.workloads import compact_field
    shape = (3, n, 40) if axis == 1 else (n, 2, 40)
    f32 = compact_field(shape, seed=n + 7).astype(np.float32)
    h = 1.0 / (n - 1)
    want = compact.derivative_banded(f32.astype(np.float64), h, axis)
    fd = torch.tensor(f32, device=dev)
    plan = Plan(n, COMPACT_COEFFS)
    keep, nv = current(1, n, precision)
    try:
        for v in range(nv):
            _set(1, precision, n, v)
            out = torch.full(shape, float("nan"), dtype=torch.float64, device=dev)
            core.compact_derivative(plan, fd, out, h, axis, precision)
            err = np.abs(out.cpu().numpy() - want).max() / np.abs(want).max()
            assert err < TOL[precision], (v, err)
    finally:
        _set(1, precision, n, keep)


@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("n", [64, 96, 99, 100, 128, 256, 512, 768, 1024, 2048])
def test_every_contiguous_variant_fp32_storage(dev, precision, n):
    """FP32-stored fields along contiguous x-lines, shape (3, 40, n): every
    contiguous launch variant (wave-coalesced, direct-chunk and staged kernels,
    with FP32 input rows; n = 99 leaves every row unaligned to 16 bytes) equals
    the oracle's derivative of the same FP32 values in FP64.  This is the
    reference's RHS + CR path (compact/rhs.py:21-36, neato.py:34-123) on FP32
    storage."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.autotune import _set, current
    from ntamd.workloads import compact_field
    shape = (3, 40, n)
    f32 = compact_field(shape, seed=n + 11).astype(np.float32)
    h = 1.0 / (n - 1)
    want = compact.derivative_banded(f32.astype(np.float64), h, 2)
    fd = torch.tensor(f32, device=dev)
    plan = Plan(n, COMPACT_COEFFS)
    keep, nv = current(0, n, precision)
    outs = []
    try:
        for v in range(nv):
            _set(0, precision, n, v)
            out = torch.full(shape, float("nan"), dtype=torch.float64, device=dev)
            core.compact_derivative(plan, fd, out, h, 2, precision)
            got = out.cpu().numpy()
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err < TOL[precision], (v, err)
            outs.append(got)
    finally:
        _set(0, precision, n, keep)
    scale = np.abs(want).max()
    for v in range(1, nv):
        assert np.abs(outs[v] - outs[0]).max() / scale < TOL[precision], v


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("precision", ["fp64", "fp32_refine"])
@pytest.mark.parametrize("n", [48, 100, 128, 256, 384, 512, 1024, 2048])
def test_many_lines_every_contiguous_variant(dev, precision, n, storage):
    """More lines than the device holds wavefronts of the persistent kernel at
    once (9001 lines: every wavefront of contiguous variant 4 walks several
    lines, with the next one prefetched), and a line count that leaves the
    last wavefront of the segmented shapes (2 or 4 lines per wavefront at 48,
    100, 128, 256 and 384 points) partly empty: every contiguous variant
    against the oracle on 300 sampled lines (compact/rhs.py + neato.py solve,
    FP64 or FP32 storage)."""
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.autotune import _set, current
    from ntamd.workloads import compact_field
    shape = (1, 9001, n)
    f = compact_field(shape, seed=n + 3)
    if storage == "f32":
        f = f.astype(np.float32)
    h = 1.0 / (n - 1)
    rows = np.random.default_rng(n).choice(9001, 300, replace=False)
    rows[:2] = (0, 9000)
    want = compact.derivative_banded(f[:, rows, :].astype(np.float64), h, 2)
    fd = torch.tensor(f, device=dev)
    plan = Plan(n, COMPACT_COEFFS)
    keep, nv = current(0, n, precision)
    try:
        for v in range(nv):
            _set(0, precision, n, v)
            out = torch.full(shape, float("nan"), dtype=torch.float64, device=dev)
            core.compact_derivative(plan, fd, out, h, 2, precision)
            got = out.cpu().numpy()
            assert np.isfinite(got).all(), v
            err = np.abs(got[:, rows, :] - want).max() / np.abs(want).max()
            assert err < TOL[precision], (v, err)
    finally:
        _set(0, precision, n, keep)


@pytest.mark.parametrize("N", [34, 130, 256, 300, 512, 700, 1024, 2048])
@pytest.mark.parametrize("layout", [2, 3])
def test_every_variant_adi(dev, N, layout):
    """ADI steps under every launch shape of one ADI tuning layout (2: the
    x-implicit half step on contiguous lines, 3: the y-implicit half step on
    strided lines, with its halo lines), the other layout at its default, equal
    the oracle's Peaceman-Rachford step."""
    from ntamd import adi_coeffs, Plan, core
    from ntamd.autotune import _set, current
    from oracle import adi
    r = 0.7
    u0 = adi.gaussian_ic(N, seed=N)
    want = adi.pr_run(u0, r, 1)
    plan = Plan(N, adi_coeffs(r))
    keep, nv = current(layout, N, "fp64")
    try:
        for v in range(nv):
            _set(layout, "fp64", N, v)
            u = torch.tensor(u0, device=dev)
            w = torch.empty_like(u)
            core.adi_prepare(u, w)               # the work buffer carries u's boundary values
            core.adi_steps(plan, u, w, r, 1, "fp64", 1)
            err = np.abs(u.cpu().numpy() - want).max() / np.abs(want).max()
            assert err < 1e-12, (v, err)
    finally:
        _set(layout, "fp64", N, keep)


def test_autotune_times_all_four_layouts(dev):
    """ntamd.autotune.tune() on one small length: it times every variant of the
    four tuning layouts (derivatives on layouts 0/1, ADI steps on 2/3), applies
    the fastest, and reports each choice; the table is put back afterwards so
    the tests after this one see the defaults."""
    from ntamd import autotune
    lengths = (64, 48)
    keep = {(lay, p, n): autotune.current(lay, n, p)[0]
            for lay in range(4) for p in ("fp64", "fp32_refine") for n in lengths}
    try:
        res = autotune.tune(ns=(64,), points=1 << 16, reps=1)
        for (lay, p, n) in keep:
            entry = res[f"{lay}:{p}:{n}"]
            nv = autotune.current(lay, n, p)[1]
            assert len(entry["seconds"]) == nv and 0 <= entry["variant"] < nv
            assert autotune.current(lay, n, p)[0] == entry["variant"]
            assert all(t > 0 for t in entry["seconds"])
    finally:
        for (lay, p, n), v in keep.items():
            autotune._set(lay, p, n, v)
