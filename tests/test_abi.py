"""The C ABI without a GPU: the library loads, exports every declared symbol,
and its host-side plan logic (level table, reference arrays, validation)
agrees with the oracle.  No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import adi, cr
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ntoeplitz.h")


def _declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nt_[a-z0-9_]+)\s*\(", text)))


def test_library_loads_and_exports_every_header_symbol():
    from ntamd import _lib
    lib = _lib.load()
    syms = _declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), f"libntoeplitz.so does not export {s}"
    # the ctypes signature table covers the header exactly
    assert sorted(_lib.SIGNATURES) == syms


def test_version_and_error_string():
    from ntamd import _lib
    lib = _lib.load()
    assert lib.nt_version() >= 100
    h = ctypes.c_void_p()
    co = (ctypes.c_double * 7)(1, 2, .25, 1, .25, 2, 1)
    rc = lib.nt_plan_create(ctypes.c_int64(1), co, ctypes.byref(h))
    assert rc == _lib.NT_EINVAL
    assert b"out of range" in lib.nt_last_error()
    rc = lib.nt_plan_create(ctypes.c_int64(8), co, ctypes.byref(h))
    assert rc == _lib.NT_OK and h.value
    assert lib.nt_last_error() == b""
    lib.nt_plan_destroy(h)


COEFFS = {
    "compact": (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0),
    "adi": adi.pr_coeffs(0.5),
    "general": (2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7),
}


@pytest.mark.parametrize("name", sorted(COEFFS))
@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 8, 13, 64, 96, 100, 511, 512, 1000, 2048, 4096])
def test_plan_levels_match_oracle(name, n):
    from ntamd import Plan, NtSingular
    if name == "compact" and n == 3:
        with pytest.raises(NtSingular):           # det [[1,2,0],[.25,1,.25],[0,2,1]] = 0
            Plan(n, COEFFS[name])
        return
    p = Plan(n, COEFFS[name])
    lv = cr.level_systems(n, COEFFS[name])
    got = p.levels()
    assert got.shape == (len(lv), 8)
    want = np.array([[L[k] for k in ("m", "bF", "cF", "a", "b", "c", "aL", "bL")] for L in lv])
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=0)


@pytest.mark.parametrize("name", sorted(COEFFS))
@pytest.mark.parametrize("n", [4, 8, 64, 512, 2048])
def test_plan_reference_arrays_match_neato(name, n):
    from ntamd import Plan
    got = Plan(n, COEFFS[name]).reference_arrays()
    want = cr.reference_coefficients(n, COEFFS[name])
    for g, w in zip(got, want):
        np.testing.assert_allclose(g, w, rtol=1e-14, atol=1e-300)


def test_plan_rejects_bad_input():
    from ntamd import Plan, NtInvalidArgument, NtSingular
    with pytest.raises(NtInvalidArgument):
        Plan(1, COEFFS["compact"])
    with pytest.raises(NtInvalidArgument):
        Plan(5000, COEFFS["compact"])
    with pytest.raises(NtInvalidArgument):
        Plan(8, (1, 2, float("nan"), 1, 1, 1, 1))
    with pytest.raises(NtSingular):
        Plan(8, (0.0, 1.0, 1.0, 2.0, 1.0, 1.0, 1.0))      # b1 = 0 pivot
    with pytest.raises(NtSingular):
        Plan(3, COEFFS["compact"])                        # the 3x3 compact matrix is singular
    with pytest.raises(NtInvalidArgument):
        Plan(8, (1, 2, 3))


def test_reference_arrays_need_power_of_two():
    from ntamd import Plan, NtInvalidArgument
    with pytest.raises(NtInvalidArgument):
        Plan(96, COEFFS["compact"]).reference_arrays()


def test_tensor_layer_validates_before_any_launch():
    import torch
    from ntamd import Plan, core, NtInvalidArgument
    p = Plan(8, COEFFS["compact"])
    cpu = torch.zeros(4, 8, dtype=torch.float64)
    with pytest.raises(NtInvalidArgument):
        core.solve(p, cpu)                       # not a device tensor
    with pytest.raises(TypeError):
        core.solve(p, np.zeros(8))
    with pytest.raises(NtInvalidArgument):
        core.precision_code("fp16")


def test_dropin_constructor_checks():
    from neato import NearToeplitzSolver
    with pytest.raises(ValueError):
        NearToeplitzSolver(1, 4, COEFFS["compact"])
    with pytest.raises(ValueError):
        NearToeplitzSolver(4096, 4, COEFFS["compact"])
    s = NearToeplitzSolver(96, 10, COEFFS["compact"])   # non power of two is accepted
    assert s.plan.nlevels == len(cr.level_systems(96, COEFFS["compact"]))


def test_tuning_table_set_get_and_validation():
    """The launch-shape table is host state: defaults are in range, set/get
    round-trips per (layout, precision, n bucket), bad arguments are rejected."""
    import ctypes
    from ntamd import _lib
    lib = _lib.load()
    v, nv = ctypes.c_int(), ctypes.c_int()
    for layout in (0, 1, 2, 3):
        for prec in (_lib.NT_PREC_F64, _lib.NT_PREC_F32_REFINE):
            for n in (2, 64, 100, 512, 2048):
                assert lib.nt_tune_get(layout, prec, n, ctypes.byref(v), ctypes.byref(nv)) == 0
                assert 0 <= v.value < nv.value
    keep = ctypes.c_int()
    assert lib.nt_tune_get(1, 1, 700, ctypes.byref(keep), ctypes.byref(nv)) == 0
    other = (keep.value + 1) % nv.value
    p2 = ctypes.c_int()
    assert lib.nt_tune_get(1, 1, 1024, ctypes.byref(p2), None) == 0
    assert lib.nt_tune_set(1, 1, 700, other) == 0
    assert lib.nt_tune_get(1, 1, 900, ctypes.byref(v), None) == 0 and v.value == other         # same slot
    assert lib.nt_tune_get(1, 1, 1024, ctypes.byref(v), None) == 0 and v.value == p2.value       # power of two
    assert lib.nt_tune_set(1, 1, 700, keep.value) == 0
    # the ADI layouts (2, 3) have slots of their own
    adi_keep = ctypes.c_int()
    assert lib.nt_tune_get(2, 0, 2048, ctypes.byref(adi_keep), ctypes.byref(nv)) == 0 and nv.value == 4
    assert lib.nt_tune_get(0, 0, 2048, ctypes.byref(p2), None) == 0
    assert lib.nt_tune_set(2, 0, 2048, (adi_keep.value + 1) % 4) == 0
    assert lib.nt_tune_get(0, 0, 2048, ctypes.byref(v), None) == 0 and v.value == p2.value
    assert lib.nt_tune_set(2, 0, 2048, adi_keep.value) == 0
    assert lib.nt_tune_get(3, 0, 64, None, ctypes.byref(nv)) == 0 and nv.value == 4
    assert lib.nt_tune_set(4, 0, 64, 0) == _lib.NT_EINVAL
    assert lib.nt_tune_set(-1, 0, 64, 0) == _lib.NT_EINVAL
    assert lib.nt_tune_set(0, 5, 64, 0) == _lib.NT_EINVAL
    assert lib.nt_tune_set(0, 0, 1, 0) == _lib.NT_EINVAL
    assert lib.nt_tune_set(0, 0, 64, 5) == _lib.NT_EINVAL
    assert lib.nt_tune_set(2, 0, 64, 4) == _lib.NT_EINVAL
    assert lib.nt_tune_set(1, 0, 64, 6) == _lib.NT_EINVAL
    assert lib.nt_tune_set(3, 0, 64, 4) == _lib.NT_EINVAL
    assert b"variant" in lib.nt_last_error()
    assert lib.nt_tune_get(1, 0, 64, None, ctypes.byref(nv)) == 0 and nv.value == 6


def test_noalloc_batch_limits_without_a_gpu():
    """nt_batch_run_noalloc: the image limit of the header is the one the Python
    layer states, and empty or oversized batches are refused before any device
    call (so this runs without a GPU)."""
    from ntamd import _lib
    from ntamd.batch import NtBatchJob
    lib = _lib.load()
    text = open(HEADER).read()
    m = re.search(r"#define NT_BATCH_NOALLOC_IMAGE_MAX \(\(size_t\)(\d+)\)", text)
    assert m and int(m.group(1)) == _lib.NT_BATCH_NOALLOC_IMAGE_MAX
    assert _lib.NT_BATCH_NOALLOC_IMAGE_MAX + 256 == 1 << 20
    h = ctypes.c_void_p()
    co = (ctypes.c_double * 7)(1, 2, .25, 1, .25, 2, 1)
    assert lib.nt_plan_create(ctypes.c_int64(64), co, ctypes.byref(h)) == 0
    pv = (ctypes.c_void_p * 1)(h.value)
    try:
        assert lib.nt_batch_run_noalloc((NtBatchJob * 1)(), 0, pv, 1, 0, None) == _lib.NT_EINVAL
        njobs = 1
        while lib.nt_batch_image_bytes(njobs, 1) <= _lib.NT_BATCH_NOALLOC_IMAGE_MAX:
            njobs += 1000
        arr = (NtBatchJob * njobs)()
        assert lib.nt_batch_run_noalloc(arr, njobs, pv, 1, 0, None) == _lib.NT_EINVAL
        assert b"static scratch" in lib.nt_last_error()
    finally:
        lib.nt_plan_destroy(h)


def test_compact_derivative_refuses_overlapping_buffers():
    """f and out sharing any byte (not only the same base address) are refused
    before a launch: the tile kernels read one tile while storing another."""
    from ntamd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    co = (ctypes.c_double * 7)(1, 2, .25, 1, .25, 2, 1)
    assert lib.nt_plan_create(ctypes.c_int64(64), co, ctypes.byref(h)) == 0
    try:
        base = 1 << 32                      # addresses only: nothing is dereferenced
        npts = 4 * 4 * 64
        for f_dtype, esize in ((_lib.NT_F64, 8), (_lib.NT_F32, 4)):
            for off in (8, npts * esize - 8, -npts * 8 + 8):
                rc = lib.nt_compact_derivative(h, ctypes.c_void_p(base), f_dtype, ctypes.c_void_p(base + off),
                                               _lib.NT_F64, 4, 4, 64, 2, 0.1, 0, 1, None)
                assert rc == _lib.NT_EINVAL and b"overlap" in lib.nt_last_error(), (f_dtype, off)
    finally:
        lib.nt_plan_destroy(h)
