"""Host side of the partitioned-solve plan (no GPU): secondary solutions and errors."""
import numpy as np
import pytest

from oracle import adi, partition

COMPACT = (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0)


@pytest.mark.parametrize("coeffs", [COMPACT, adi.pr_coeffs(0.5), (2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7)])
@pytest.mark.parametrize("n,P", [(2048, 16), (512, 4), (96, 3), (64, 1), (100, 2)])
def test_secondary_solutions_match_misc_py(coeffs, n, P):
    from ntamd.partition import PartitionPlan
    plan = PartitionPlan(n, P, coeffs)
    m = n // P
    for p in range(P):
        xu, xl = plan.secondary(p)
        wu, wl = partition.solve_secondary_systems(m, p, P, coeffs)
        np.testing.assert_allclose(xu, wu, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(xl, wl, rtol=1e-12, atol=1e-15)


def test_partition_plan_rejects_bad_shapes():
    from ntamd.partition import PartitionPlan
    from ntamd import NtInvalidArgument
    with pytest.raises(NtInvalidArgument):
        PartitionPlan(100, 3, COMPACT)          # not divisible
    with pytest.raises(NtInvalidArgument):
        PartitionPlan(2048, 32, COMPACT)        # P > 16
    with pytest.raises(NtInvalidArgument):
        PartitionPlan(4096, 16, COMPACT)        # n > 2048
    with pytest.raises(NtInvalidArgument):
        PartitionPlan(24, 8, COMPACT)           # m = 3 < 4


def test_part_solve_axis_validates_before_any_launch():
    """nt_part_solve_axis refuses bad arguments before it touches a device:
    axis, extents, producer, line length against the plan, the ADI field shape,
    aliasing, and in-place solves (addresses only: nothing is dereferenced)."""
    import ctypes
    from ntamd import _lib
    from ntamd.partition import PartitionPlan
    lib = _lib.load()
    plan = PartitionPlan(64, 4, COMPACT)
    h = plan._h
    a, b = ctypes.c_void_p(1 << 32), ctypes.c_void_p(1 << 33)
    cases = [
        (0, a, a, 4, 64, 8, 3, 0.0, "axis"),            # axis 3
        (0, a, a, 4, 0, 8, 1, 0.0, "extents"),          # ny = 0
        (5, a, b, 4, 64, 8, 1, 0.1, "producer"),        # producer 5
        (1, a, b, 4, 63, 8, 1, 0.1, "line length"),     # n != plan n
        (2, a, b, 2, 64, 64, 1, 0.5, "ADI"),            # ADI needs nz == 1
        (1, a, a, 4, 64, 8, 1, 0.1, "alias"),           # derivative in place
        (0, a, b, 4, 64, 8, 1, 0.0, "in place"),        # solve out of place
    ]
    for producer, src, dst, nz, ny, nx, axis, p1, what in cases:
        rc = lib.nt_part_solve_axis(h, producer, src, dst, ctypes.c_int64(nz), ctypes.c_int64(ny),
                                    ctypes.c_int64(nx), axis, ctypes.c_double(p1), None)
        assert rc == _lib.NT_EINVAL, what
    with pytest.raises(_lib.NtInvalidArgument):
        lib_call = _lib.call
        lib_call("nt_part_solve_axis", None, 0, a, a, ctypes.c_int64(1), ctypes.c_int64(1),
                 ctypes.c_int64(64), 2, ctypes.c_double(0.0), None)
