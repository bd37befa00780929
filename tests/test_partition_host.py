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
