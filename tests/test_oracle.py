"""Pins of the CPU oracle (no GPU): each oracle function against SciPy, brute
force, closed forms or the reference's own known answers."""
import numpy as np
import pytest

from oracle import adi, compact, cr, partition

COEFF_SETS = {
    "compact": (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0),
    "adi_r05": adi.pr_coeffs(0.5),
    "adi_ref": adi.reference_coeffs(1.0 / 63, 0.1),
    "general": (2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7),
    "neg": (-1.5, 0.4, 0.3, -2.2, 0.6, -0.5, 1.9),
}
SIZES = [2, 4, 5, 6, 7, 8, 9, 11, 13, 16, 17, 31, 32, 33, 63, 64, 65, 96, 100, 127, 128,
         160, 255, 512, 513, 1000, 1023, 2048]


def _nonsingular(n, co):
    A = cr.dense_matrix(n, co)
    return np.linalg.cond(A) < 1e12


# The reference's unscaled ADI matrix (run-adi.py:50-53) has identity end rows
# next to interior rows of size ~8e3, condition number ~4e4: LAPACK gtsv's
# pivoting then loses ~1e-12 on the end values, so its tolerance is wider.
TOL = {"adi_ref": 1e-10}


@pytest.mark.parametrize("name", sorted(COEFF_SETS))
def test_cr_matches_scipy_banded(name):
    co = COEFF_SETS[name]
    tol = TOL.get(name, 1e-12)
    rng = np.random.default_rng(11)
    for n in SIZES:
        if n <= 64 and not _nonsingular(n, co):
            continue
        d = rng.standard_normal((6, n))
        x = cr.cr_solve(co, d)
        xr = cr.banded_solve(co, d)
        err = np.abs(x - xr).max() / np.abs(xr).max()
        assert err < tol, (name, n, err)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 16])
def test_cr_matches_dense_brute_force(n):
    co = COEFF_SETS["general"]
    rng = np.random.default_rng(n)
    d = rng.standard_normal((3, n))
    x = cr.cr_solve(co, d)
    xd = np.linalg.solve(cr.dense_matrix(n, co), d.T).T
    assert np.abs(x - xd).max() < 1e-13 * max(1.0, np.abs(xd).max())


def test_cr_residual_is_tiny():
    co = COEFF_SETS["compact"]
    d = np.random.default_rng(3).standard_normal((4, 512))
    x = cr.cr_solve(co, d)
    assert np.abs(cr.matvec(co, x) - d).max() < 1e-13


def test_reference_coefficients_hand_values():
    # compact matrix (1, 2, 1/4, 1, 1/4, 2, 1): level-0 factors by hand
    a, b, c, k1, k2, bf, k1f, k1l = cr.reference_coefficients(64, COEFF_SETS["compact"])
    assert k1[0] == pytest.approx(0.25) and k2[0] == pytest.approx(0.25)
    assert a[0] == pytest.approx(-1 / 16) and c[0] == pytest.approx(-1 / 16)
    assert b[0] == pytest.approx(7 / 8)
    assert bf[0] == pytest.approx(7 / 16)
    assert k1f[0] == pytest.approx(0.25)
    assert k1l[0] == pytest.approx(2.0)
    a4, b4, *_ = cr.reference_coefficients(4, COEFF_SETS["compact"])
    assert a4[-1] == pytest.approx(-0.5) and b4[-1] == pytest.approx(0.5)   # a_last, b_last


@pytest.mark.parametrize("n", [4, 8, 64, 512, 2048])
@pytest.mark.parametrize("name", ["compact", "adi_r05", "general"])
def test_level_systems_reproduce_reference_arrays(n, name):
    co = COEFF_SETS[name]
    a, b, c, k1, k2, bf, k1f, k1l = cr.reference_coefficients(n, co)
    lv = cr.level_systems(n, co)
    assert len(lv) == int(np.log2(n))            # levels 0 .. log2(n)-1
    for i in range(len(lv) - 1):
        s, t = lv[i], lv[i + 1]
        np.testing.assert_allclose(
            [k1[i], k2[i], k1f[i], k1l[i], a[i], b[i], c[i], bf[i]],
            [s["a"] / s["b"], s["c"] / s["b"], s["a"] / s["bF"], s["aL"] / s["b"],
             t["a"], t["b"], t["c"], t["bF"]], rtol=1e-14, atol=1e-300)
    assert a[-1] == pytest.approx(lv[-1]["aL"], rel=1e-14)
    assert b[-1] == pytest.approx(lv[-1]["bL"], rel=1e-14)


@pytest.mark.parametrize("n", [8, 9, 10, 11, 16, 17, 33])
def test_level_systems_match_numerical_elimination(n):
    """One CR step applied to the dense matrix yields the next level's matrix."""
    co = COEFF_SETS["general"]
    lv = cr.level_systems(n, co)
    A = cr.dense_matrix(n, co)
    for s in range(1, len(lv)):
        m = A.shape[0]
        keep = np.arange(1, m, 2)
        B = np.zeros((len(keep), len(keep)))
        for r, j in enumerate(keep):
            row = A[j].copy()
            for nb in (j - 1, j + 1):
                if 0 <= nb < m and A[j, nb] != 0:
                    row = row - (A[j, nb] / A[nb, nb]) * A[nb]
            B[r] = row[keep]
        L = lv[s]
        k = len(keep)
        assert k == L["m"]
        assert B[0, 0] == pytest.approx(L["bF"], rel=1e-13)
        assert B[-1, -1] == pytest.approx(L["bL"], rel=1e-13)
        if k >= 3:
            assert B[1, 0] == pytest.approx(L["a"], rel=1e-13)
            assert B[1, 1] == pytest.approx(L["b"], rel=1e-13)
            assert B[1, 2] == pytest.approx(L["c"], rel=1e-13)
        if k >= 2:
            assert B[0, 1] == pytest.approx(L["cF"], rel=1e-13)
            assert B[-1, -2] == pytest.approx(L["aL"], rel=1e-13)
        A = B


def test_compact_derivative_exact_on_cubics():
    N = 17
    xs = np.linspace(0, 1, N)
    h = xs[1] - xs[0]
    Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
    f = X ** 3 - 2 * Y ** 3 + 0.5 * Z ** 3 + X * Y * Z + X ** 2
    for axis, true in [(2, 3 * X ** 2 + Y * Z + 2 * X), (1, -6 * Y ** 2 + X * Z),
                       (0, 1.5 * Z ** 2 + X * Y)]:
        np.testing.assert_allclose(compact.derivative(f, h, axis), true, atol=2e-12)


def test_compact_reference_known_answer_xyz():
    """experiments/run-parallel-compact.py:89-90,126-127,163-164: f = x*y*z."""
    N = 32
    xs = np.linspace(0, 1, N)
    Z, Y, X = np.meshgrid(xs, xs, xs, indexing="ij")
    f = X * Y * Z
    h = xs[1] - xs[0]
    np.testing.assert_allclose(compact.derivative(f, h, 2), Y * Z, rtol=1e-7, atol=1e-13)
    np.testing.assert_allclose(compact.derivative(f, h, 1), X * Z, rtol=1e-7, atol=1e-13)
    np.testing.assert_allclose(compact.derivative(f, h, 0), X * Y, rtol=1e-7, atol=1e-13)


def test_compact_convergence_orders():
    """experiments/run-compact-errors.py: interior error ratio ~16, boundary ~8."""
    errs_mid, errs_bnd = [], []
    for N in (32, 64, 128, 256):
        xs = np.linspace(2.0, 6.0, N)
        h = xs[1] - xs[0]
        f = (np.sin(xs) / xs ** 3)[None, None, :]
        true = (xs * np.cos(xs) - 3 * np.sin(xs)) / xs ** 4
        d = compact.derivative(f, h, 2)[0, 0]
        errs_mid.append(abs(d[N // 2] - true[N // 2]))
        errs_bnd.append(abs(d[0] - true[0]))
    r_mid = [errs_mid[i] / errs_mid[i + 1] for i in range(3)]
    r_bnd = [errs_bnd[i] / errs_bnd[i + 1] for i in range(3)]
    assert 12 < r_mid[-1] < 20, r_mid
    assert 6 < r_bnd[-1] < 10, r_bnd


def test_compact_matches_banded():
    f = np.random.default_rng(5).standard_normal((6, 7, 40))
    for axis in (0, 1, 2):
        if f.shape[axis] < 4:
            continue
        a = compact.derivative(f, 0.1, axis)
        b = compact.derivative_banded(f, 0.1, axis)
        assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()


def test_permute_is_numpy_transpose():
    a = np.arange(2 * 3 * 4, dtype=float).reshape(2, 3, 4)
    for p in [(0, 2, 1), (1, 2, 0), (2, 0, 1), (1, 0, 2), (2, 1, 0), (0, 1, 2)]:
        b = compact.permute(a, p)
        for idx in np.ndindex(b.shape):
            src = [0, 0, 0]
            for d, ax in enumerate(p):
                src[ax] = idx[d]
            assert b[idx] == a[tuple(src)]
    # reference round trips of run-compact.py:70,73 and :83,86
    np.testing.assert_array_equal(compact.permute(compact.permute(a, (0, 2, 1)), (0, 2, 1)), a)
    np.testing.assert_array_equal(compact.permute(compact.permute(a, (1, 2, 0)), (2, 0, 1)), a)


def test_adi_reference_form_equals_r_form():
    N = 34
    h = 1.0 / (N - 1)
    r = 0.5
    u0 = adi.gaussian_ic(N)
    a = adi.reference_run(u0, h, r * h * h, 3)
    b = adi.pr_run(u0, r, 3)
    assert np.abs(a - b).max() < 1e-13


def test_adi_half_step_matches_per_line_scipy():
    N = 20
    r = 0.7
    u = np.random.default_rng(2).random((N, N))
    v = adi.pr_step(u, r, solver=cr.cr_solve)
    w = adi.pr_step(u, r, solver=cr.banded_solve)
    assert np.abs(v - w).max() < 1e-13


@pytest.mark.parametrize("k,l,r", [(1, 1, 0.5), (3, 5, 0.5), (7, 2, 2.0), (10, 10, 0.1)])
def test_adi_eigenmode_closed_form(k, l, r):
    N = 40
    xs = np.linspace(0, 1, N)
    Y, X = np.meshgrid(xs, xs, indexing="ij")
    u = np.sin(k * np.pi * X) * np.sin(l * np.pi * Y)
    g = adi.eigen_amplification(N, k, l, r)
    v = adi.pr_run(u, r, 3)
    assert np.abs(v - g ** 3 * u).max() < 1e-12


def test_adi_dirichlet_boundary_carried():
    """run-adi.py:43-47: boundary values never change (non-zero Dirichlet data)."""
    N = 24
    u = np.full((N, N), 0.5)
    u[:, 0] = 1.0
    h = 1 / (N - 1)
    out = adi.reference_run(u, h, 0.1, 4)
    np.testing.assert_array_equal(out[:, 0], 1.0)
    np.testing.assert_array_equal(out[0, 1:], 0.5)
    assert np.abs(out - adi.pr_run(u, 0.1 / (h * h), 4)).max() < 1e-12


@pytest.mark.parametrize("P,m", [(2, 8), (4, 16), (16, 128), (3, 10), (8, 64)])
@pytest.mark.parametrize("name", ["compact", "adi_r05", "general"])
def test_partition_equals_monolithic(P, m, name):
    co = COEFF_SETS[name]
    d = np.random.default_rng(P * m).standard_normal((5, P * m))
    x = partition.partitioned_solve(co, d, P)
    xr = cr.banded_solve(co, d)
    assert np.abs(x - xr).max() < 1e-12 * np.abs(xr).max()


def test_secondary_systems_are_homogeneous_solutions():
    co = COEFF_SETS["compact"]
    for rank, size in [(0, 4), (1, 4), (3, 4)]:
        xU, xL = partition.solve_secondary_systems(32, rank, size, co)
        sub = partition.subdomain_coeffs(co, rank, size)
        A = cr.dense_matrix(32, sub)
        eU = np.zeros(32)
        eL = np.zeros(32)
        eU[0] = -(0.0 if rank == 0 else co[2])
        eL[-1] = -(0.0 if rank == size - 1 else co[4])
        np.testing.assert_allclose(A @ xU, eU, atol=1e-14)
        np.testing.assert_allclose(A @ xL, eL, atol=1e-14)


def test_partition_known_answer_xyz():
    """The reference's distributed check (f = x*y*z -> y*z) on one line cut in P pieces."""
    N, P = 64, 4
    xs = np.linspace(0, 1, N)
    h = xs[1] - xs[0]
    f = (xs * 0.3 * 0.7)[None, :]
    rhs = compact.compact_rhs(f, h)
    x = partition.partitioned_solve(compact.COMPACT_COEFFS, rhs, P)
    np.testing.assert_allclose(x, 0.3 * 0.7, rtol=1e-7)


def test_thomas_matches_scipy():
    n = 32
    rng = np.random.default_rng(9)
    a = rng.random(n); b = 3 + rng.random(n); c = rng.random(n)
    d = rng.standard_normal((n, 5))
    x = partition.thomas(a, b, c, d)
    xr = cr.tridiagonal_solve(a, b, c, d)
    np.testing.assert_allclose(x, xr, rtol=1e-12, atol=1e-14)


def test_ghosted_rhs_matches_compact_rhs():
    f = np.random.default_rng(4).standard_normal((5, 6, 12))
    g = np.zeros((7, 8, 14))
    g[1:-1, 1:-1, 1:-1] = f
    np.testing.assert_allclose(partition.ghosted_rhs(g, 0.2, 0, 1), compact.compact_rhs(f, 0.2),
                               rtol=1e-14)
