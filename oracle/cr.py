"""Near-Toeplitz cyclic reduction, restated on the CPU in NumPy FP64.

TEST INFRASTRUCTURE ONLY (see the package docstring).

A near-Toeplitz tridiagonal matrix of size n is described by the 7-tuple
``(b1, c1, ai, bi, ci, an, bn)`` (reference ``neato.py:136-138,175-177``):
first row ``[b1, c1]``, interior rows ``(ai, bi, ci)``, last row ``[an, bn]``.

Cyclic reduction (CR) eliminates the even-numbered equations of the current
level, keeping the odd ones (reference ``neato.py:34-51`` keeps rows 2t+1 of
level 0).  After one reduction the kept equations again form a near-Toeplitz
system, so every level is described by the same 7 numbers plus its equation
count m.  The reference restricts n to powers of two (``neato.py:145``) and
therefore always has m even; ``level_systems`` below also covers odd m (the
last equation of the level is then eliminated, and the last kept equation
picks up a special right factor ``c/bL``).  Reduction stops when m <= 3 and the
top system is solved directly (2x2 Cramer as in ``neato.py:78-93``, or a 3x3
elimination).
"""
import numpy as np
from scipy.linalg import solve_banded

__all__ = [
    "reference_coefficients", "level_systems", "cr_solve", "banded_solve",
    "dense_matrix", "tridiagonal_solve", "matvec",
]


def reference_coefficients(system_size, coeffs):
    """The eight level arrays of the reference, restated for power-of-2 n.

    Follows ``neato.py:199-274`` literally: arrays of length log2(n), entries
    ``0 .. log2(n)-2`` hold the factors of each reduction, and the last slot of
    ``a`` and ``b`` holds the final last-row values (``neato.py:270-272``).
    Returns ``(a, b, c, k1, k2, b_first, k1_first, k1_last)``.
    """
    n = int(system_size)
    nlog = int(round(np.log2(n)))
    if 2 ** nlog != n or n < 4:
        raise ValueError("reference coefficients need a power-of-2 n >= 4")
    b1, c1, ai, bi, ci, an, bn = [float(v) for v in coeffs]
    out = {k: np.zeros(nlog) for k in
           ("a", "b", "c", "k1", "k2", "b_first", "k1_first", "k1_last")}
    a_last = b_last = 0.0
    for i in range(nlog - 1):
        if i == 0:
            pa, pb, pc, pbf, pcf = ai, bi, ci, b1, c1
            out["k1_last"][i] = an / bi
            a_last = -ai * out["k1_last"][i]
            b_last = bn - ci * out["k1_last"][i]
        else:
            pa, pb, pc = out["a"][i - 1], out["b"][i - 1], out["c"][i - 1]
            pbf, pcf = out["b_first"][i - 1], out["c"][i - 1]
            out["k1_last"][i] = a_last / pb
            a_last, b_last = (-pa * out["k1_last"][i],
                              b_last - pc * out["k1_last"][i])
        out["k1"][i] = pa / pb
        out["k2"][i] = pc / pb
        out["a"][i] = -pa * out["k1"][i]
        out["b"][i] = pb - pc * out["k1"][i] - pa * out["k2"][i]
        out["c"][i] = -pc * out["k2"][i]
        out["k1_first"][i] = pa / pbf
        out["b_first"][i] = pb - pcf * out["k1_first"][i] - pa * out["k2"][i]
    out["a"][-1] = a_last
    out["b"][-1] = b_last
    return tuple(out[k] for k in
                 ("a", "b", "c", "k1", "k2", "b_first", "k1_first", "k1_last"))


def level_systems(n, coeffs):
    """Near-Toeplitz description of every CR level, for any n >= 2.

    Returns a list of dicts with keys ``m, bF, cF, a, b, c, aL, bL``; entry 0 is
    the input matrix and entry s is the system left after reduction s.  The
    list ends at the first level with m <= 3.
    """
    n = int(n)
    if n < 2:
        raise ValueError("n must be >= 2")
    b1, c1, ai, bi, ci, an, bn = [float(v) for v in coeffs]
    lv = [dict(m=n, bF=b1, cF=c1, a=ai, b=bi, c=ci, aL=an, bL=bn)]
    while lv[-1]["m"] >= 4:
        p = lv[-1]
        m, a, b, c = p["m"], p["a"], p["b"], p["c"]
        k1, k2, k1f = a / b, c / b, a / p["bF"]
        na, nb, nc = -k1 * a, b - k1 * c - k2 * a, -k2 * c
        nbf = b - k1f * p["cF"] - k2 * a
        if m % 2 == 0:          # the last row is kept; its left neighbour is interior
            k1l = p["aL"] / b
            nal, nbl = -k1l * a, p["bL"] - k1l * c
        else:                   # the last row is eliminated into row m-2
            k2l = c / p["bL"]
            nal, nbl = -k1 * a, b - k1 * c - k2l * p["aL"]
        lv.append(dict(m=m // 2, bF=nbf, cF=nc, a=na, b=nb, c=nc, aL=nal, bL=nbl))
    return lv


def _row_coeffs(L, j):
    """(a_j, b_j, c_j) of row j of level system L (vectorised over j)."""
    m = L["m"]
    j = np.asarray(j)
    a = np.where(j == m - 1, L["aL"], L["a"])
    b = np.where(j == 0, L["bF"], np.where(j == m - 1, L["bL"], L["b"]))
    c = np.where(j == 0, L["cF"], L["c"])
    a = np.where(j == 0, 0.0, a)
    c = np.where(j == m - 1, 0.0, c)
    return a, b, c


def cr_solve(coeffs, d):
    """Solve the near-Toeplitz systems ``A x = d`` row-wise by cyclic reduction.

    ``d`` has shape ``(..., n)``; each trailing line is one system.  Phase order
    follows ``neato.py:34-123``: forward reductions (the first fused with the
    load), the direct solve of the top system, back substitution level by level,
    the last substitution (level 0 even rows) at store time.  Works in place on
    a FP64 copy, one original row index per equation, as the reference does.
    """
    d = np.array(d, dtype=np.float64, copy=True)
    n = d.shape[-1]
    lv = level_systems(n, coeffs)
    top = len(lv) - 1
    # forward reduction s: level s-1 -> s.  Level-(s-1) equation j is row h*(j+1)-1.
    for s in range(1, top + 1):
        L = lv[s - 1]
        m, h = L["m"], 2 ** (s - 1)
        j = np.arange(1, m, 2)                          # kept equations
        rows = h * (j + 1) - 1
        a, b, c = _row_coeffs(L, j)
        al, bl, cl = _row_coeffs(L, j - 1)
        k1 = a / bl
        jr = np.minimum(j + 1, m - 1)
        ar, br, cr = _row_coeffs(L, jr)
        k2 = np.where(j + 1 <= m - 1, c / br, 0.0)
        right = np.where(j + 1 <= m - 1, rows + h, rows)    # dummy index when absent
        d[..., rows] = d[..., rows] - k1 * d[..., rows - h] - k2 * d[..., right]
    # top system
    T = lv[top]
    H = 2 ** top
    m = T["m"]
    r = [H * (j + 1) - 1 for j in range(m)]
    if m == 1:
        d[..., r[0]] = d[..., r[0]] / T["bF"]
    elif m == 2:
        det = T["bF"] * T["bL"] - T["cF"] * T["aL"]
        d0, d1 = d[..., r[0]].copy(), d[..., r[1]].copy()
        d[..., r[0]] = (d0 * T["bL"] - T["cF"] * d1) / det
        d[..., r[1]] = (T["bF"] * d1 - d0 * T["aL"]) / det
    else:  # m == 3: eliminate rows 0 and 2 into row 1
        d0, d1, d2 = (d[..., r[0]].copy(), d[..., r[1]].copy(), d[..., r[2]].copy())
        k1f, k2l = T["a"] / T["bF"], T["c"] / T["bL"]
        x1 = (d1 - k1f * d0 - k2l * d2) / (T["b"] - k1f * T["cF"] - k2l * T["aL"])
        d[..., r[1]] = x1
        d[..., r[0]] = (d0 - T["cF"] * x1) / T["bF"]
        d[..., r[2]] = (d2 - T["aL"] * x1) / T["bL"]
    # back substitution: level s-1 even equations from their kept neighbours
    for s in range(top, 0, -1):
        L = lv[s - 1]
        m, h = L["m"], 2 ** (s - 1)
        j = np.arange(0, m, 2)
        rows = h * (j + 1) - 1
        a, b, c = _row_coeffs(L, j)
        left = np.where(j >= 1, rows - h, rows)
        right = np.where(j + 1 <= m - 1, rows + h, rows)
        xl = np.where(j >= 1, d[..., left], 0.0)
        xr = np.where(j + 1 <= m - 1, d[..., right], 0.0)
        d[..., rows] = (d[..., rows] - a * xl - c * xr) / b
    return d


def dense_matrix(n, coeffs):
    """Dense n x n near-Toeplitz matrix from the 7-tuple."""
    b1, c1, ai, bi, ci, an, bn = [float(v) for v in coeffs]
    A = np.zeros((n, n))
    i = np.arange(1, n - 1)
    A[i, i - 1], A[i, i], A[i, i + 1] = ai, bi, ci
    A[0, 0], A[0, 1] = b1, c1
    A[n - 1, n - 2], A[n - 1, n - 1] = an, bn
    return A


def tridiagonal_solve(a, b, c, rhs):
    """SciPy banded solve (LAPACK gtsv path), as ``parallelcompact/misc.py:5-19``.

    ``a`` lower diagonal (first entry ignored), ``b`` diagonal, ``c`` upper
    diagonal (last entry ignored); ``rhs`` of shape (n,) or (n, k).
    """
    ab = np.vstack([np.append(0.0, c[:-1]), b, np.append(a[1:], 0.0)])
    return solve_banded((1, 1), ab, rhs)


def banded_solve(coeffs, d):
    """Solve every trailing line of ``d`` with one multi-RHS SciPy banded call."""
    d = np.asarray(d, dtype=np.float64)
    n = d.shape[-1]
    b1, c1, ai, bi, ci, an, bn = [float(v) for v in coeffs]
    a = np.full(n, ai); b = np.full(n, bi); c = np.full(n, ci)
    b[0], c[0] = b1, c1
    a[-1], b[-1] = an, bn
    flat = d.reshape(-1, n).T
    x = tridiagonal_solve(a, b, c, flat)
    return np.ascontiguousarray(x.T).reshape(d.shape)


def matvec(coeffs, x):
    """``A x`` for the near-Toeplitz matrix, along the trailing axis."""
    b1, c1, ai, bi, ci, an, bn = [float(v) for v in coeffs]
    x = np.asarray(x, dtype=np.float64)
    y = np.empty_like(x)
    y[..., 1:-1] = ai * x[..., :-2] + bi * x[..., 1:-1] + ci * x[..., 2:]
    y[..., 0] = b1 * x[..., 0] + c1 * x[..., 1]
    y[..., -1] = an * x[..., -2] + bn * x[..., -1]
    return y
