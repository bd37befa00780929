"""Write the golden fixtures of tests/golden/ (inputs + expected outputs).

Expected outputs come from the SciPy banded solve (LAPACK gtsv) and closed
forms, with inputs from seeded NumPy generators.  Nothing from the reference
is imported or run: the reference is Python 2 + PyCUDA and ships no fixtures.

  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import adi, compact, cr, partition  # noqa: E402

COMPACT = (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0)


def field64(seed=0):
    N = 64
    x = np.linspace(0.0, 1.0, N)
    f = (np.sin(2 * np.pi * x)[None, None, :] * np.sin(2 * np.pi * x)[None, :, None]
         * np.sin(2 * np.pi * x)[:, None, None])
    return f + np.random.default_rng(seed).normal(0.0, 1e-3, size=(N, N, N))


def main():
    rng = np.random.default_rng(2024)
    # 1. line solves: several sizes and matrices, two RHS each
    cases = [("compact", COMPACT, n) for n in (4, 8, 13, 64, 96, 512)]
    cases += [("adi_r05", adi.pr_coeffs(0.5), n) for n in (7, 100, 2048)]
    cases += [("general", (2.0, 0.3, 0.5, 3.0, 0.7, 0.2, 1.7), n) for n in (2, 33, 257)]
    out = {}
    for i, (name, co, n) in enumerate(cases):
        d = rng.standard_normal((2, n))
        out[f"c{i}_name"] = np.array(name)
        out[f"c{i}_coeffs"] = np.array(co, dtype=np.float64)
        out[f"c{i}_d"] = d
        out[f"c{i}_x"] = cr.banded_solve(co, d)
    out["ncases"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "cr_lines.npz"), **out)

    # 2. config 1 (64^3, seed 0): sampled x-, y- and z-lines of input and d/dx, d/dy, d/dz
    f = field64(0)
    h = 1.0 / 63
    idx = rng.integers(0, 64, size=(24, 2))
    g = {"idx": idx, "h": np.array(h)}
    for axis in (2, 1, 0):
        d = compact.derivative_banded(f, h, axis)
        lines_f, lines_d = [], []
        for i, j in idx:
            sl = [int(i), int(j)]
            sl.insert(axis, slice(None))
            lines_f.append(f[tuple(sl)])
            lines_d.append(d[tuple(sl)])
        g[f"f_axis{axis}"] = np.array(lines_f)
        g[f"d_axis{axis}"] = np.array(lines_d)
        g[f"sum_axis{axis}"] = np.array(d.sum())
        g[f"sumsq_axis{axis}"] = np.array((d * d).sum())
    np.savez_compressed(os.path.join(HERE, "compact64.npz"), **g)

    # 3. ADI: 66 x 66 Gaussian + U[0,1e-3) (seed 1), r = 0.5, 5 steps
    u0 = adi.gaussian_ic(66, seed=1)
    np.savez_compressed(os.path.join(HERE, "adi66.npz"), u0=u0, r=np.array(0.5),
                        steps=np.array(5), u=adi.pr_run(u0, 0.5, 5, solver=cr.banded_solve))

    # 4. partition method: 2 lines of 2048 cut into 16 subdomains (ADI matrix r = 0.5)
    co = adi.pr_coeffs(0.5)
    d = rng.standard_normal((2, 2048))
    np.savez_compressed(os.path.join(HERE, "partition2048.npz"), d=d, coeffs=np.array(co),
                        P=np.array(16), x=cr.banded_solve(co, d),
                        xR0=cr.banded_solve(partition.subdomain_coeffs(co, 0, 16), d[:, :128]))
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
