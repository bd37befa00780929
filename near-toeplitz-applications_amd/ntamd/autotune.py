"""Launch-shape autotuner for the line kernels.

For each layout, precision mode and line-length bucket (powers of two in a
slot of their own) the library has several equivalent launch shapes
(``nt_tune_set``): contiguous lines can stream through a wave-private
LDS transpose, load per-lane chunks directly, or go through a workgroup tile;
strided lines can use tiles of 16 or of 32 adjacent lines, or the persistent
whole-tile kernels.  Layouts 0 and 1 serve the solves and derivatives, 2 and 3
the ADI half steps (x-implicit on contiguous lines, y-implicit on strided ones).  ``tune()`` times
every variant on a synthetic field of the given shape class, keeps the fastest,
and ``save()``/``apply()`` persist the choice in a JSON cache keyed by device
name and library version.  Variants give the same results (the strided ones to
within one rounding), so tuning never changes numerics beyond that.
"""
import ctypes
import json
import os

import torch

from . import _lib
from .core import Plan, adi_prepare, adi_steps, compact_derivative, precision_code

__all__ = ["tune", "apply", "save", "load", "current", "DEFAULT_CACHE"]

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
COMPACT = (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0)


def _device_key():
    name = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "cpu"
    return f"{name}|abi{_lib.load().nt_version()}"


def current(layout, n, precision="fp64"):
    """(variant, number of variants) in use for lines of length n."""
    v, nv = ctypes.c_int(), ctypes.c_int()
    _lib.check(_lib.load().nt_tune_get(int(layout), precision_code(precision), ctypes.c_int64(int(n)),
                                       ctypes.byref(v), ctypes.byref(nv)))
    return v.value, nv.value


def _set(layout, precision, n, variant):
    _lib.check(_lib.load().nt_tune_set(int(layout), precision_code(precision), ctypes.c_int64(int(n)), int(variant)))


def _time(fn, reps):
    for _ in range(2):
        fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps * 1e-3


def tune(ns=(64, 128, 256, 512, 1024, 2048), points=1 << 27, reps=10, precisions=("fp64", "fp32_refine"),
         verbose=False, adi=True):
    """Time every variant for contiguous (layout 0) and strided (layout 1) lines
    at each precision mode, on fields of about ``points`` points, and apply the
    fastest; with ``adi``, also the ADI half steps (layouts 2 and 3) on n x n
    fields.  The table has two slots per length bucket: powers of two (timed at
    n, where the levels are known at compile time) and other lengths (timed at
    3n/4, the general path).  Returns
    ``{"<layout>:<precision>:<n>": {"variant": v, "seconds": [...]}}``."""
    dev = torch.device("cuda")
    result = {}
    for n in ns:
        for length in ([n] if n < 16 else [n, 3 * n // 4]):
            # the other two extents: multiples of 64, as in real grids (aligned
            # rows let the strided kernels use 16-byte row loads)
            other = max(64, int((points / length) ** 0.5) // 64 * 64)
            plan = Plan(length, COMPACT)
            for layout, axis, shape in ((0, 2, (other, other, length)), (1, 1, (other, length, other))):
                f = torch.randn(shape, dtype=torch.float64, device=dev)
                out = torch.empty_like(f)
                for prec in precisions:
                    _, nv = current(layout, length, prec)
                    times = []
                    for v in range(nv):
                        _set(layout, prec, length, v)
                        times.append(_time(lambda: compact_derivative(plan, f, out, 1.0 / (length - 1), axis, prec),
                                           reps))
                    best = min(range(nv), key=lambda i: times[i])
                    _set(layout, prec, length, best)
                    result[f"{layout}:{prec}:{length}"] = {"variant": best, "seconds": times}
                    if verbose:
                        print(f"layout {layout} {prec} n {length}: " +
                              ", ".join(f"v{i} {t * 1e6:.1f}us" for i, t in enumerate(times)), flush=True)
                del f, out
    if adi:
        from . import adi_coeffs
        for n in ns:
            for length in ([n] if n < 16 else [n, 3 * n // 4]):
                if length < 8:
                    continue
                plan = Plan(length, adi_coeffs(0.5))
                u = torch.randn((length, length), dtype=torch.float64, device=dev)
                w = torch.empty_like(u)
                adi_prepare(u, w)
                for layout in (2, 3):
                    for prec in precisions:
                        _, nv = current(layout, length, prec)
                        times = []
                        for v in range(nv):
                            _set(layout, prec, length, v)
                            times.append(_time(lambda: adi_steps(plan, u, w, 0.5, 4, prec), reps) / 4)
                        best = min(range(nv), key=lambda i: times[i])
                        _set(layout, prec, length, best)
                        result[f"{layout}:{prec}:{length}"] = {"variant": best, "seconds": times}
                        if verbose:
                            print(f"layout {layout} (ADI) {prec} n {length}: " +
                                  ", ".join(f"v{i} {t * 1e6:.1f}us" for i, t in enumerate(times)), flush=True)
                del u, w
    return result


def save(result, path=DEFAULT_CACHE):
    data = {}
    if os.path.exists(path):
        with open(path) as fh:
            data = json.load(fh)
    data[_device_key()] = {k: v["variant"] for k, v in result.items()}
    with open(path, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)


def load(path=DEFAULT_CACHE):
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh).get(_device_key(), {})


def apply(path=DEFAULT_CACHE):
    """Set the cached variants for this device; returns what was applied."""
    table = load(path)
    for key, v in table.items():
        layout, prec, n = key.split(":")
        _set(int(layout), prec, int(n), int(v))
    return table
