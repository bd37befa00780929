"""Thin, validating Python layer over the C ABI (torch tensors in, nothing allocated).

Device arrays are PyTorch-ROCm tensors; launches go on
``torch.cuda.current_stream()``.  Every function checks dtype, device, shape and
contiguity before the call (the reference checks none of these).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import (NT_F32, NT_F64, NT_PREC_F32_REFINE, NT_PREC_F64,  # noqa: F401
                   NtError, NtInvalidArgument, NtLibraryError, NtSingular)

PRECISIONS = {"fp64": NT_PREC_F64, "fp32_refine": NT_PREC_F32_REFINE}
_DTYPES = {torch.float64: NT_F64, torch.float32: NT_F32}
MAX_N = 2048


def precision_code(precision):
    if isinstance(precision, int):
        return precision
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise NtInvalidArgument(_lib.NT_EINVAL,
                                f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}")


def stream_handle(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def device_ptr(t, name, dtypes=(torch.float64,), contiguous=True):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor, got {type(t).__name__}")
    if not t.is_cuda:
        raise NtInvalidArgument(_lib.NT_EINVAL, f"{name} must be a device (HIP) tensor")
    if t.dtype not in dtypes:
        raise NtInvalidArgument(_lib.NT_EINVAL, f"{name} has dtype {t.dtype}, expected one of {dtypes}")
    if contiguous and not t.is_contiguous():
        raise NtInvalidArgument(_lib.NT_EINVAL, f"{name} must be C-contiguous")
    return ctypes.c_void_p(t.data_ptr())


def storage_elems_from(t):
    """Elements of t's storage from t.data_ptr() to the end of the storage."""
    st = t.untyped_storage()
    return (st.data_ptr() + st.nbytes() - t.data_ptr()) // t.element_size()


class Plan:
    """A near-Toeplitz matrix of size n: ``coeffs = (b1, c1, ai, bi, ci, an, bn)``.

    Host-side only: the FP64 level table lives in the plan and is passed to the
    kernels as a launch argument, so a plan holds no device memory.
    """

    def __init__(self, n, coeffs):
        coeffs = [float(c) for c in coeffs]
        if len(coeffs) != 7:
            raise NtInvalidArgument(_lib.NT_EINVAL, "coeffs must have 7 entries (b1, c1, ai, bi, ci, an, bn)")
        self.n = int(n)
        self.coeffs = tuple(coeffs)
        lib = _lib.load()
        h = ctypes.c_void_p()
        arr = (ctypes.c_double * 7)(*coeffs)
        _lib.check(lib.nt_plan_create(ctypes.c_int64(self.n), arr, ctypes.byref(h)))
        self._h = h
        self._lib = lib
        nl = ctypes.c_int()
        _lib.check(lib.nt_plan_info(h, None, ctypes.byref(nl)))
        self.nlevels = nl.value

    @property
    def handle(self):
        return self._h

    def levels(self):
        """(nlevels, 8) array: m, bF, cF, a, b, c, aL, bL of every CR level."""
        out = (ctypes.c_double * (8 * self.nlevels))()
        _lib.check(self._lib.nt_plan_levels(self._h, out, 8 * self.nlevels))
        return np.frombuffer(out, dtype=np.float64).reshape(self.nlevels, 8).copy()

    def reference_arrays(self):
        """The reference's (a, b, c, k1, k2, b_first, k1_first, k1_last) arrays."""
        L = int(round(np.log2(self.n)))
        out = (ctypes.c_double * (8 * L))()
        _lib.check(self._lib.nt_plan_reference_arrays(self._h, out, 8 * L))
        return tuple(np.frombuffer(out, dtype=np.float64).reshape(8, L).copy())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.nt_plan_destroy(h)
            except Exception:
                pass
            self._h = None


def solve(plan, x, nrhs=None, precision="fp64", sweeps=1, stream=None):
    """Solve ``nrhs`` contiguous systems in place starting at ``x.data_ptr()``."""
    ptr = device_ptr(x, "x", contiguous=False)
    n = plan.n
    avail = storage_elems_from(x)
    if nrhs is None:
        if not x.is_contiguous() or x.numel() % n:
            raise NtInvalidArgument(_lib.NT_EINVAL, "x must hold whole contiguous systems of length n")
        nrhs = x.numel() // n
    if nrhs * n > avail:
        raise NtInvalidArgument(
            _lib.NT_EINVAL,
            f"{nrhs} systems of length {n} need {nrhs * n} elements, only {avail} remain in x's storage")
    _lib.call("nt_solve", plan.handle, ptr, NT_F64, ctypes.c_int64(nrhs), precision_code(precision),
              int(sweeps), stream_handle(stream))
    return x


def _shape3(t, name):
    if t.dim() != 3:
        raise NtInvalidArgument(_lib.NT_EINVAL, f"{name} must be 3-D (nz, ny, nx), got shape {tuple(t.shape)}")
    return tuple(int(s) for s in t.shape)


def solve_axis(plan, x, axis, precision="fp64", sweeps=1, stream=None):
    """In-place solve of every line along ``axis`` (0=z, 1=y, 2=x) of a 3-D field."""
    ptr = device_ptr(x, "x")
    nz, ny, nx = _shape3(x, "x")
    _lib.call("nt_solve_axis", plan.handle, ptr, NT_F64, nz, ny, nx, int(axis),
              precision_code(precision), int(sweeps), stream_handle(stream))
    return x


def compact_derivative(plan, f, out, dx, axis, precision="fp64", sweeps=1, stream=None):
    """Fused 4th-order compact derivative along ``axis`` of a 3-D field."""
    fp = device_ptr(f, "f", dtypes=(torch.float64, torch.float32))
    op = device_ptr(out, "out")
    shape = _shape3(f, "f")
    if tuple(out.shape) != shape:
        raise NtInvalidArgument(_lib.NT_EINVAL, f"out shape {tuple(out.shape)} != f shape {shape}")
    _lib.call("nt_compact_derivative", plan.handle, fp, _DTYPES[f.dtype], op, NT_F64,
              shape[0], shape[1], shape[2], int(axis), float(dx), precision_code(precision),
              int(sweeps), stream_handle(stream))
    return out


def compact_rhs(f, rhs, dx, stream=None):
    fp = device_ptr(f, "f")
    rp = device_ptr(rhs, "rhs")
    shape = _shape3(f, "f")
    if tuple(rhs.shape) != shape:
        raise NtInvalidArgument(_lib.NT_EINVAL, "rhs must have f's shape")
    _lib.call("nt_compact_rhs", fp, rp, shape[0], shape[1], shape[2], float(dx), stream_handle(stream))
    return rhs


def permute(a, b, permutation, stream=None):
    ap_ = device_ptr(a, "a_d")
    bp = device_ptr(b, "b_d")
    shape = _shape3(a, "a_d")
    perm = [int(p) for p in permutation]
    if sorted(perm) != [0, 1, 2]:
        raise NtInvalidArgument(_lib.NT_EINVAL, f"bad permutation {permutation}")
    want = tuple(shape[p] for p in perm)
    if tuple(b.shape) != want:
        raise NtInvalidArgument(_lib.NT_EINVAL, f"b_d shape {tuple(b.shape)} != {want}")
    sh = (ctypes.c_int64 * 3)(*shape)
    pm = (ctypes.c_int * 3)(*perm)
    _lib.call("nt_permute", ap_, bp, sh, pm, stream_handle(stream))
    return b


def _square(t, N, name):
    if t.numel() < N * N:
        raise NtInvalidArgument(_lib.NT_EINVAL, f"{name} holds {t.numel()} elements, need {N}x{N}")


def adi_rhs(u, d, dx, dt, N, stream=None):
    up = device_ptr(u, "u")
    dp = device_ptr(d, "d")
    _square(u, N, "u")
    _square(d, N, "d")
    _lib.call("nt_adi_rhs", up, dp, float(dx), float(dt), int(N), stream_handle(stream))
    return d


def transpose(a, b, N, stream=None):
    ap_ = device_ptr(a, "a")
    bp = device_ptr(b, "b")
    _square(a, N, "a")
    _square(b, N, "b")
    _lib.call("nt_transpose", ap_, bp, int(N), stream_handle(stream))
    return b


def adi_prepare(u, work, stream=None):
    N = int(u.shape[-1])
    _square(u, N, "u")
    _square(work, N, "work")
    _lib.call("nt_adi_prepare", device_ptr(u, "u"), device_ptr(work, "work"), N, stream_handle(stream))


def adi_steps(plan, u, work, r, nsteps, precision="fp64", sweeps=1, stream=None):
    N = int(u.shape[-1])
    if u.dim() != 2 or u.shape[0] != N:
        raise NtInvalidArgument(_lib.NT_EINVAL, "u must be square N x N")
    if tuple(work.shape) != (N, N):
        raise NtInvalidArgument(_lib.NT_EINVAL, "work must be N x N")
    _lib.call("nt_adi_steps", plan.handle, device_ptr(u, "u"), device_ptr(work, "work"), N, float(r),
              int(nsteps), precision_code(precision), int(sweeps), stream_handle(stream))
    return u


# ---- partition method, one step at a time (parallelcompact/rhs.py, reduced.py, sum.py) ----

def part_rhs_ghosted(f_local, x, dx, mx, npx, stream=None):
    """Compact RHS along x of a width-1 ghosted subdomain array:
    f_local (nz+2, ny+2, nx+2) -> x (nz, ny, nx); closures on rank 0 / npx-1."""
    fp = device_ptr(f_local, "f_local")
    xp = device_ptr(x, "x")
    nz, ny, nx = _shape3(x, "x")
    if tuple(f_local.shape) != (nz + 2, ny + 2, nx + 2):
        raise NtInvalidArgument(_lib.NT_EINVAL,
                                f"f_local shape {tuple(f_local.shape)} != {(nz + 2, ny + 2, nx + 2)}")
    _lib.call("nt_part_rhs_ghosted", fp, xp, nz, ny, nx, float(dx), int(mx), int(npx), stream_handle(stream))
    return x


def part_params(xU, xL, xR, alpha, beta, stream=None):
    """alpha, beta (P, nz, ny) of the reduced system for P subdomains on this
    device: xU, xL (P, m) secondary solutions, xR (P, nz, ny, m)."""
    up, lp, rp = device_ptr(xU, "xU"), device_ptr(xL, "xL"), device_ptr(xR, "xR")
    ap_, bp = device_ptr(alpha, "alpha"), device_ptr(beta, "beta")
    if xR.dim() != 4:
        raise NtInvalidArgument(_lib.NT_EINVAL, f"xR must be (P, nz, ny, m), got shape {tuple(xR.shape)}")
    P, nz, ny, m = (int(s) for s in xR.shape)
    for t, name in ((xU, "xU"), (xL, "xL")):
        if tuple(t.shape) != (P, m):
            raise NtInvalidArgument(_lib.NT_EINVAL, f"{name} shape {tuple(t.shape)} != {(P, m)}")
    for t, name in ((alpha, "alpha"), (beta, "beta")):
        if tuple(t.shape) != (P, nz, ny):
            raise NtInvalidArgument(_lib.NT_EINVAL, f"{name} shape {tuple(t.shape)} != {(P, nz, ny)}")
    _lib.call("nt_part_params", up, lp, rp, ap_, bp, P, nz, ny, m, stream_handle(stream))
    return alpha, beta


def part_sum(xU, xL, xR, alpha, beta, stream=None):
    """xR (nz, ny, nx) += alpha[z, y] xU[x] + beta[z, y] xL[x], in place."""
    up, lp, rp = device_ptr(xU, "x_UH"), device_ptr(xL, "x_LH"), device_ptr(xR, "x_R")
    ap_, bp = device_ptr(alpha, "alpha"), device_ptr(beta, "beta")
    nz, ny, nx = _shape3(xR, "x_R")
    if xU.numel() != nx or xL.numel() != nx:
        raise NtInvalidArgument(_lib.NT_EINVAL, f"x_UH and x_LH must hold nx = {nx} values")
    if alpha.numel() != nz * ny or beta.numel() != nz * ny:
        raise NtInvalidArgument(_lib.NT_EINVAL, f"alpha and beta must hold nz*ny = {nz * ny} values")
    _lib.call("nt_part_sum", up, lp, rp, ap_, bp, nz, ny, nx, stream_handle(stream))
    return xR
