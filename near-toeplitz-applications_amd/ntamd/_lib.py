"""ctypes binding of libntoeplitz.so (C ABI declared in include/ntoeplitz.h).

The library is built in-tree by ``csrc/build.py`` into ``lib/libntoeplitz.so``.
There is no fallback: if the library is missing or fails to load, every call
raises ``NtLibraryError``.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libntoeplitz.so")

NT_OK = 0
NT_EINVAL = -1
NT_ESINGULAR = -2
NT_EHIP = -3
NT_EUNSUPPORTED = -4

NT_F64 = 0
NT_F32 = 1
NT_PREC_F64 = 0
NT_PREC_F32_REFINE = 1
NT_BATCH_NOALLOC_IMAGE_MAX = 1048320   # bytes of job image the allocation-free batch holds

# every symbol include/ntoeplitz.h declares, with its ctypes signature
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_intp = ctypes.POINTER(ctypes.c_int)

SIGNATURES = {
    "nt_last_error": (ctypes.c_char_p, []),
    "nt_version": (_int, []),
    "nt_tune_set": (_int, [_int, _int, _i64, _int]),
    "nt_tune_get": (_int, [_int, _int, _i64, _intp, _intp]),
    "nt_plan_create": (_int, [_i64, _dp, _pp]),
    "nt_plan_destroy": (_int, [_vp]),
    "nt_plan_info": (_int, [_vp, _i64p, _intp]),
    "nt_plan_levels": (_int, [_vp, _dp, _int]),
    "nt_plan_reference_arrays": (_int, [_vp, _dp, _int]),
    "nt_solve": (_int, [_vp, _vp, _int, _i64, _int, _int, _vp]),
    "nt_solve_axis": (_int, [_vp, _vp, _int, _i64, _i64, _i64, _int, _int, _int, _vp]),
    "nt_compact_derivative": (_int, [_vp, _vp, _int, _vp, _int, _i64, _i64, _i64, _int, _dbl,
                                     _int, _int, _vp]),
    "nt_compact_rhs": (_int, [_vp, _vp, _i64, _i64, _i64, _dbl, _vp]),
    "nt_permute": (_int, [_vp, _vp, _i64p, _intp, _vp]),
    "nt_adi_rhs": (_int, [_vp, _vp, _dbl, _dbl, _i64, _vp]),
    "nt_transpose": (_int, [_vp, _vp, _i64, _vp]),
    "nt_adi_prepare": (_int, [_vp, _vp, _i64, _vp]),
    "nt_adi_steps": (_int, [_vp, _vp, _vp, _i64, _dbl, _int, _int, _int, _vp]),
    "nt_part_plan_create": (_int, [_i64, _int, _dp, _pp]),
    "nt_part_plan_destroy": (_int, [_vp]),
    "nt_part_plan_secondary": (_int, [_vp, _int, _dp, _dp, _int]),
    "nt_part_solve_lines": (_int, [_vp, _int, _vp, _vp, _i64, _dbl, _dbl, _vp]),
    "nt_part_solve_axis": (_int, [_vp, _int, _vp, _vp, _i64, _i64, _i64, _int, _dbl, _vp]),
    "nt_part_rhs_ghosted": (_int, [_vp, _vp, _i64, _i64, _i64, _dbl, _int, _int, _vp]),
    "nt_part_params": (_int, [_vp, _vp, _vp, _vp, _vp, _int, _i64, _i64, _i64, _vp]),
    "nt_part_sum": (_int, [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _vp]),
    "nt_batch_image_bytes": (ctypes.c_size_t, [_int, _int]),
    "nt_batch_pack": (_int, [_vp, _int, _pp, _int, _vp, ctypes.c_size_t, _i64p]),
    "nt_batch_launch": (_int, [_vp, _int, _int, _i64, _vp, _int, _vp]),
    "nt_batch_run_noalloc": (_int, [_vp, _int, _pp, _int, _int, _vp]),
}


class NtLibraryError(RuntimeError):
    """The HIP library is missing or could not be loaded."""


class NtError(RuntimeError):
    """A library call returned an error code."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class NtInvalidArgument(NtError, ValueError):
    pass


class NtSingular(NtError, ArithmeticError):
    pass


_lib = None


def load():
    """Load (once) and return the ctypes library handle."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NtLibraryError(
            f"{LIB_PATH} is missing: build it with `python csrc/build.py` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as e:
        raise NtLibraryError(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    """Raise the matching exception for a non-zero return code."""
    if rc == NT_OK:
        return
    msg = load().nt_last_error().decode(errors="replace")
    if rc == NT_EINVAL:
        raise NtInvalidArgument(rc, msg)
    if rc == NT_ESINGULAR:
        raise NtSingular(rc, msg)
    raise NtError(rc, msg)


def call(name, *args):
    check(getattr(load(), name)(*args))
