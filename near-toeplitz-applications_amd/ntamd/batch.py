"""Mixed-size persistent batches: many fields, one launch (nt_batch_* in the C ABI).

    batch = MixedBatch([dict(f=f1, out=o1, axis=2, dx=h1),
                        dict(f=f2, out=o2, axis=0, dx=h2, precision="fp32_refine", sweeps=1),
                        ...])
    batch.run()            # one persistent launch on the current stream

The job table and the level tables of the distinct plans are packed once on the
host by the library, copied to one device buffer, and reused by every run(); a
run zeroes the work counter with a kernel (graph-capturable) and launches.
Fields must stay alive, at the same addresses, for the life of the batch.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .core import Plan, device_ptr, precision_code, stream_handle

__all__ = ["NtBatchJob", "MixedBatch", "ImagePool", "job_table", "run_noalloc"]

COMPACT = (1.0, 2.0, 0.25, 1.0, 0.25, 2.0, 1.0)


class NtBatchJob(ctypes.Structure):
    """Mirror of ``nt_batch_job`` (include/ntoeplitz.h)."""
    _fields_ = [("in_", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("nz", ctypes.c_int64), ("ny", ctypes.c_int64), ("nx", ctypes.c_int64),
                ("dx", ctypes.c_double), ("axis", ctypes.c_int), ("producer", ctypes.c_int),
                ("in_dtype", ctypes.c_int), ("plan_slot", ctypes.c_int),
                ("precision", ctypes.c_int), ("sweeps", ctypes.c_int)]


def job_table(jobs):
    """The ``nt_batch_job`` array, the distinct plans and the tensors to keep
    alive for a list of job dicts (see ``MixedBatch``), validated."""
    if not jobs:
        raise ValueError("empty batch")
    keep = []
    plans, slot = [], {}
    arr = (NtBatchJob * len(jobs))()
    dev = None
    for i, j in enumerate(jobs):
        out = j["out"]
        solve = bool(j.get("solve", False))
        f = out if solve else j["f"]
        fp = device_ptr(f, "f", dtypes=(torch.float64, torch.float32))
        op = device_ptr(out, "out")
        if f.dim() != 3 or tuple(out.shape) != tuple(f.shape):
            raise _lib.NtInvalidArgument(_lib.NT_EINVAL, f"job {i}: f/out must be 3-D, same shape")
        dev = out.device if dev is None else dev
        axis = int(j["axis"])
        n = int(f.shape[axis])
        coeffs = tuple(float(c) for c in j.get("coeffs", COMPACT))
        key = (n, coeffs)
        if key not in slot:
            slot[key] = len(plans)
            plans.append(Plan(n, coeffs))
        a = arr[i]
        a.in_, a.out = fp.value, op.value
        a.nz, a.ny, a.nx = (int(s) for s in f.shape)
        a.dx = float(j.get("dx", 1.0))
        a.axis = axis
        a.producer = 0 if solve else 1
        a.in_dtype = _lib.NT_F32 if f.dtype == torch.float32 else _lib.NT_F64
        a.plan_slot = slot[key]
        a.precision = precision_code(j.get("precision", "fp64"))
        a.sweeps = int(j.get("sweeps", 1))
        keep += [f, out]
    return arr, plans, keep, dev


def run_noalloc(jobs, grid=0):
    """Run the jobs with ``nt_batch_run_noalloc`` on the current stream: the same
    kernel as ``MixedBatch.run``, with the job image and work counter in the
    library's static device array, so the call allocates no device or pinned
    memory (for callers that cannot).  Calls overlapping on one GPU race on
    that array; calls on one stream are ordered."""
    lib = _lib.load()
    arr, plans, keep, _ = job_table(jobs)
    pv = (ctypes.c_void_p * len(plans))(*[p.handle.value for p in plans])
    _lib.check(lib.nt_batch_run_noalloc(arr, len(jobs), pv, len(plans), int(grid), stream_handle()))
    return keep


class MixedBatch:
    """A fixed set of compact-derivative (or in-place solve) jobs run by one launch.

    Each job is a dict: ``f`` (3-D FP64 or FP32 device tensor), ``out`` (FP64,
    same shape), ``axis`` (0 z, 1 y, 2 x), ``dx``; optional ``precision``
    ("fp64" | "fp32_refine"), ``sweeps``, ``coeffs`` (default: the compact
    matrix), ``solve=True`` for an in-place solve of ``out`` (then ``f`` is
    ignored).  Lines must have 16..512 points.
    """

    def __init__(self, jobs, grid=0):
        lib = _lib.load()
        arr, plans, self._keep, dev = job_table(jobs)
        self.plans = plans
        pv = (ctypes.c_void_p * len(plans))(*[p.handle.value for p in plans])
        nbytes = lib.nt_batch_image_bytes(len(jobs), len(plans))
        host = np.zeros(nbytes, dtype=np.uint8)
        total = ctypes.c_int64()
        _lib.check(lib.nt_batch_pack(arr, len(jobs), pv, len(plans), host.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_size_t(nbytes), ctypes.byref(total)))
        self.total_tiles = total.value
        self.njobs, self.nplans = len(jobs), len(plans)
        # asynchronous upload from pinned memory on the current stream; the
        # pinned copy lives as long as the batch, and run() orders itself
        # after the upload whatever stream it is called on
        self._pinned = torch.from_numpy(host).pin_memory()
        self.image = self._pinned.to(dev, non_blocking=True)
        if self.image.data_ptr() % 256:
            raise RuntimeError("device image is not 256-byte aligned")
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev)
        self._uploaded = torch.cuda.Event()
        self._uploaded.record()
        self.grid = int(grid)

    def run(self):
        ev = self._uploaded
        if ev is not None:
            if ev.query():
                self._uploaded = None
            elif torch.cuda.is_current_stream_capturing():
                ev.synchronize()                  # a capture may not wait on an outside event
                self._uploaded = None
            else:
                torch.cuda.current_stream().wait_event(ev)
        self.counter.zero_()
        _lib.call("nt_batch_launch", ctypes.c_void_p(self.image.data_ptr()), self.njobs, self.nplans,
                  ctypes.c_int64(self.total_tiles), ctypes.c_void_p(self.counter.data_ptr()),
                  self.grid, stream_handle())


class ImagePool:
    """Reusable job-image slots for batches built again and again (a serving
    loop builds one per scheduling pass).  Each slot is a pinned host buffer,
    a device buffer and a work counter, allocated once.  ``run(jobs)`` packs the
    jobs straight into a free slot's pinned buffer, uploads it on the current
    stream, clears the counter with a kernel and launches; the slot is free
    again once the launch's completion event has passed.  With no free slot,
    ``run`` waits for the oldest one."""

    def __init__(self, slots=8, slot_bytes=1 << 20, device="cuda"):
        self.slot_bytes = int(slot_bytes)
        dev = torch.device(device)
        self.host = [torch.empty(self.slot_bytes, dtype=torch.uint8, pin_memory=True) for _ in range(slots)]
        self.dev = [torch.empty(self.slot_bytes, dtype=torch.uint8, device=dev) for _ in range(slots)]
        self.counter = [torch.zeros(1, dtype=torch.int32, device=dev) for _ in range(slots)]
        for d in self.dev:
            if d.data_ptr() % 256:
                raise RuntimeError("device image slot is not 256-byte aligned")
        self.events = [None] * slots
        self._order = []          # slots in use, oldest first

    def _slot(self):
        for i, ev in enumerate(self.events):
            if ev is None or ev.query():
                self.events[i] = None
                if i in self._order:
                    self._order.remove(i)
                return i
        i = self._order.pop(0)    # all in flight: wait for the oldest
        self.events[i].synchronize()
        self.events[i] = None
        return i

    def run(self, jobs, grid=0):
        """Pack, upload and launch ``jobs`` (as ``MixedBatch``) on the current
        stream; returns (tensors to keep alive until done, completion event)."""
        lib = _lib.load()
        arr, plans, keep, _ = job_table(jobs)
        pv = (ctypes.c_void_p * len(plans))(*[p.handle.value for p in plans])
        nbytes = lib.nt_batch_image_bytes(len(jobs), len(plans))
        if nbytes > self.slot_bytes:
            raise _lib.NtInvalidArgument(_lib.NT_EINVAL, f"batch image of {nbytes} bytes exceeds the "
                                                         f"pool's {self.slot_bytes}-byte slots")
        i = self._slot()
        total = ctypes.c_int64()
        _lib.check(lib.nt_batch_pack(arr, len(jobs), pv, len(plans), ctypes.c_void_p(self.host[i].data_ptr()),
                                     ctypes.c_size_t(nbytes), ctypes.byref(total)))
        self.dev[i][:nbytes].copy_(self.host[i][:nbytes], non_blocking=True)
        self.counter[i].zero_()
        _lib.call("nt_batch_launch", ctypes.c_void_p(self.dev[i].data_ptr()), len(jobs), len(plans),
                  ctypes.c_int64(total.value), ctypes.c_void_p(self.counter[i].data_ptr()), int(grid),
                  stream_handle())
        ev = torch.cuda.Event()
        ev.record()
        self.events[i] = ev
        self._order.append(i)
        return (keep, plans), ev
