"""Continuous-batching server for compact-derivative and ADI requests.

No reference counterpart: the reference runs one field at a time from its
drivers (experiments/run-compact.py, run-adi.py).  A request is

* ``deriv``: the compact derivative along ``axis`` of a field of shape
  (nz, ny, nx), or
* ``adi``: ``steps`` Peaceman-Rachford steps on a side x side grid at ratio r,

each in FP64 or FP32 + refinement.  A request's input either crosses from the
host (``upload``: a slice of a pinned host pool, copied on the request's stream
before its solve, so the copy is part of its latency and its device copy part of
its admission) or is made on the device (a seeded random-phase trigonometric
field, or a Gaussian for ADI: data a previous GPU step left resident).
``Server.run`` replays a request stream with its arrival times on the wall
clock:

* arrived requests wait until their field fits in the HBM budget (admission by
  free HBM: the budget is a fraction of ``torch.cuda.mem_get_info`` minus what
  admitted requests hold);
* on every scheduling pass the newly admitted small derivative requests go to
  the GPU together as one mixed persistent launch, whose job image goes to a
  reused slot of an ``ImagePool`` (no per-pass allocation or pinning); larger
  derivatives and ADI runs get their own fused launches; every launch group
  goes to the stream of a pool with the fewest bytes of work in flight, so new
  work overlaps work already on the device;
* completions are observed with events; memory is released and the latency
  (completion - arrival) recorded.

Plans (the factored level tables) are cached by (n, coefficients) and fields
stay on the device from admission to completion.  ``mode="sequential"`` is
the baseline: one request at a time, plain FP64, each waited for before the
next starts.  The numerics of a request do not depend on what it is batched
with, so results are identical alone and in the mix.
"""
import collections
import dataclasses
import math
import time

import numpy as np
import torch

from .batch import ImagePool
from .core import Plan, adi_prepare, adi_steps, compact_derivative
from . import adi_coeffs, COMPACT_COEFFS

__all__ = ["Request", "make_requests", "Server", "HostPool", "trig_field", "adi_field", "percentile"]


@dataclasses.dataclass
class Request:
    rid: int
    kind: str                      # "deriv" | "adi"
    arrival: float                 # seconds after the start of the stream
    precision: str = "fp64"        # "fp64" | "fp32_refine"
    seed: int = 0
    shape: tuple = (0, 0, 0)       # deriv
    axis: int = 2                  # deriv
    side: int = 0                  # adi
    steps: int = 0                 # adi
    r: float = 0.0                 # adi
    upload: bool = False           # input copied from the host (pinned) instead of made on the device
    admitted: float = math.nan
    done: float = math.nan
    result: object = None

    @property
    def latency(self):
        return self.done - self.arrival

    def nbytes(self):
        """Device bytes the request holds while admitted: input and result (ADI:
        field and work buffer)."""
        if self.kind == "deriv":
            return int(np.prod(self.shape)) * 16
        return self.side * self.side * 16

    def points(self):
        return int(np.prod(self.shape)) if self.kind == "deriv" else self.side * self.side


def make_requests(count, rate, seed=2, deriv_share=0.5, fp32_share=0.5, deriv_sizes=(32, 512),
                  adi_sides=(256, 2048), adi_steps=(1, 20), r_range=(0.1, 2.0), upload_share=0.0):
    """Poisson arrivals at ``rate`` requests/s; per-request kind, sizes, axis,
    precision and seeds drawn from numpy.random.default_rng(seed); a share
    ``upload_share`` of them (drawn from a second generator, so the rest of
    the stream does not change) bring their input from the host."""
    rng = np.random.default_rng(seed)
    up = np.random.default_rng(seed + 7919).random(count) < upload_share
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    reqs = []
    for i in range(count):
        prec = "fp32_refine" if rng.random() < fp32_share else "fp64"
        if rng.random() < deriv_share:
            shape = tuple(int(s) for s in rng.integers(deriv_sizes[0], deriv_sizes[1] + 1, size=3))
            reqs.append(Request(i, "deriv", float(t[i]), prec, int(rng.integers(1 << 30)), shape=shape,
                                axis=int(rng.integers(0, 3))))
        else:
            reqs.append(Request(i, "adi", float(t[i]), prec, int(rng.integers(1 << 30)),
                                side=int(rng.integers(adi_sides[0], adi_sides[1] + 1)),
                                steps=int(rng.integers(adi_steps[0], adi_steps[1] + 1)),
                                r=float(rng.uniform(*r_range))))
        reqs[-1].upload = bool(up[i])
    return reqs


class HostPool:
    """Pinned host memory holding request inputs that cross to the device:
    ``points`` FP64 values drawn once from numpy.random.default_rng(seed)
    (normal, sigma 1); a request's input is the first prod(shape) of them,
    viewed in its shape."""

    def __init__(self, points, seed=11):
        vals = np.random.default_rng(seed).standard_normal(int(points))
        self.data = torch.from_numpy(vals).pin_memory()
        self.points = int(points)

    def view(self, shape):
        k = int(np.prod(shape))
        if k > self.points:
            raise ValueError(f"host pool holds {self.points} points, the request needs {k}")
        return self.data[:k].view(*shape)

    def grow(self, points, seed=11):
        if points > self.points:
            self.__init__(points, seed)

    _shared = None

    @classmethod
    def shared(cls, points):
        """One pool per process, grown to the largest request seen (servers share it)."""
        if cls._shared is None:
            cls._shared = cls(points)
        else:
            cls._shared.grow(points)
        return cls._shared


def trig_field(shape, seed, device):
    """Seeded random-phase field sin(2pi(kx x + px)) sin(2pi(ky y + py)) sin(2pi(kz z + pz))
    on [0,1]^3 with endpoints, wavenumbers 1..4."""
    g = np.random.default_rng(seed)
    k = g.integers(1, 5, size=3)
    ph = g.random(3)
    nz, ny, nx = shape
    lin = lambda n: torch.linspace(0.0, 1.0, n, dtype=torch.float64, device=device)  # noqa: E731
    sz = torch.sin(2 * math.pi * (k[0] * lin(nz) + ph[0]))
    sy = torch.sin(2 * math.pi * (k[1] * lin(ny) + ph[1]))
    sx = torch.sin(2 * math.pi * (k[2] * lin(nx) + ph[2]))
    return sz[:, None, None] * sy[None, :, None] * sx[None, None, :]


def adi_field(side, seed, device):
    """Seeded initial condition for an ADI request: a Gaussian bump at a random
    centre, zero Dirichlet boundary."""
    g = np.random.default_rng(seed)
    cx, cy = g.uniform(0.3, 0.7, size=2)
    x = torch.linspace(0.0, 1.0, side, dtype=torch.float64, device=device)
    u = torch.exp(-((x[None, :] - cx) ** 2 + (x[:, None] - cy) ** 2) / 0.01)
    u[0, :] = 0.0
    u[-1, :] = 0.0
    u[:, 0] = 0.0
    u[:, -1] = 0.0
    return u.contiguous()


def admit(waiting, used, budget, max_batch):
    """Admission by free HBM, first come first served: pop requests from the
    front of ``waiting`` while their fields fit in ``budget`` bytes beside the
    ``used`` bytes of requests in flight, at most ``max_batch`` of them.  The
    head of the queue blocks those behind it (no overtaking).  A request larger
    than the whole budget is admitted alone once nothing else is in flight.
    Returns (admitted requests, new used bytes)."""
    admitted = []
    while waiting and len(admitted) < max_batch:
        need = waiting[0].nbytes()
        if used + need > budget and (used > 0 or admitted):
            break                       # wait for memory to free up
        admitted.append(waiting.popleft())
        used += need
    return admitted, used


def percentile(values, q):
    return float(np.percentile(np.asarray(values, dtype=np.float64), q))


class Server:
    """Replays a request stream on one device.

    ``streams`` HIP streams (default priority) carry the work; each launch
    group goes to the stream with the fewest bytes of work in flight.
    Derivative requests of at most ``small_points`` points admitted in the same
    pass are packed into one persistent batch launch (there the per-launch cost
    dominates; a lone small request is a batch of one), its job image in a
    reused ``ImagePool`` slot; larger ones, and ADI runs, get their own fused
    launches.
    """

    def __init__(self, device="cuda", hbm_fraction=0.5, hbm_budget=None, max_batch=64, streams=4,
                 small_points=1 << 21, keep_results=False):
        self.device = torch.device(device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(free * hbm_fraction)
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.small_points = int(small_points)
        self.keep_results = keep_results
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(max(1, int(streams)))]
        self._busy = [0] * len(self.streams)
        self._plans = {}
        # job-image slots for the batched route: a max_batch image holds at most
        # max_batch jobs and as many distinct plans (about 2 KB each)
        self.images = ImagePool(slots=4 * len(self.streams), slot_bytes=self.max_batch * 2304 + 4096,
                                device=self.device)
        self.host_pool = None

    # -- plan cache (host level tables; shared by all streams) ----------------------
    def plan(self, n, coeffs):
        key = (int(n), tuple(coeffs))
        p = self._plans.get(key)
        if p is None:
            p = self._plans[key] = Plan(n, coeffs)
        return p

    def _batchable(self, q):
        """Small derivative whose lines fit the persistent batch kernel (16..512)."""
        return (q.kind == "deriv" and math.prod(q.shape) <= self.small_points
                and 16 <= q.shape[q.axis] <= 512)

    def _stream(self, nbytes):
        """The stream with the fewest bytes of work still in flight; ``nbytes`` is
        added to it until the launch group completes."""
        i = min(range(len(self.streams)), key=lambda k: self._busy[k])
        self._busy[i] += nbytes
        return i

    # -- launching: each returns (event, stream index, [(request, result)], keep-alive) --
    def _input(self, r, shape):
        """The request's input on the device, on the current stream: copied from
        the pinned host pool (upload) or made in place."""
        if r.upload:
            t = torch.empty(shape, dtype=torch.float64, device=self.device)
            t.copy_(self.host_pool.view(shape), non_blocking=True)
            return t
        if r.kind == "deriv":
            return trig_field(shape, r.seed, self.device).contiguous()
        return adi_field(r.side, r.seed, self.device)

    def _launch_derivs(self, reqs, batched, precision=None):
        si = self._stream(sum(r.nbytes() for r in reqs))
        with torch.cuda.stream(self.streams[si]):
            jobs, items = [], []
            for r in reqs:
                f = self._input(r, r.shape)
                out = torch.empty(r.shape, dtype=torch.float64, device=self.device)
                jobs.append(dict(f=f, out=out, axis=r.axis, dx=1.0 / (r.shape[r.axis] - 1),
                                 precision=precision or r.precision, sweeps=1))
                items.append((r, out))
            if batched:
                keep, ev = self.images.run(jobs)
                return ev, si, items, keep
            for j in jobs:
                compact_derivative(self.plan(j["f"].shape[j["axis"]], COMPACT_COEFFS), j["f"], j["out"],
                                   j["dx"], j["axis"], j["precision"], j["sweeps"])
            ev = torch.cuda.Event()
            ev.record()
        return ev, si, items, jobs

    def _launch_adi(self, r, precision=None):
        si = self._stream(r.nbytes() * r.steps)
        with torch.cuda.stream(self.streams[si]):
            u = self._input(r, (r.side, r.side))
            work = torch.empty_like(u)
            adi_prepare(u, work)
            adi_steps(self.plan(r.side, adi_coeffs(r.r)), u, work, r.r, r.steps, precision or r.precision, 1)
            ev = torch.cuda.Event()
            ev.record()
        return ev, si, [(r, u)], work

    def warmup(self):
        """Send one small request down every route (batched and fused derivative on
        each axis, ADI; both precisions) through both modes, so that one-time costs
        (code-object loading on first launch, allocator growth, pinned-memory pool)
        are paid before a timed run."""
        side = min(512, int(round(self.small_points ** (1.0 / 3.0))) + 2)
        reqs = []
        for prec in ("fp64", "fp32_refine"):
            for axis in range(3):
                reqs.append(Request(len(reqs), "deriv", 0.0, prec, 1, shape=(32, 32, 32), axis=axis))
                reqs.append(Request(len(reqs), "deriv", 0.0, prec, 2, shape=(side, side, side), axis=axis))
            reqs.append(Request(len(reqs), "adi", 0.0, prec, 3, side=256, steps=2, r=0.5))
        up = [Request(len(reqs) + i, q.kind, 0.0, q.precision, q.seed, shape=q.shape, axis=q.axis,
                      side=q.side, steps=q.steps, r=q.r, upload=True) for i, q in enumerate(reqs)]
        for mode in ("continuous", "sequential"):
            self.run(reqs + up, mode=mode)

    # -- the scheduling loop ------------------------------------------------------------
    def run(self, requests, mode="continuous", time_scale=1.0):
        """Serve ``requests`` (arrival times in seconds, scaled by time_scale);
        returns a stats dict.  mode: "continuous" | "sequential"."""
        if mode not in ("continuous", "sequential"):
            raise ValueError("mode must be 'continuous' or 'sequential'")
        reqs = sorted(requests, key=lambda q: q.arrival)
        for q in reqs:
            q.done = q.admitted = math.nan
            q.result = None
        need = max([q.points() for q in reqs if q.upload] or [0])
        if need:                  # pinned host inputs exist before the clock starts
            self.host_pool = HostPool.shared(need)
        pending = collections.deque(reqs)
        waiting = collections.deque()
        inflight = []              # (event, stream index, [(request, result tensor)], keep-alive)
        used = 0
        batches = 0
        self._busy = [0] * len(self.streams)
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731
        while pending or waiting or inflight:
            now = clock()
            while pending and pending[0].arrival * time_scale <= now:
                waiting.append(pending.popleft())
            progressed = False
            still = []
            for ev, si, items, keep in inflight:
                if ev.query():
                    t = clock()
                    for q, res in items:
                        q.done = t
                        used -= q.nbytes()
                        self._busy[si] -= q.nbytes() * (q.steps if q.kind == "adi" else 1)
                        if self.keep_results:
                            q.result = res
                    progressed = True
                else:
                    still.append((ev, si, items, keep))
            inflight = still
            if mode == "continuous":
                admitted, used = admit(waiting, used, self.budget, self.max_batch)
                t_adm = clock()
                for q in admitted:
                    q.admitted = t_adm
                small = [q for q in admitted if self._batchable(q)]
                for q in admitted:
                    if q.kind == "adi":
                        inflight.append(self._launch_adi(q))
                    elif not self._batchable(q):
                        inflight.append(self._launch_derivs([q], batched=False))
                if small:   # the route depends on the request only, so alone == in the mix
                    inflight.append(self._launch_derivs(small, batched=True))
                    batches += 1
                progressed = progressed or bool(admitted)
            else:
                if not inflight and waiting:
                    q = waiting.popleft()
                    used += q.nbytes()
                    q.admitted = clock()
                    if q.kind == "deriv":
                        inflight.append(self._launch_derivs([q], batched=False, precision="fp64"))
                    else:
                        inflight.append(self._launch_adi(q, precision="fp64"))
                    progressed = True
            if not progressed:
                time.sleep(20e-6)
        torch.cuda.synchronize(self.device)
        lat = [q.latency for q in reqs]
        span = max(q.done for q in reqs) - min(q.arrival * time_scale for q in reqs)
        return {"mode": mode, "requests": len(reqs), "uploads": sum(1 for q in reqs if q.upload),
                "p50_ms": 1e3 * percentile(lat, 50),
                "p99_ms": 1e3 * percentile(lat, 99), "mean_ms": 1e3 * float(np.mean(lat)),
                "throughput_rps": len(reqs) / span if span > 0 else math.inf, "batches": batches,
                "streams": len(self.streams), "budget_bytes": self.budget}
