"""Benchmark of the near-Toeplitz hot path on one MI355X.

Headline workload: 4th-order compact d/dx of a 512^3 FP64 field
f = sin(2 pi x) sin(2 pi y) sin(2 pi z) + N(0, 1e-3) noise (numpy
default_rng(0), generated on the host and copied to the device), i.e. 262,144
near-Toeplitz systems of size 512, one fused launch per step (RHS stencil +
cyclic reduction, each point read once and written once).  Metric: grid points
per second and the fraction of HBM peak.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--quick] [--case NAME ...]

Prints ONE compact JSON line on stdout (a few KB).  Every figure it carries
is also in a longer record written to stderr.  Other configurations are
summarised under "extras": d/dy and d/dz ("axes") and FP32 + refinement
sweeps ("refine") on the headline's field, timed before the headline; ADI
2048^2 x 200 steps in a hipGraph and the partitioned ADI form ("adi"), a
1024^3 field stored in FP32 ("fp32_storage"), 64 mixed-size grids in one
persistent launch ("batch"), and a Poisson request stream served with
continuous batching ("serving"), timed after it.  --quick skips them; --case
runs only the named ones.  "strided_sizes" (d/dz and d/dy of 2^26-point fields
with 96..480-point lines) runs only when named.

Every result's error is measured against the CPU oracle (SciPy banded solve
of the same matrix, oracle/) on lines copied from the same device field, with
the sample size stated.  The CPU baseline times that oracle path on lines of
the benchmarked field, on one core and on every host core.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "near-toeplitz-applications_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "compact_ddx_512_fp64_traffic.json")
METRIC = "Grid pts/s & % HBM peak, 512^3 compact d/dx; p50/p99 served vs plain FP64"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sig(x, digits=4):
    """Round floats to `digits` significant digits, recursively (compact JSON)."""
    if isinstance(x, float):
        if x == 0.0 or not math.isfinite(x):
            return x
        return float(f"{x:.{digits}g}")
    if isinstance(x, dict):
        return {k: sig(v, digits) for k, v in x.items()}
    if isinstance(x, (list, tuple)):
        return [sig(v, digits) for v in x]
    return x


def summary(values):
    """{max, p50, count} of a list of per-grid errors."""
    if not values:
        return None
    return {"max": max(values), "p50": statistics.median(values), "count": len(values)}


# ---------------------------------------------------------------------------
# oracle checks (the CPU oracle is the checker here, never the thing timed on GPU)

def gather_lines(t, axis, idx_a, idx_b):
    """Lines along `axis` of a 3-D torch tensor at the index pairs (idx_a, idx_b)
    of the other two axes, in order, as a host float64 array (lines, n)."""
    import torch
    a = torch.as_tensor(idx_a, device=t.device)
    b = torch.as_tensor(idx_b, device=t.device)
    if axis == 2:
        g = t[a, b, :]
    elif axis == 1:
        g = t[a, :, b]          # separated index arrays: their dimension comes first
    else:
        g = t[:, a, b].t()
    return g.double().cpu().numpy()


def oracle_err(f, out, axis, h, nlines=256, seed=0):
    """Max relative error of `out` against the oracle's derivative of `f` on
    `nlines` lines drawn at random (numpy default_rng(seed)) along `axis`."""
    import numpy as np
    from oracle.compact import COMPACT_COEFFS, compact_rhs
    from oracle.cr import banded_solve
    others = [d for d in range(3) if d != axis]
    rng = np.random.default_rng(seed)
    ia = rng.integers(0, f.shape[others[0]], nlines)
    ib = rng.integers(0, f.shape[others[1]], nlines)
    src = gather_lines(f, axis, ia, ib)
    got = gather_lines(out, axis, ia, ib)
    want = banded_solve(COMPACT_COEFFS, compact_rhs(src, h))
    return float(np.abs(got - want).max() / np.abs(want).max())


# ---------------------------------------------------------------------------
# CPU baseline: NumPy RHS + one scipy.linalg.solve_banded call (LAPACK gtsv),
# the reference's CPU path (parallelcompact/misc.py:5-19), on lines of the
# benchmarked field.

def _pool_init():
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)
    globals()["_POOL_LIMIT"] = True


def _pool_task(args):
    """Worker: the compact derivative of lines lo..hi of a shared (lines, n)
    array into the shared output; returns the seconds it took."""
    import numpy as np
    from multiprocessing import shared_memory
    from oracle.compact import COMPACT_COEFFS, compact_rhs
    from oracle.cr import banded_solve
    name_in, name_out, shape, lo, hi, h = args
    si = shared_memory.SharedMemory(name=name_in)
    so = shared_memory.SharedMemory(name=name_out)
    try:
        src = np.ndarray(shape, dtype=np.float64, buffer=si.buf)
        dst = np.ndarray(shape, dtype=np.float64, buffer=so.buf)
        t0 = time.perf_counter()
        dst[lo:hi] = banded_solve(COMPACT_COEFFS, compact_rhs(src[lo:hi], h))
        return time.perf_counter() - t0
    finally:
        del src, dst
        si.close()
        so.close()


def host_lines(f, axis, nlines, seed):
    """`nlines` lines along `axis` of the host field f, as a (nlines, n) array."""
    import numpy as np
    others = [d for d in range(3) if d != axis]
    rng = np.random.default_rng(seed)
    ia = rng.integers(0, f.shape[others[0]], nlines)
    ib = rng.integers(0, f.shape[others[1]], nlines)
    if axis == 2:
        return np.ascontiguousarray(f[ia, ib, :])
    if axis == 1:
        return np.ascontiguousarray(f[ia, :, ib])
    return np.ascontiguousarray(f[:, ia, ib].T)


def cpu_baseline(f, h, sample_lines=65536, reps=3, pool_workers=None):
    """The oracle path timed on the host: per direction, `sample_lines` lines of
    the benchmarked field f on one core (BLAS limited to one thread), and d/dx
    of every x-line of f split over a process pool of every available host core
    (at most 16).  Runs before the GPU is touched: the pool is started as
    child processes of a process that has not initialised HIP."""
    import numpy as np
    from threadpoolctl import threadpool_info, threadpool_limits
    from oracle.compact import COMPACT_COEFFS, compact_rhs
    from oracle.cr import banded_solve
    n = f.shape[2]
    single = {}
    with threadpool_limits(1):
        blas_threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
        for axis, name in ((2, "ddx"), (1, "ddy"), (0, "ddz")):
            lines = host_lines(f, axis, sample_lines, seed=70 + axis)
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                banded_solve(COMPACT_COEFFS, compact_rhs(lines, h))
                times.append(time.perf_counter() - t0)
            med = statistics.median(times)
            single[name] = sample_lines * n / med
            log(f"[bench] CPU 1 core {name}: {single[name]:.3e} pts/s "
                f"({sample_lines} lines, median of {reps})")
    out = {"value": single["ddx"], "unit": "grid pts/s", "cores": 1, "kind": "port",
           "sample": f"{sample_lines} lines per direction copied from the benchmarked {n}^3 field "
                     f"(seeded draw), NumPy RHS + one scipy.linalg.solve_banded call (LAPACK gtsv), "
                     f"BLAS limited to 1 thread, median of {reps}",
           "blas_threads": blas_threads, "host_cores_available": len(os.sched_getaffinity(0)),
           "single_core_pts_per_s": single, "all_cores": None}
    workers = pool_workers or min(16, len(os.sched_getaffinity(0)))
    if workers > 1:
        try:
            out["all_cores"] = cpu_pool(f, h, workers, reps)
        except Exception as e:          # the all-core leg is optional
            log(f"[bench] CPU pool leg skipped: {e!r}")
    if out["all_cores"] is not None:
        ac = out["all_cores"]
        out.update(value=ac["pts_per_s"], cores=workers,
                   sample=f"d/dx of all {f.shape[0] * f.shape[1]} x-lines of the benchmarked {n}^3 field over "
                          f"{workers} processes x 1 thread (NumPy RHS + solve_banded, LAPACK gtsv), median "
                          f"of {reps}; single_core: {sample_lines} lines per direction, 1 thread")
    return out


def cpu_pool(f, h, workers, reps):
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from multiprocessing import shared_memory
    import numpy as np
    nl = f.shape[0] * f.shape[1]
    n = f.shape[2]
    shape = (nl, n)
    si = shared_memory.SharedMemory(create=True, size=nl * n * 8)
    so = shared_memory.SharedMemory(create=True, size=nl * n * 8)
    try:
        np.ndarray(shape, dtype=np.float64, buffer=si.buf)[:] = f.reshape(shape)
        bounds = np.linspace(0, nl, workers + 1).astype(int)
        tasks = [(si.name, so.name, shape, int(bounds[i]), int(bounds[i + 1]), h) for i in range(workers)]
        with ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn"),
                                 initializer=_pool_init) as ex:
            # one small task per worker first: interpreter start-up and imports are not timed
            warm = [(si.name, so.name, shape, i * 4, i * 4 + 4, h) for i in range(workers)]
            list(ex.map(_pool_task, warm))
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                list(ex.map(_pool_task, tasks, chunksize=1))
                times.append(time.perf_counter() - t0)
        med = statistics.median(times)
        res = {"pts_per_s": nl * n / med, "seconds": med}
        log(f"[bench] CPU {workers} cores d/dx: {res['pts_per_s']:.3e} pts/s (all {nl} x-lines, "
            f"median of {reps})")
        return res
    finally:
        si.close()
        si.unlink()
        so.close()
        so.unlink()


# ---------------------------------------------------------------------------
# GPU cases

def time_launches(fn, steps, warmup):
    """Warm up, then time `steps` launches: per-launch HIP events on the current
    stream (the stream the kernels go to) plus a wall clock around all of them."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    t0 = time.perf_counter()
    for a, b in evs:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [a.elapsed_time(b) * 1e-3 for a, b in evs]
    return {"wall_s": wall, "mean_s": sum(per) / len(per), "median_s": statistics.median(per),
            "min_s": min(per), "max_s": max(per)}


def compact_case(f, axis, precision, sweeps, steps, warmup, out=None):
    import torch
    from ntamd import COMPACT_COEFFS, Plan, core
    n = f.shape[axis]
    if out is None:
        out = torch.empty(f.shape, dtype=torch.float64, device=f.device)
    plan = Plan(n, COMPACT_COEFFS)
    h = 1.0 / (n - 1)
    t = time_launches(lambda: core.compact_derivative(plan, f, out, h, axis, precision, sweeps),
                      steps, warmup)
    pts = f.numel()
    nbytes = pts * (f.element_size() + 8)
    t["pts_per_s"] = pts / t["mean_s"]
    t["alg_bytes"] = nbytes
    t["gbps"] = nbytes / t["mean_s"] / 1e9
    t["hbm_frac"] = nbytes / t["mean_s"] / HBM_PEAK
    return t, out


def adi_case(N, r, nsteps, reps):
    """ADI 2048^2, 200 Peaceman-Rachford steps in one hipGraph (monolithic
    solves), and the one-GPU partitioned form: 16 subdomains of 128 per line
    plus the reduced 32-unknown system for both half steps."""
    import numpy as np
    import torch
    from adi.solver import ADISolver
    from ntamd.workloads import gaussian_ic
    from oracle import adi as oadi
    u0 = gaussian_ic(N, seed=0)
    u = torch.tensor(u0, device="cuda")
    s = ADISolver(N, r)
    s.capture(u, nsteps)
    s.replay()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        s.replay()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b) * 1e-3)
    t = statistics.median(ts)
    half_bytes = 16 * N * N
    # one step from the initial condition against the oracle's PR step
    u1 = torch.tensor(u0, device="cuda")
    s1 = ADISolver(N, r)
    s1.step(u1, 1)
    want = oadi.pr_run(u0, r, 1)
    err1 = float(np.abs(u1.cpu().numpy() - want).max() / np.abs(want).max())
    # the 32 MiB field stays in the 256 MB Infinity Cache: effective, not HBM, GB/s
    res = {"N": N, "r": r, "steps": nsteps, "ms_per_step": 1e3 * t / nsteps,
           "effective_gbps": 2 * nsteps * half_bytes / t / 1e9, "err_vs_oracle_1step": err1}
    from ntamd import adi_coeffs
    from ntamd.partition import PartitionPlan
    pp = PartitionPlan(N, 16, adi_coeffs(r))
    w = torch.empty_like(u)
    core_prepare(u, w)
    tx = time_launches(lambda: pp.adi_half_x(u, w, r), 20, 3)
    tx1 = time_launches(lambda: pp.adi_half_x(u, w, r, tiles=False), 20, 3)
    u2 = torch.empty_like(u)
    ty = time_launches(lambda: pp.adi_half_y(w, u2, r), 20, 3)
    res["partitioned_x_half_step_16x128_us"] = tx["mean_s"] * 1e6
    res["partitioned_x_half_step_one_row_per_workgroup_us"] = tx1["mean_s"] * 1e6
    res["partitioned_y_half_step_16x128_us"] = ty["mean_s"] * 1e6
    res.update(partitioned_adi_case(u0, r, nsteps, reps))
    return res


def core_prepare(u, w):
    from ntamd import core
    core.adi_prepare(u, w)


def partitioned_adi_case(u0, r, nsteps, reps, P=16):
    """Both half steps partitioned (P = 16 subdomains of 128 + the reduced
    32-unknown system per line), all `nsteps` steps in one hipGraph; error
    against the oracle after one step."""
    import numpy as np
    import torch
    from ntamd.partition import PartitionedADISolver
    from oracle import adi as oadi
    N = u0.shape[0]
    s = PartitionedADISolver(N, r, P)
    u = torch.tensor(u0, device="cuda")
    s.capture(u, nsteps)
    s.replay()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        s.replay()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b) * 1e-3)
    t = statistics.median(ts)
    u1 = torch.tensor(u0, device="cuda")
    PartitionedADISolver(N, r, P).step(u1, 1)
    want = oadi.pr_run(u0, r, 1)
    err = float(np.abs(u1.cpu().numpy() - want).max() / np.abs(want).max())
    return {"partitioned_ms_per_step": 1e3 * t / nsteps, "partitioned_err_vs_oracle_1step": err}


def measured_traffic():
    """HBM bytes per launch of the headline kernel from the committed rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE passes (scaled by a 1 GiB copy's readings)."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            return json.load(fh)["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def pmc_probe():
    """Three launches each of the headline d/dx, then d/dy and d/dz, then three
    copies of the same 1 GiB field: the program profiled by the FETCH_SIZE /
    WRITE_SIZE passes."""
    import torch
    from ntamd import COMPACT_COEFFS, Plan, core
    N = 512
    f = torch.randn((N, N, N), dtype=torch.float64, device="cuda")
    out = torch.empty_like(f)
    plan = Plan(N, COMPACT_COEFFS)
    for axis in (2, 1, 0):
        for _ in range(3):
            core.compact_derivative(plan, f, out, 1.0 / (N - 1), axis)
    for _ in range(3):
        out.copy_(f)
    torch.cuda.synchronize()
    print("pmc probe done", flush=True)


def strided_sizes_case(steps, warmup, sizes=(96, 160, 300, 384, 480)):
    """d/dz and d/dy of about 2^26-point FP64 fields whose lines are not a power
    of two long, with the default launch shapes (one lane per line): per length
    [us per launch, GB/s at 16 B/point, max rel err vs the oracle on 256 lines]."""
    import torch
    from ntamd.workloads import compact_field
    res = {}
    for n in sizes:
        m = int(round(((1 << 26) / n) ** 0.5))
        m -= m % 64
        for axis, name in ((0, "ddz"), (1, "ddy")):
            shape = (n, m, m) if axis == 0 else (m, n, m)
            f = torch.from_numpy(compact_field(shape, seed=1)).to("cuda")
            out = torch.empty_like(f)
            t, _ = compact_case(f, axis, "fp64", 1, steps, warmup, out=out)
            err = oracle_err(f, out, axis, 1.0 / (n - 1), 256, seed=n + axis)
            res[f"{name}_{n}"] = [t["mean_s"] * 1e6, t["gbps"], err]
            log(f"[bench] strided {name} n={n} {shape}: {t['mean_s'] * 1e6:.1f} us, "
                f"{t['gbps']:.0f} GB/s, err {err:.2e}")
            del f, out
            torch.cuda.empty_cache()
    return res


def config4_case(N=1024, reps=5):
    """1024^3 field stored in FP32 (4 GiB): d/dx, d/dy, d/dz with the FP32 solve +
    one FP64-residual sweep, against plain FP64 (the FP64-stored field, FP64
    solve), and FP32 storage with the FP64 solve.  Errors of every form are
    against the oracle on 256 lines per axis copied from the device field."""
    import torch
    from ntamd.workloads import compact_field_device
    # Blocks cached from the earlier cases go back to the driver first, so the
    # 1024^3 fields are fresh allocations.
    torch.cuda.empty_cache()
    f32 = compact_field_device(N, seed=0, dtype=torch.float32)
    f64 = f32.double()
    out = torch.empty((N, N, N), dtype=torch.float64, device=f32.device)
    h = 1.0 / (N - 1)
    res = {}
    for axis, name in ((2, "ddx"), (1, "ddy"), (0, "ddz")):
        tr, _ = compact_case(f32, axis, "fp32_refine", 1, reps, 2, out=out)
        er = oracle_err(f32, out, axis, h, 256, seed=axis)
        t64, _ = compact_case(f64, axis, "fp64", 1, reps, 2, out=out)
        e64 = oracle_err(f64, out, axis, h, 256, seed=axis)
        ts, _ = compact_case(f32, axis, "fp64", 1, reps, 2, out=out)
        es = oracle_err(f32, out, axis, h, 256, seed=axis)
        # ms and error vs the oracle of: FP32 storage + FP32 solve + 1 sweep ("fp32_refine"),
        # FP64 storage + FP64 solve ("fp64"), FP32 storage + FP64 solve ("f32_store_fp64")
        res[name] = {"ms": {"fp32_refine": tr["mean_s"] * 1e3, "fp64": t64["mean_s"] * 1e3,
                            "f32_store_fp64": ts["mean_s"] * 1e3},
                     "err": {"fp32_refine": er, "fp64": e64, "f32_store_fp64": es},
                     "speedup_vs_fp64": t64["mean_s"] / tr["mean_s"]}
        log(f"[bench] 1024^3 {name}: fp32+ref {tr['mean_s'] * 1e3:.3f} ms, fp64 {t64['mean_s'] * 1e3:.3f} ms "
            f"(x{res[name]['speedup_vs_fp64']:.2f}); fp32 storage + fp64 solve {ts['mean_s'] * 1e3:.3f} ms; "
            f"err vs oracle {er:.2e} / {e64:.2e} / {es:.2e}")
    del f32, f64, out
    torch.cuda.empty_cache()
    return res


def mixed_batch_case(count=64, seed=1, reps=5, sizes=None, precision="fp32_refine", lines_per_grid=32):
    """`count` grids, per-axis sizes uniform in `sizes` (default {64, 96, ...,
    512}, numpy default_rng(seed)), a random derivative direction each, FP32
    solve + one sweep, all in one persistent launch, against the same jobs as
    separate fused launches.  Per-grid max relative error against the oracle
    on `lines_per_grid` lines of each grid."""
    import numpy as np
    import torch
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.batch import MixedBatch
    from ntamd.workloads import compact_field_device
    rng = np.random.default_rng(seed)
    sizes = np.arange(64, 513, 32) if sizes is None else np.asarray(sizes)
    jobs, pts = [], 0
    for i in range(count):
        shape = tuple(int(s) for s in rng.choice(sizes, size=3))
        axis = int(rng.integers(0, 3))
        f = compact_field_device(shape, seed=100 + i)
        jobs.append(dict(f=f, out=torch.empty_like(f), axis=axis, dx=1.0 / (shape[axis] - 1),
                         precision=precision, sweeps=1))
        pts += int(np.prod(shape))
    b = MixedBatch(jobs)
    t = time_launches(b.run, reps, 2)
    errs = [oracle_err(j["f"], j["out"], j["axis"], j["dx"], lines_per_grid, seed=k)
            for k, j in enumerate(jobs)]
    plans = {n: Plan(n, COMPACT_COEFFS) for n in {j["f"].shape[j["axis"]] for j in jobs}}

    def separate():
        for j in jobs:
            core.compact_derivative(plans[j["f"].shape[j["axis"]]], j["f"], j["out"], j["dx"], j["axis"],
                                    j["precision"], j["sweeps"])
    ts = time_launches(separate, reps, 2)
    # medians: a launch sequence that pays a one-off cost (a code object loaded,
    # an occupancy query answered the first time) does not set the comparison
    res = {"grids": count, "points": pts, "launch_ms": t["median_s"] * 1e3,
           "gbps": pts * 16 / t["median_s"] / 1e9, "separate_launches_ms": ts["median_s"] * 1e3,
           "speedup_vs_separate": ts["median_s"] / t["median_s"],
           "err_vs_oracle_per_grid": summary(errs), "err_lines_per_grid": lines_per_grid}
    log(f"[bench] mixed batch: {count} grids, {pts / 1e6:.0f} M pts, {t['median_s'] * 1e3:.3f} ms "
        f"(min {t['min_s'] * 1e3:.3f}, max {t['max_s'] * 1e3:.3f}; {res['gbps']:.0f} GB/s); {count} separate "
        f"launches {ts['median_s'] * 1e3:.3f} ms (min {ts['min_s'] * 1e3:.3f}, max {ts['max_s'] * 1e3:.3f}); "
        f"worst grid err vs oracle {max(errs):.2e}")
    if count <= 64:
        log(f"[bench] mixed batch per-grid errors vs oracle: {[float(f'{e:.3g}') for e in errs]}")
    del jobs, b
    torch.cuda.empty_cache()
    return res


SERVE_RATES = (500.0, 2000.0, 3000.0)
SERVE_UPLOAD_RATES = (100.0, 400.0)


def serving_case(count=1000, rates=SERVE_RATES, upload_rates=SERVE_UPLOAD_RATES, seed=2):
    """Poisson arrivals (rate req/s, numpy default_rng(2)): compact derivatives
    (per-axis sizes U[32,512], random axis) and ADI runs (side U[256,2048],
    1-20 steps, r U[0.1,2]), FP64 or FP32+refinement each; continuous batching
    vs one-at-a-time plain FP64, same request stream, each on a warmed-up
    server.  Two input sources: device-resident (fields made on the device, at
    `rates`) and uploads (every input copied from pinned host memory on the
    request's stream, at `upload_rates`: about 150 MB per request crosses the
    host link)."""
    from ntamd.serve import Server, make_requests
    out = {"requests": count, "seed": seed}
    for share, rs, tag in ((0.0, rates, "rate"), (1.0, upload_rates, "upload_rate")):
        for rate in rs:
            case = {}
            for mode in ("continuous", "sequential"):
                reqs = make_requests(count, rate, seed=seed, upload_share=share)
                srv = Server()
                srv.warmup()             # first-launch costs are paid outside the timed stream
                stats = srv.run(reqs, mode=mode)
                case[mode] = stats
                log(f"[bench] serving {tag} {rate:.0f} req/s {mode}: p50 {stats['p50_ms']:.2f} ms, "
                    f"p99 {stats['p99_ms']:.2f} ms, {stats['throughput_rps']:.0f} req/s")
                del srv
                torch_empty_cache()
            out[f"{tag}_{rate:.0f}"] = case
    return out


def torch_empty_cache():
    import torch
    torch.cuda.empty_cache()


def served_summary(serving):
    """p50/p99 of the served request stream at each stated arrival rate, continuous
    batching beside one-at-a-time plain FP64 (None when the case did not run)."""
    if not serving:
        return None
    # rate_*: inputs resident on the device; upload_rate_*: every input copied from pinned host memory
    out = {"requests": serving["requests"], "arrivals": "Poisson, numpy default_rng(2)"}
    for key, case in serving.items():
        if key.startswith("rate_") or key.startswith("upload_rate_"):
            out[key] = {m: {k: case[m][k] for k in ("p50_ms", "p99_ms", "throughput_rps")}
                        for m in ("continuous", "sequential")}
    return out


def compact_extras(extras):
    """The extra cases as short arrays for the stdout line (the full record,
    with every field named, goes to stderr).  The legend travels in the line."""
    out = {}
    legend = {}
    us_err = {k: [v["us"], v["err_vs_oracle"]] for k, v in extras.items()
              if isinstance(v, dict) and "us" in v}
    if us_err:
        out["us_err"] = us_err
        legend["us_err"] = "512^3 field of the headline: [us per launch, max rel err vs oracle on 256 lines]"
    a = extras.get("adi_2048_200steps_graph")
    if a:
        out["adi_2048"] = [a["ms_per_step"], a["err_vs_oracle_1step"], a["partitioned_ms_per_step"],
                           a["partitioned_x_half_step_16x128_us"], a["partitioned_y_half_step_16x128_us"],
                           a["partitioned_err_vs_oracle_1step"]]
        legend["adi_2048"] = ("200 steps in one hipGraph: [ms/step, err after 1 step; partitioned 16x128 + "
                              "reduced 32: ms/step, x half step us, y half step us, err after 1 step]")
    c = extras.get("fp32_storage_1024")
    if c:
        out["fp32_storage_1024"] = {ax: [r["ms"]["fp32_refine"], r["ms"]["fp64"], r["ms"]["f32_store_fp64"],
                                         r["speedup_vs_fp64"], max(r["err"].values())]
                                    for ax, r in c.items()}
        legend["fp32_storage_1024"] = ("[ms FP32 storage + FP32 solve + 1 sweep, ms plain FP64, ms FP32 storage "
                                       "+ FP64 solve, speedup of the first over plain FP64, worst err vs oracle]")
    for key in ("mixed_batch_64", "mixed_batch_small_1024"):
        b = extras.get(key)
        if b:
            out[key] = [b["grids"], b["launch_ms"], b["separate_launches_ms"], b["err_vs_oracle_per_grid"]["max"]]
            legend["mixed_batch"] = "[grids, ms one persistent launch, ms separate launches, worst grid err vs oracle]"
    if out:
        out["legend"] = legend
    return out


def compact_served(served):
    """p50/p99 ms and requests/s per arrival rate: continuous beside sequential."""
    if not served:
        return None
    out = {"requests": served["requests"],
           "fields": "[p50_ms, p99_ms, req/s] continuous | sequential plain FP64; upload_* = inputs "
                     "copied from pinned host memory"}
    for key, case in served.items():
        if key.startswith("rate_") or key.startswith("upload_rate_"):
            out[key] = {m: [case[m]["p50_ms"], case[m]["p99_ms"], case[m]["throughput_rps"]]
                        for m in ("continuous", "sequential")}
    return out


EXTRA_CASES = ("axes", "refine", "adi", "fp32_storage", "batch", "serving")
# run only when named with --case
OPTIONAL_CASES = ("strided_sizes",)
EXTRA_WARMUP = 30          # untimed launches of each extra 512^3 case


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="headline only, no extras/CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run the short program profiled for HBM traffic and exit")
    ap.add_argument("--case", action="append", choices=EXTRA_CASES + OPTIONAL_CASES,
                    help="run only these extra cases (repeatable)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a derivative's lines "
              "shard across GPUs without communication; run one process per GPU on its share)",
              file=sys.stderr)
        sys.exit(2)

    if args.pmc_probe:
        import torch
        from ntamd import _lib
        _lib.load()
        pmc_probe()
        return

    import numpy as np
    from ntamd.workloads import compact_field
    N = 512
    h = 1.0 / (N - 1)
    t0 = time.perf_counter()
    f_host = compact_field(N, seed=0)
    log(f"[bench] 512^3 field generated on the host in {time.perf_counter() - t0:.1f} s")
    cpu = None
    if not (args.no_cpu or args.quick):
        # before the GPU is initialised: the all-core leg starts child processes
        cpu = cpu_baseline(f_host, h)

    import torch
    from ntamd import _lib, autotune
    _lib.load()
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    tuned = autotune.apply()          # launch shapes chosen by the autotuner, if cached
    f64 = torch.from_numpy(f_host).to("cuda")
    del f_host

    extras = {}
    cases = set(args.case) if args.case else (set() if args.quick else set(EXTRA_CASES))
    short = max(10, args.steps // 2)
    out = torch.empty_like(f64)
    o2 = torch.empty_like(out)

    def line(t, err):
        return {"us": t["mean_s"] * 1e6, "gbps": t["gbps"], "hbm_frac": t["hbm_frac"], "err_vs_oracle": err}
    # The extra 512^3 cases take 30 untimed launches each: after a switch to
    # another kernel, launches 6-10 of a run come out up to 1.5x slower (FP32 +
    # refinement d/dx: 338, 345, 362, 433, 500 us, then steady at 342-350 us over
    # 300 launches; FP64 steady at 328-344 us throughout), so a few warm-up
    # launches would time that transient instead of the kernel.
    xw = max(args.warmup, EXTRA_WARMUP)
    if "axes" in cases:
        for name, axis in (("ddy_fp64", 1), ("ddz_fp64", 0)):
            t, _ = compact_case(f64, axis, "fp64", 1, short, xw, out=o2)
            extras[name] = line(t, oracle_err(f64, o2, axis, h, 256, seed=axis))
            log(f"[bench] {name}: {t['mean_s'] * 1e6:.1f} us, {t['gbps']:.0f} GB/s, "
                f"err {extras[name]['err_vs_oracle']:.2e}")
    if "refine" in cases:
        for name, axis, sweeps in (("ddx_fp32_refine1", 2, 1), ("ddx_fp32_refine2", 2, 2),
                                   ("ddy_fp32_refine1", 1, 1), ("ddz_fp32_refine1", 0, 1)):
            t, _ = compact_case(f64, axis, "fp32_refine", sweeps, short, xw, out=o2)
            extras[name] = line(t, oracle_err(f64, o2, axis, h, 256, seed=10 + axis))
            log(f"[bench] {name}: {t['mean_s'] * 1e6:.1f} us, err {extras[name]['err_vs_oracle']:.2e}")
    # The headline runs after the other cases on this field: the device has
    # been streaming for tens of ms, past the first launches after start-up,
    # which run a few percent slower.  It still takes exactly `warmup` untimed
    # and `steps` timed launches.
    log(f"[bench] headline: compact d/dx {N}^3 FP64, steps={args.steps} warmup={args.warmup}")
    head, out = compact_case(f64, 2, "fp64", 1, args.steps, args.warmup, out=out)
    head_err = oracle_err(f64, out, 2, h, 1024, seed=0)
    log(f"[bench] d/dx: {head['mean_s'] * 1e6:.1f} us/launch, {head['gbps']:.0f} GB/s, "
        f"err vs oracle {head_err:.2e} (1024 lines)")
    del out, o2, f64
    torch.cuda.empty_cache()
    if "adi" in cases:
        a = adi_case(2048, 0.5, 200, 3)
        extras["adi_2048_200steps_graph"] = a
        log(f"[bench] ADI 2048^2: {a['ms_per_step']:.3f} ms/step (err 1 step {a['err_vs_oracle_1step']:.1e}); "
            f"partitioned (16 x 128 + reduced 32): x half step {a['partitioned_x_half_step_16x128_us']:.1f} us, "
            f"y half step {a['partitioned_y_half_step_16x128_us']:.1f} us, both in a graph "
            f"{a['partitioned_ms_per_step']:.3f} ms/step (err 1 step {a['partitioned_err_vs_oracle_1step']:.1e})")
    if "fp32_storage" in cases:
        extras["fp32_storage_1024"] = config4_case()
    if "batch" in cases:
        extras["mixed_batch_64"] = mixed_batch_case()
        extras["mixed_batch_small_1024"] = mixed_batch_case(count=1024, sizes=np.arange(16, 65, 8), reps=5,
                                                            lines_per_grid=8)
    if "serving" in cases:
        extras["serving"] = serving_case()
    if "strided_sizes" in cases:
        extras["strided_sizes"] = strided_sizes_case(short, xw)

    pts = N ** 3
    rec = {
        "metric": METRIC,
        "value": pts / head["mean_s"],
        "unit": "grid pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["mean_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (sin(2pi x)sin(2pi y)sin(2pi z) + N(0,1e-3), numpy default_rng(0), host-made)",
        "config": {"workload": "compact 4th-order d/dx, 512^3 FP64, 262144 systems x 512",
                   "n": N, "systems": N * N, "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": head["gbps"], "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": head["hbm_frac"], "traffic": measured_traffic(),
                     "alg_bytes_per_launch": head["alg_bytes"], "kernel_mean_us": head["mean_s"] * 1e6},
        "max_rel_err_vs_oracle": head_err,
        "err_sample": "1024 x-lines of the same device field, oracle = SciPy banded (LAPACK gtsv)",
        "served": served_summary(extras.get("serving")),
        "cpu_baseline": cpu,
        "extras": {k: v for k, v in extras.items() if k != "serving"},
    }
    if tuned:
        rec["tuning"] = tuned
    full = json.dumps(rec)
    log(f"[bench] detail {full}")
    # The stdout line: the same record with the extra cases, the served stream and
    # the CPU baseline in short form (about 2.5 KB).
    line_rec = dict(rec)
    line_rec["served"] = compact_served(rec["served"])
    line_rec["extras"] = compact_extras(rec["extras"])
    if cpu:
        line_rec["cpu_baseline"] = {k: cpu[k] for k in ("value", "unit", "cores", "kind", "blas_threads",
                                                        "single_core_pts_per_s") if k in cpu}
        line_rec["cpu_baseline"]["sample"] = (
            "d/dx of all x-lines of the benchmarked 512^3 field, one LAPACK gtsv thread per process; "
            "single core: 65536 lines per direction")
    small = json.dumps(sig(line_rec), separators=(",", ":"))
    log(f"[bench] stdout line: {len(small)} bytes")
    print(small, flush=True)


if __name__ == "__main__":
    main()
