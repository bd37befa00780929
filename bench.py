"""Benchmark of the near-Toeplitz hot path on one MI355X.

Headline workload: 4th-order compact d/dx of a 512^3 FP64 field
f = sin(2 pi x) sin(2 pi y) sin(2 pi z) + N(0, 1e-3) noise (seed 0), i.e.
262,144 near-Toeplitz systems of size 512, one fused launch per step (RHS
stencil + cyclic reduction, each point read once and written once).
Metric: grid points per second and the fraction of HBM peak.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--quick] [--case NAME ...]

Prints ONE JSON line.  Other configurations are timed after the headline and
reported under "extras": d/dy and d/dz ("axes"), FP32 + refinement sweeps
along x and one sweep along y and z ("refine"), ADI 2048^2 x 200 steps in a
hipGraph ("adi"), a 1024^3 field stored in FP32 ("fp32_storage"), 64
mixed-size grids in one persistent launch
("batch"), and a Poisson request stream served with continuous batching
("serving").  --quick skips them; --case runs only the named ones.
"""
import argparse
import json
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


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


def compact_case(N, axis, precision, sweeps, in_dtype, steps, warmup, field=None):
    import torch
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.workloads import compact_field_device
    f = field if field is not None else compact_field_device(N, seed=0, dtype=in_dtype)
    out = torch.empty(f.shape, dtype=torch.float64, device=f.device)
    plan = Plan(N, COMPACT_COEFFS)
    h = 1.0 / (N - 1)
    t = time_launches(lambda: core.compact_derivative(plan, f, out, h, axis, precision, sweeps),
                      steps, warmup)
    pts = N ** 3
    nbytes = pts * (f.element_size() + 8)
    t["pts_per_s"] = pts / t["mean_s"]
    t["alg_bytes"] = nbytes
    t["gbps"] = nbytes / t["mean_s"] / 1e9
    t["hbm_frac"] = nbytes / t["mean_s"] / HBM_PEAK
    return t, f, out


def cpu_baseline(N, sample_lines, reps=3):
    """SciPy banded solve (LAPACK gtsv) + NumPy RHS on a sample of the 512^3
    d/dx workload's lines: the CPU path of the same matrices."""
    import numpy as np
    from oracle.compact import COMPACT_COEFFS, compact_rhs
    from oracle.cr import banded_solve
    rng = np.random.default_rng(0)
    x = np.linspace(0.0, 1.0, N)
    ys = rng.random(sample_lines)
    zs = rng.random(sample_lines)
    f = (np.sin(2 * np.pi * zs)[:, None] * np.sin(2 * np.pi * ys)[:, None]
         * np.sin(2 * np.pi * x)[None, :]) + rng.normal(0.0, 1e-3, (sample_lines, N))
    h = 1.0 / (N - 1)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        banded_solve(COMPACT_COEFFS, compact_rhs(f, h))
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    try:
        from threadpoolctl import threadpool_info
        blas_threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        blas_threads = None
    return {"value": sample_lines * N / med, "unit": "grid pts/s", "cores": 1,
            "kind": "port",
            "sample": f"{sample_lines} x-lines of the 512^3 d/dx shape, y and z drawn at random "
                      f"({sample_lines * N} pts), NumPy RHS + one scipy.linalg.solve_banded "
                      f"call (LAPACK gtsv, single-threaded), median of {reps}",
            "seconds": med, "host_cores_available": len(os.sched_getaffinity(0)),
            "blas_threads": blas_threads}


def adi_case(N, r, nsteps, reps):
    import torch
    from adi.solver import ADISolver
    from ntamd.workloads import gaussian_ic
    u = torch.tensor(gaussian_ic(N, seed=0), device="cuda")
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
    # the one-GPU partitioned form: 16 subdomains of 128 per line plus the reduced
    # 32-unknown system, one launch per x-implicit half step
    from ntamd import adi_coeffs
    from ntamd.partition import PartitionPlan
    pp = PartitionPlan(N, 16, adi_coeffs(r))
    w = torch.empty_like(u)
    tp = time_launches(lambda: pp.adi_half_x(u, w, r), 20, 3)
    return {"N": N, "r": r, "steps": nsteps, "s_per_run": t, "ms_per_step": 1e3 * t / nsteps,
            "effective_gbps": 2 * nsteps * half_bytes / t / 1e9,
            "partitioned_x_half_step_16x128_us": tp["mean_s"] * 1e6,
            "note": "32 MiB field stays in the 256 MB Infinity Cache: effective, not HBM, GB/s"}


def measured_traffic():
    """HBM bytes per launch of the headline kernel from the committed rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE passes (scaled by a 1 GiB copy's readings)."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            return json.load(fh)["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def pmc_probe():
    """Three headline launches, then three copies of the same 1 GiB field: the
    program profiled by the FETCH_SIZE / WRITE_SIZE passes."""
    import torch
    from ntamd import COMPACT_COEFFS, Plan, core
    N = 512
    f = torch.randn((N, N, N), dtype=torch.float64, device="cuda")
    out = torch.empty_like(f)
    plan = Plan(N, COMPACT_COEFFS)
    for _ in range(3):
        core.compact_derivative(plan, f, out, 1.0 / (N - 1), 2)
    for _ in range(3):
        out.copy_(f)
    torch.cuda.synchronize()
    print("pmc probe done", flush=True)


def config4_case(N=1024, reps=5):
    """1024^3 field stored in FP32 (4 GiB): d/dx, d/dy, d/dz with the FP32 solve +
    one FP64-residual sweep, against plain FP64 (FP64-stored field, FP64 solve).
    Errors are relative to the FP64 solve of the same (FP32-stored) values, which
    is timed too: FP32 storage with FP64 arithmetic."""
    import torch
    from ntamd import COMPACT_COEFFS, Plan, core
    from ntamd.workloads import compact_field_device
    # Blocks cached from the earlier cases go back to the driver first, so the
    # 1024^3 fields are fresh allocations.
    torch.cuda.empty_cache()
    f32 = compact_field_device(N, seed=0, dtype=torch.float32)
    f64 = f32.double()
    out_r = torch.empty((N, N, N), dtype=torch.float64, device=f32.device)
    out_64 = torch.empty_like(out_r)
    plan = Plan(N, COMPACT_COEFFS)
    h = 1.0 / (N - 1)
    res = {}
    for axis, name in ((2, "ddx"), (1, "ddy"), (0, "ddz")):
        tr = time_launches(lambda: core.compact_derivative(plan, f32, out_r, h, axis, "fp32_refine", 1),
                           reps, 2)
        t64 = time_launches(lambda: core.compact_derivative(plan, f64, out_64, h, axis, "fp64", 1), reps, 2)
        ref64 = torch.empty_like(out_r)
        # FP64 solve of the same FP32-stored values: the error reference, and
        # timed as the third form (FP32 storage, FP64 arithmetic)
        ts = time_launches(lambda: core.compact_derivative(plan, f32, ref64, h, axis, "fp64", 1), reps, 2)
        err = ((out_r - ref64).abs().max() / ref64.abs().max()).item()
        res[name] = {"fp32_refine_s": tr["mean_s"], "fp64_s": t64["mean_s"],
                     "speedup_vs_fp64": t64["mean_s"] / tr["mean_s"], "max_rel_err_vs_fp64": err,
                     "fp32_refine_gbps": N ** 3 * 12 / tr["mean_s"] / 1e9,
                     "fp64_gbps": N ** 3 * 16 / t64["mean_s"] / 1e9,
                     "fp32_storage_fp64_solve_s": ts["mean_s"],
                     "fp32_storage_fp64_solve_speedup_vs_fp64": t64["mean_s"] / ts["mean_s"]}
        log(f"[bench] 1024^3 {name}: fp32+ref {tr['mean_s'] * 1e3:.3f} ms, fp64 {t64['mean_s'] * 1e3:.3f} ms, "
            f"x{res[name]['speedup_vs_fp64']:.2f}, err {err:.2e}; fp32 storage + fp64 solve "
            f"{ts['mean_s'] * 1e3:.3f} ms, x{res[name]['fp32_storage_fp64_solve_speedup_vs_fp64']:.2f}")
        del ref64
    del f32, f64, out_r, out_64
    torch.cuda.empty_cache()
    return res


def mixed_batch_case(count=64, seed=1, reps=5, sizes=None, precision="fp32_refine"):
    """64 grids, per-axis sizes uniform in {64, 96, ..., 512} (seeded), random
    derivative direction each, FP32 solve + one sweep, all in one persistent launch,
    against the same jobs as separate fused launches.  Per-grid max relative
    error against the plain FP64 kernel."""
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
    plans = {n: Plan(n, COMPACT_COEFFS) for n in {j["f"].shape[j["axis"]] for j in jobs}}

    def separate():
        for j in jobs:
            core.compact_derivative(plans[j["f"].shape[j["axis"]]], j["f"], j["out"], j["dx"], j["axis"],
                                    j["precision"], j["sweeps"])
    ts = time_launches(separate, reps, 2)
    errs = []
    for j in jobs:
        ref = torch.empty_like(j["out"])
        n = j["f"].shape[j["axis"]]
        core.compact_derivative(Plan(n, COMPACT_COEFFS), j["f"], ref, j["dx"], j["axis"], "fp64", 1)
        errs.append(((j["out"] - ref).abs().max() / ref.abs().max()).item())
    # medians: a launch sequence that pays a one-off cost (a code object loaded,
    # an occupancy query answered the first time) does not set the comparison
    res = {"grids": count, "points": pts, "launch_s": t["median_s"], "pts_per_s": pts / t["median_s"],
           "gbps": pts * 16 / t["median_s"] / 1e9, "separate_launches_s": ts["median_s"],
           "launch_s_min_max": [t["min_s"], t["max_s"]], "separate_s_min_max": [ts["min_s"], ts["max_s"]],
           "max_rel_err_vs_fp64_per_grid_max": max(errs),
           "max_rel_err_vs_fp64_per_grid": errs}
    log(f"[bench] mixed batch: {count} grids, {pts / 1e6:.0f} M pts, {t['median_s'] * 1e3:.3f} ms "
        f"(min {t['min_s'] * 1e3:.3f}, max {t['max_s'] * 1e3:.3f}; {res['gbps']:.0f} GB/s); {count} separate "
        f"launches {ts['median_s'] * 1e3:.3f} ms (min {ts['min_s'] * 1e3:.3f}, max {ts['max_s'] * 1e3:.3f}); "
        f"worst grid err {max(errs):.2e}")
    del jobs, b
    torch.cuda.empty_cache()
    return res


def serving_case(count=1000, rates=(500.0, 2000.0), seed=2):
    """Poisson arrivals (rate req/s, numpy default_rng(2)): compact derivatives
    of random-phase trig fields (per-axis sizes U[32,512], random axis) and ADI
    runs (side U[256,2048], 1-20 steps, r U[0.1,2]), FP64 or FP32+refinement
    each; continuous batching vs one-at-a-time plain FP64, same request stream,
    each on a warmed-up server."""
    from ntamd.serve import Server, make_requests
    out = {"requests": count, "seed": seed}
    for rate in rates:
        case = {}
        for mode in ("continuous", "sequential"):
            reqs = make_requests(count, rate, seed=seed)
            srv = Server()
            srv.warmup()                 # first-launch costs are paid outside the timed stream
            stats = srv.run(reqs, mode=mode)
            case[mode] = stats
            log(f"[bench] serving {rate:.0f} req/s {mode}: p50 {stats['p50_ms']:.2f} ms, "
                f"p99 {stats['p99_ms']:.2f} ms, {stats['throughput_rps']:.0f} req/s")
        out[f"rate_{rate:.0f}"] = case
    return out


def served_summary(serving):
    """p50/p99 of the served request stream at each stated arrival rate, continuous
    batching beside one-at-a-time plain FP64 (None when the case did not run)."""
    if not serving:
        return None
    out = {"requests": serving["requests"], "arrivals": "Poisson, numpy default_rng(2)"}
    for key, case in serving.items():
        if key.startswith("rate_"):
            out[key] = {m: {k: case[m][k] for k in ("p50_ms", "p99_ms", "throughput_rps")}
                        for m in ("continuous", "sequential")}
    return out


EXTRA_CASES = ("axes", "refine", "adi", "fp32_storage", "batch", "serving")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="headline only, no extras/CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run the short program profiled for HBM traffic and exit")
    ap.add_argument("--case", action="append", choices=EXTRA_CASES,
                    help="run only these extra cases after the headline (repeatable)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (a derivative's lines "
              "shard across GPUs without communication; run one process per GPU on its share)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from ntamd import _lib, autotune
    _lib.load()
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(1)
    tuned = autotune.apply()          # launch shapes chosen by the autotuner, if cached

    if args.pmc_probe:
        pmc_probe()
        return

    N = 512
    log(f"[bench] headline: compact d/dx {N}^3 FP64, steps={args.steps} warmup={args.warmup}")
    head, f64, out = compact_case(N, 2, "fp64", 1, torch.float64, args.steps, args.warmup)
    log(f"[bench] d/dx: {head['mean_s'] * 1e6:.1f} us/launch, {head['gbps']:.0f} GB/s")

    extras = {}
    cases = set(args.case) if args.case else (set() if args.quick else set(EXTRA_CASES))
    short = max(10, args.steps // 2)
    if "axes" in cases:
        for name, axis in (("ddy_fp64", 1), ("ddz_fp64", 0)):
            t, _, _ = compact_case(N, axis, "fp64", 1, torch.float64, short, 3, field=f64)
            extras[name] = {k: t[k] for k in ("mean_s", "pts_per_s", "gbps", "hbm_frac")}
            log(f"[bench] {name}: {t['mean_s'] * 1e6:.1f} us, {t['gbps']:.0f} GB/s")
    if "refine" in cases:
        for name, sweeps in (("ddx_fp32_refine1", 1), ("ddx_fp32_refine2", 2)):
            t, _, o2 = compact_case(N, 2, "fp32_refine", sweeps, torch.float64, short, 3, field=f64)
            err = ((o2 - out).abs().max() / out.abs().max()).item()
            extras[name] = {k: t[k] for k in ("mean_s", "pts_per_s", "gbps", "hbm_frac")}
            extras[name]["max_rel_err_vs_fp64"] = err
            log(f"[bench] {name}: {t['mean_s'] * 1e6:.1f} us, err {err:.2e}")
        # the strided axes with one sweep, against the FP64 kernel of the same axis
        from ntamd import COMPACT_COEFFS, Plan, core
        ref = torch.empty_like(out)
        for name, axis in (("ddy_fp32_refine1", 1), ("ddz_fp32_refine1", 0)):
            core.compact_derivative(Plan(N, COMPACT_COEFFS), f64, ref, 1.0 / (N - 1), axis, "fp64", 1)
            t, _, o2 = compact_case(N, axis, "fp32_refine", 1, torch.float64, short, 3, field=f64)
            err = ((o2 - ref).abs().max() / ref.abs().max()).item()
            extras[name] = {k: t[k] for k in ("mean_s", "pts_per_s", "gbps", "hbm_frac")}
            extras[name]["max_rel_err_vs_fp64"] = err
            log(f"[bench] {name}: {t['mean_s'] * 1e6:.1f} us, err {err:.2e}")
            del o2
        del ref
    del out, f64
    torch.cuda.empty_cache()
    if "adi" in cases:
        a = adi_case(2048, 0.5, 200, 3)
        extras["adi_2048_200steps_graph"] = a
        log(f"[bench] ADI 2048^2: {a['ms_per_step']:.3f} ms/step; partitioned x half step "
            f"(16 x 128 + reduced 32): {a['partitioned_x_half_step_16x128_us']:.1f} us")
    if "fp32_storage" in cases:
        extras["fp32_storage_1024"] = config4_case()
    if "batch" in cases:
        extras["mixed_batch_64"] = mixed_batch_case()
        extras["mixed_batch_small_1024"] = mixed_batch_case(count=1024, sizes=np.arange(16, 65, 8), reps=5)
    if "serving" in cases:
        extras["serving"] = serving_case()

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (SciPy banded) ...")
        cpu = cpu_baseline(N, 262144)     # every x-line of the 512^3 workload, ~3 s per repetition
        log(f"[bench] CPU: {cpu['value']:.3e} pts/s")

    pts = N ** 3
    rec = {
        "metric": "Grid pts/s & % HBM peak, 512^3 compact d/dx; p50/p99 served vs plain FP64",
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
        "data": "synthetic (sin(2pi x)sin(2pi y)sin(2pi z) + N(0,1e-3), torch seed 0, on device)",
        "config": {"workload": "compact 4th-order d/dx, 512^3 FP64, 262144 systems x 512",
                   "n": N, "systems": N * N, "coeffs": [1, 2, 0.25, 1, 0.25, 2, 1],
                   "precision": "fp64"},
        "roofline": {"bound": "hbm", "achieved": head["gbps"], "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": head["hbm_frac"], "traffic": measured_traffic(),
                     "traffic_source": "profiles/compact_ddx_512_fp64.md (rocprofv3 --pmc)",
                     "alg_bytes_per_launch": head["alg_bytes"],
                     "kernel_mean_s": head["mean_s"], "kernel_median_s": head["median_s"]},
        "hbm_peak_frac": head["hbm_frac"],
        "served": served_summary(extras.get("serving")),
        "cpu_baseline": cpu,
        "wall_s_timed_region": head["wall_s"],
        "extras": extras,
        "tuning": tuned,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
