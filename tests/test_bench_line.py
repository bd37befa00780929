"""The bench's stdout record stays one short JSON line (CPU only).

A synthetic full record with every extra case is folded the way bench.main()
folds it; the line must parse, keep the contract's keys and stay under 3 KB.
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _extras():
    line = {"us": 425.123456, "gbps": 5042.1, "hbm_frac": 0.63, "err_vs_oracle": 4.8e-16}
    ex = {k: dict(line) for k in ("ddy_fp64", "ddz_fp64", "ddx_fp32_refine1", "ddx_fp32_refine2",
                                  "ddy_fp32_refine1", "ddz_fp32_refine1")}
    ex["adi_2048_200steps_graph"] = {
        "N": 2048, "r": 0.5, "steps": 200, "ms_per_step": 0.0495, "effective_gbps": 2710.0,
        "err_vs_oracle_1step": 6.7e-16, "partitioned_x_half_step_16x128_us": 30.6,
        "partitioned_x_half_step_one_row_per_workgroup_us": 54.0, "partitioned_y_half_step_16x128_us": 46.9,
        "partitioned_ms_per_step": 0.0723, "partitioned_err_vs_oracle_1step": 6.7e-16}
    ax = {"ms": {"fp32_refine": 2.3, "fp64": 3.2, "f32_store_fp64": 2.4},
          "err": {"fp32_refine": 2.1e-14, "fp64": 2.5e-16, "f32_store_fp64": 2.5e-16}, "speedup_vs_fp64": 1.39}
    ex["fp32_storage_1024"] = {"ddx": ax, "ddy": ax, "ddz": ax}
    b = {"grids": 64, "points": 1663631360, "launch_ms": 7.9, "gbps": 3361.0, "separate_launches_ms": 8.49,
         "speedup_vs_separate": 1.07, "err_vs_oracle_per_grid": {"max": 1.4e-13, "p50": 3e-14, "count": 64},
         "err_lines_per_grid": 32}
    ex["mixed_batch_64"] = b
    ex["mixed_batch_small_1024"] = dict(b, grids=1024)
    return ex


def _served():
    case = {m: {"p50_ms": 0.72, "p99_ms": 5.4, "throughput_rps": 2156.0, "mean_ms": 1.0}
            for m in ("continuous", "sequential")}
    serving = {"requests": 1000}
    for key in ("rate_500", "rate_2000", "rate_3000", "upload_rate_100", "upload_rate_400"):
        serving[key] = case
    return bench.served_summary(serving)


def test_stdout_line_is_short_and_complete():
    served = _served()
    ex = _extras()
    rec = {"metric": bench.METRIC, "value": 3.7e11, "unit": "grid pts/s", "n_gpus": 1, "steps": 20,
           "warmup": 5, "ms_per_step": 0.3587, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic", "config": {"workload": "w"},
           "roofline": {"bound": "hbm", "achieved": 5987.0, "peak": 8000.0, "unit": "GB/s", "frac": 0.748,
                        "traffic": 2.148e9},
           "served": served, "cpu_baseline": {"value": 5.7e8, "unit": "grid pts/s", "cores": 16, "kind": "port"},
           "extras": ex}
    line_rec = dict(rec, served=bench.compact_served(served), extras=bench.compact_extras(ex))
    line = json.dumps(bench.sig(line_rec), separators=(",", ":"))
    assert len(line) < 3000, len(line)
    back = json.loads(line)
    for key in ("metric", "value", "unit", "roofline", "cpu_baseline", "served", "extras"):
        assert key in back
    assert back["served"]["rate_2000"]["continuous"] == [0.72, 5.4, 2156.0]
    assert back["extras"]["us_err"]["ddz_fp64"] == [425.1, 4.8e-16]
    assert back["extras"]["fp32_storage_1024"]["ddx"][3] == 1.39
    assert back["extras"]["mixed_batch_64"][:3] == [64, 7.9, 8.49]
    assert "legend" in back["extras"]


def test_compact_helpers_accept_missing_cases():
    assert bench.compact_served(None) is None
    assert bench.compact_extras({}) == {}
