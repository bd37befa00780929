"""Build libntoeplitz.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: python build.py [-j JOBS] [--verbose]
Each translation unit compiles in parallel into ../_build/, then links into
../lib/libntoeplitz.so.  Rebuilds every time it is called.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libntoeplitz.so")

SOURCES = [
    "nt_plan.cpp",
    "nt_api.hip",
    "nt_aux_kernels.hip",
    "nt_k_ident_contig.hip",
    "nt_k_ident_strided.hip",
    "nt_k_compact64_contig.hip",
    "nt_k_compact64_strided.hip",
    "nt_k_compact32_contig.hip",
    "nt_k_compact32_strided.hip",
    "nt_k_adi_contig.hip",
    "nt_k_adi_strided.hip",
    "nt_part.hip", "nt_part_ops.hip",
    "nt_batch.hip",
]

ARCH = os.environ.get("NT_OFFLOAD_ARCH", "gfx950")
COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # packed FP32 instructions (v_pk_*_f32) stay off: some of their forms have
    # given wrong results on these GPUs without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def hipcc():
    return shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def compile_one(src, verbose, extra):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [hipcc()] + COMMON + extra + lang + ["-c", os.path.join(HERE, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{res.stdout}\n{res.stderr}")
    if res.stderr.strip() and verbose:
        print(res.stderr, flush=True)
    return obj, res.stderr


def build(jobs=8, verbose=False, extra=()):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    extra = list(extra)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(compile_one, s, verbose, extra) for s in SOURCES]
        objs = [f.result()[0] for f in futs]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}"] + objs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--resource-usage", action="store_true",
                    help="print per-kernel VGPR/SGPR/LDS/scratch usage")
    a = ap.parse_args()
    extra = ["-Rpass-analysis=kernel-resource-usage"] if a.resource_usage else []
    print(build(a.j, a.verbose or a.resource_usage, extra))
    sys.exit(0)
