"""The host plan builder (csrc/nt_plan.cpp) under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/host/plan_asan.cpp, built with g++, builds
plans for every n in [2, 4096] over six matrix families, feeds the builder
invalid input and exact-length error buffers, and checks the level table's
structure.  Any sanitizer report fails the run (halt on error, no recovery)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "near-toeplitz-applications_amd", "csrc")
HARNESS = os.path.join(ROOT, "tests", "host", "plan_asan.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_plan_builder_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "plan_asan")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", CSRC, HARNESS, os.path.join(CSRC, "nt_plan.cpp"), "-o", exe]
    build = subprocess.run(cmd, capture_output=True, text=True)
    if build.returncode != 0 and ("-lasan" in build.stderr or "-lubsan" in build.stderr):
        pytest.skip("sanitizer runtime not available: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stdout[-2000:] + run.stderr[-4000:]
    assert "runtime error" not in run.stderr and "AddressSanitizer" not in run.stderr, run.stderr[-4000:]
    assert "failures 0" in run.stdout, run.stdout
