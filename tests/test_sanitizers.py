"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the FP64
oracle (every operator, chunk width and thread count on small geometries) and
the product's host planner (plan.cpp over accepted and rejected geometries).
Both are built with g++ from source and run as child processes; a sanitizer
report makes the process exit non-zero."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _run(exe):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    return r.stdout


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_selftest")
    subprocess.run(["g++", *SAN, "-fopenmp", "-o", exe, os.path.join(ROOT, "oracle", "selftest.cpp"),
                    os.path.join(ROOT, "oracle", "ct_oracle.cpp")], check=True, timeout=300)
    assert "oracle selftest ok" in _run(exe)


def test_host_plan_under_sanitizers(tmp_path):
    exe = str(tmp_path / "plan_selftest")
    subprocess.run(["g++", *SAN, "-o", exe, os.path.join(ROOT, "tests", "native", "plan_sanitize.cpp"),
                    os.path.join(ROOT, "ctlib_amd", "csrc", "plan.cpp")], check=True, timeout=300)
    assert "plan selftest ok" in _run(exe)
