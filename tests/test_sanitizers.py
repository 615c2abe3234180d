"""The host library (mesh generation, geometry, layout packing, the C ABI)
and the oracle built with AddressSanitizer and UndefinedBehaviorSanitizer,
driven by the CPU tests that exercise them: any out-of-bounds access, use
after free, leak in the C++ classes' lifetimes or undefined arithmetic
aborts the run.  (The kernels are the normal build; nothing here touches a
GPU.)"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "muscl-scheme-on-spherical-mesh-wip_amd", "csrc")
SAN_LIB = os.path.join(ROOT, "muscl-scheme-on-spherical-mesh-wip_amd", "_build_san", "libsplash_amd_san.so")
SAN_ORACLE = os.path.join(ROOT, "oracle", "liboracle_san.so")
CPU_TESTS = ["tests/test_capi.py", "tests/test_oracle_mesh.py", "tests/test_oracle_mesh_numbering.py",
             "tests/test_oracle_physics.py", "tests/test_oracle_source_reconstruct.py"]


def _runtime(name):
    out = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_cpu_suite_under_asan_ubsan():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not installed")
    jobs = str(min(16, os.cpu_count() or 1))
    subprocess.run(["make", "-s", "-j", jobs, "-C", CSRC, "sanitize"], check=True, capture_output=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, capture_output=True)
    env = dict(os.environ, SPLASH_AMD_LIB=SAN_LIB, SPLASH_ORACLE_LIB=SAN_ORACLE, LD_PRELOAD=f"{asan} {ubsan}",
               ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"]
                       + CPU_TESTS, cwd=ROOT, env=env, capture_output=True, text=True, timeout=1500)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
