"""Host builds of the oracle and of the C-ABI layer with AddressSanitizer and
UndefinedBehaviorSanitizer (tests/sanitize/Makefile), each run through a
driver that calls every entry point: the oracle on small inputs (results
cross-checked), the C ABI on invalid arguments (rejected before any HIP call,
so no GPU is needed).  A memory error or undefined behaviour aborts the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="4")


def build(target):
    if not shutil.which("gcc") or not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("needs gcc and hipcc")
    r = subprocess.run(["make", "-s", "-C", SAN, target], capture_output=True, text=True,
                       timeout=900)
    if r.returncode != 0 and "build the library first" in r.stdout + r.stderr:
        pytest.skip("libtsm objects not built (python -m tsm_benchmark_amd.build)")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return os.path.join(SAN, target)


@pytest.mark.parametrize("target", ["_build/oracle_checks", "_build/capi_checks"])
def test_sanitized_driver(target):
    exe = build(target)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
    assert "checks passed" in r.stdout
