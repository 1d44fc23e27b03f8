"""Host-side AddressSanitizer + UndefinedBehaviorSanitizer builds: the CPU
oracle (every function, small seeded scenes, invariants) and the host side of
the HIP library's C ABI (sizes, layouts, argument checks), built by
tests/sanitize/Makefile and run as their own processes."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.fixture(scope="module")
def built():
    if not shutil.which("g++") or not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("needs g++ and hipcc")
    subprocess.run(["make", "-s", "-C", HERE, "-j2"], check=True, capture_output=True, text=True)
    return os.path.join(HERE, "_build")


@pytest.mark.parametrize("driver", ["oracle_driver", "host_abi_driver"])
def test_sanitized_driver(built, driver):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    res = subprocess.run([os.path.join(built, driver)], capture_output=True, text=True, env=env, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "ok" in res.stdout
    assert "runtime error" not in res.stderr and "AddressSanitizer" not in res.stderr
