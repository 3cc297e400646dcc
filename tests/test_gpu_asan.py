"""The library's host code under AddressSanitizer (tests/asan/driver.cpp
against the host-ASan build, libmadsdp_hip_asan.so): solver set-up and
teardown, every component combination, multi-block recorded stages, the
serving loop's worker threads with 24 concurrent requests, the API's error
paths.  Kernels are not instrumented (device ASan is not available); the
HIP runtime's own allocations are exempt from leak checking."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "asan", "driver")


def test_host_asan_driver(gpu):
    if not os.path.exists(DRIVER):
        pytest.skip("ASan driver not built (madsdp_amd.build.build_asan_driver)")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:protect_shadow_gap=0:halt_on_error=1")
    print("host-ASan driver", flush=True)
    r = subprocess.run([DRIVER], capture_output=True, text=True, env=env, timeout=240)
    out = r.stdout + r.stderr
    print(out[-3000:], flush=True)
    assert "AddressSanitizer" not in out
    assert r.returncode == 0 and "ASAN DRIVER OK" in r.stdout
