"""The request loop (nevpt2_amd/csrc/nv2_serve.cpp) on the CPU, under ThreadSanitizer and
AddressSanitizer, with stub evaluation contexts (tests/serve_stub): several submitting threads,
mixed shapes, pooled and evicted contexts under a tight HBM budget, one slot and four, and requests
whose context run fails or throws.  Regression test for two faults of the loop: an exception in a
request ended the process (std::terminate, exit status 134, no HIP message), and contexts evicted
for a request that then did not fit were never freed."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "nevpt2_amd", "csrc", "nv2_serve.cpp"),
       os.path.join(ROOT, "tests", "serve_stub", "stub_ctx.cpp"),
       os.path.join(ROOT, "tests", "serve_stub", "driver.cpp")]


@pytest.mark.parametrize("san", ["thread", "address,undefined"])
def test_request_loop_under_sanitizer(tmp_path, san):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "serve_driver")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=" + san, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           *SRC, "-pthread", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1")
    for _ in range(3):
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
        assert r.returncode == 0 and "OK" in r.stdout, r.stdout + r.stderr[-3000:]
        assert "WARNING: ThreadSanitizer" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
        assert "leaked" not in r.stderr
