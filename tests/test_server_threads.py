"""The continuous-batching server's threads (csrc/drh_server.cpp) on the CPU: the server
source is compiled with g++ against host-only stand-ins for the HIP runtime and for the
drh_* calls it makes (tests/server_stub/), under ThreadSanitizer and under AddressSanitizer,
and driven through uneven jobs, a failing job, budget refusal and servers destroyed while
admissions and steps are in flight (tests/server_stub/stress.cpp).  A data race, a use after
free, a leaked job handle or a wrong result fails the test.

The stand-ins write to the handle on every call without locking, so two threads using one
handle at once show up as a race; a job set up by an admitter after the stepper had stopped
shows up as a handle left alive after drh_server_destroy (the shutdown race this harness was
written for)."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
STUB = os.path.join(HERE, "server_stub")
SERVER = os.path.join(HERE, "..", "diffrender_tutorials_amd", "csrc", "drh_server.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
@pytest.mark.parametrize("admitters", [1, 4])
@pytest.mark.parametrize("sanitizer", ["thread", "address"])
def test_server_threading_under_sanitizer(tmp_path, sanitizer, admitters):
    exe = str(tmp_path / f"stress_{sanitizer}")
    cmd = ["g++", "-std=c++17", "-O1", "-g", f"-fsanitize={sanitizer}", "-fno-omit-frame-pointer",
           "-include", os.path.join(STUB, "tsan_cv.h"), "-I", STUB,
           os.path.join(STUB, "stress.cpp"), os.path.join(STUB, "drh_stub.cpp"), SERVER, "-o", exe, "-lpthread"]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 second_deadlock_stack=1", ASAN_OPTIONS="detect_leaks=1",
               DRH_SERVER_ADMITTERS=str(admitters))
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0 and "server stress ok" in r.stdout, (r.stdout[-2000:], r.stderr[-6000:])
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-6000:]
