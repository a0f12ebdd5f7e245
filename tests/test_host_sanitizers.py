"""The library's host code under AddressSanitizer + UndefinedBehaviorSanitizer
(and LeakSanitizer) and ThreadSanitizer, on the CPU: batch planning, the
request loop's scheduler / worker / submitting threads with pause, cancel and
admission rejection, codec set-up, per-call paths and the stream encoder, built
against a host stand-in for the HIP runtime whose GF(2^8) launchers compute
their results on the host (reedsolomon_amd/csrc/hoststub).  Each program
checks its own results and exits 0 only when every check passed and the
sanitizer reported nothing."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
STUB = os.path.join(os.path.dirname(HERE), "reedsolomon_amd", "csrc", "hoststub")
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not os.path.isdir(STUB) or shutil.which("g++") is None,
                                reason="host sanitizer sources or g++ not present")


def _make(target):
    r = subprocess.run(["make", "-j8", target], cwd=STUB, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]


def _run(prog, env):
    r = subprocess.run([os.path.join(STUB, "_build", prog)], cwd=STUB, capture_output=True, text=True,
                       timeout=600, env={**os.environ, **env})
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "Sanitizer" not in out, out[-6000:]
    for step in ("per-call paths", "batch plans", "request loop", "streams"):
        assert f"{step} done (0 failures)" in out, out[-3000:]


def test_asan_ubsan_host_paths():
    _make("asan")
    _run("rsx_host_asan", {"ASAN_OPTIONS": "detect_leaks=1",
                           "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"})


@pytest.mark.skipif(not os.path.exists(CLANG), reason="needs ROCm's clang++ for its TSan runtime")
def test_tsan_request_loop_and_planners():
    _make("tsan")
    _run("rsx_host_tsan", {"TSAN_OPTIONS": "halt_on_error=1:second_deadlock_stack=1"})
