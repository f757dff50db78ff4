"""The C ABI's host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

tests/sanitize/Makefile builds the library with both sanitizers on its host code (`-Xarch_host`) and
the driver tests/sanitize/abi_host_check.cpp, which makes every entry point refuse bad arguments,
runs the host-only calls (layouts for every dims up to 1024, autotuner text, instrumentation, truncated
buffers) and, without a GPU, checks that device calls fail with an error code.  The first build takes
a few minutes; make keeps the objects in tests/sanitize/_build."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
OUT = os.path.join(SAN, "_build")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC) or shutil.which("make") is None, reason="needs hipcc and make")
def test_abi_host_paths_under_asan_ubsan():
    jobs = str(max(1, min(8, len(os.sched_getaffinity(0)))))
    r = subprocess.run(["make", "-j" + jobs, "-C", SAN, "OUT=" + OUT], capture_output=True, text=True, timeout=1200)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:protect_shadow_gap=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([os.path.join(OUT, "abi_host_check")], capture_output=True, text=True, timeout=300, env=env)
    out = p.stdout + p.stderr
    assert p.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert " 0 failures" in out, out[-2000:]
