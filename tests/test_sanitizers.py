"""CPU sanitizer runs of the host code: the library's host sources (C ABI,
threaded BVH builder, host traversal) with the C++ adapter in host-only mode,
and the oracle's threaded entry points, under ThreadSanitizer and under
AddressSanitizer + UndefinedBehaviorSanitizer.  Recipe:
tests/cpp/sanitize/Makefile (no GPU code is instrumented)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "cpp", "sanitize")
BIN = os.path.join(ROOT, "tests", "cpp", "bin")
REPORTS = ("WARNING: ThreadSanitizer", "ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer")


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(os.path.join(ROOT, "kazen_proto_amd", "_build", "trace.o")):
        pytest.skip("library objects not built (run __graft_entry__.build())")
    r = subprocess.run(["make", "-s", "-j4", "-C", SAN], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.parametrize("prog,args", [("adapter_tsan", ["host"]), ("adapter_asan", ["host"]),
                                       ("oracle_tsan", []), ("oracle_asan", [])])
def test_sanitized_run_is_clean(built, prog, args):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([os.path.join(BIN, prog)] + args, capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    print(out[-2000:])
    assert r.returncode == 0
    assert not any(k in out for k in REPORTS)
    assert " 0\n" in r.stdout   # mismatch counts are zero
