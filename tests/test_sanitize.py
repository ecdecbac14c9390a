"""The library's host code (CPU backend, graph-file readers and writer, error
reporting) built with AddressSanitizer + UndefinedBehaviorSanitizer and with
ThreadSanitizer (tests/harness/Makefile, g++ only), then run over random
multigraphs and small files, each compared with the oracle
(tests/harness/sanitize_host.cpp).  Any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

HARNESS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "harness")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-C", HARNESS, "-j4"], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return os.path.join(HARNESS, "_san")


@pytest.mark.parametrize("kind", ["asan", "tsan"])
def test_host_code_under_sanitizer(built, kind, tmp_path):
    env = dict(os.environ, TMPDIR=str(tmp_path),
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(built, f"sanitize_{kind}")], capture_output=True,
                       text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "all checks passed" in r.stdout
    assert "Sanitizer" not in out and "runtime error" not in out, out[-4000:]
