"""The oracle's own CPU tests, re-run against the oracle built under
AddressSanitizer + UndefinedBehaviorSanitizer with every finding fatal
(`make -C oracle sanitize`).  The checker has to be free of memory errors and
undefined behaviour for its verdicts to mean anything; inputs include the
non-terminating line of the walk-guard test and non-finite coordinates."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime():
    try:
        out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_tests_pass_under_asan_and_ubsan():
    asan = _runtime()
    if asan is None:
        pytest.skip("gcc's AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ)
    env["OCHRE_ORACLE_SO"] = os.path.join(ROOT, "oracle", "_san", "libochre_oracle.so")
    # the interpreter is not instrumented: the runtime is preloaded into this
    # child only, and leaks of the interpreter's own allocations are not ours
    env["LD_PRELOAD"] = asan
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    files = [os.path.join(ROOT, "tests", f) for f in
             ("test_oracle.py", "test_oracle_stroke.py", "test_oracle_adjoint.py", "test_oracle_atlas.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        *files], cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "error" not in r.stdout.lower().split("passed")[-1], tail
