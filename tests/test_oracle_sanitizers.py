"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle's own tests (tests/test_oracle.py, tests/test_oracle_interp.py) run
again in a subprocess against a sanitised build of oracle/*.c (host code,
gcc), with the ASan runtime preloaded into the interpreter.  Any invalid
access or undefined behaviour aborts that run."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    try:
        p = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True,
                           timeout=60).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_tests_under_asan_and_ubsan(tmp_path):
    asan = _runtime("libasan.so")
    if not asan or not shutil.which("make"):
        pytest.skip("gcc's AddressSanitizer runtime or make is not available")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize",
                           "OUT=" + str(tmp_path)], timeout=300)
    env = dict(os.environ, OAK_ORACLE_LIB=str(tmp_path / "liboak_oracle_san.so"),
               LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu", "tests/test_oracle.py", "tests/test_oracle_interp.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert " passed" in r.stdout
