"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/Makefile's liboracle_san.so is the same sources built with
-fsanitize=address,undefined and no recovery; the oracle's own tests (the
closed-form, golden and statistical pins, the intersection and the outlier
finder) run on it in a child process with libasan preloaded, so any
out-of-bounds access, use after free or undefined operation in the oracle
ends the child with an error.
"""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE = os.path.join(ROOT, "oracle")


def _runtime(name):
    try:
        out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True,
                             text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return out if os.path.isabs(out) and os.path.exists(out) else None


@pytest.fixture(scope="module")
def sanitized_env():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("no libasan for g++")
    subprocess.run(["make", "-s", "-C", ORACLE, "liboracle_san.so"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, KFO_LIBRARY=os.path.join(ORACLE, "liboracle_san.so"),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4",
               PYTHONPATH=ROOT + os.pathsep + HERE)
    return env


def test_sanitized_build_is_the_one_loaded(sanitized_env):
    out = subprocess.run([sys.executable, "-c", "import oracle; print(oracle._lib()._name)"],
                         env=sanitized_env, cwd=ROOT, capture_output=True, text=True)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.strip().endswith("liboracle_san.so"), out.stdout


def test_oracle_tests_pass_under_sanitizers(sanitized_env):
    files = [os.path.join(HERE, f) for f in ("test_oracle.py", "test_intersect_oracle.py",
                                             "test_outliers.py")]
    out = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                          "-m", "not gpu", *files], env=sanitized_env, cwd=ROOT,
                         capture_output=True, text=True, timeout=900)
    tail = (out.stdout + out.stderr)[-3000:]
    assert out.returncode == 0, tail
    assert "AddressSanitizer" not in tail and "runtime error" not in tail, tail
