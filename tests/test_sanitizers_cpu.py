"""The host C code (the library's C layer: catalogs, text output, QPM
loader threads, tuning cache; and the FP64 oracle) under AddressSanitizer
and UndefinedBehaviorSanitizer: the CPU test suites run against builds made
with -fsanitize=address,undefined (host code only), with the sanitizer
runtimes preloaded into the Python process.  Any report fails the run."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    p = subprocess.check_output(["gcc", f"-print-file-name={name}"], text=True).strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_cpu_suites_under_asan_and_ubsan():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc sanitizer runtimes not installed")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"])
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "fftgal_amd", "csrc"), "asan"])
    env = dict(os.environ)
    env.update(
        LD_PRELOAD=f"{asan}:{ubsan}",
        # Python's own allocations are not freed at exit: leak reports off
        ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=86",
        UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1:exitcode=87",
        FFTGAL_LIB=os.path.join(ROOT, "fftgal_amd", "csrc", "_build", "asan",
                                "libfftgal_amd.so"),
        FFTGAL_ORACLE_LIB=os.path.join(ROOT, "oracle", "_asan", "liboracle.so"),
    )
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_cpu.py"),
                        os.path.join(ROOT, "tests", "test_library_cpu.py")],
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in tail and "runtime error:" not in tail, tail
