"""Sanitizer runs of the host-only code (CPU, no GPU): builds of tests/sanitize/Makefile.

* The PLY readers (csrc/ply.cpp) under AddressSanitizer + UBSan on valid files of every layout
  and thousands of hostile mutations of them (tests/sanitize/ply_fuzz.cpp): refusals are fine,
  out-of-bounds access, undefined behaviour and allocations sized by unchecked counts are not
  (the last is what this run found in an earlier reader: a vertex count far beyond the data).
* The drop-in ThreadPool (host/src/ThreadPool.cpp) under ThreadSanitizer
  (tests/sanitize/threadpool_stress.cpp).
* The oracle (oracle/splat_oracle.cpp) under AddressSanitizer + UBSan: its own test file runs
  against the sanitizer build, with the ASan runtime preloaded into the interpreter.
A report makes the program exit non-zero (halt_on_error / -fno-sanitize-recover).
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
OUT = os.path.join(SAN, "_build")
REPORTS = ("ERROR: AddressSanitizer", "runtime error:", "WARNING: ThreadSanitizer")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                                reason="needs g++ and make")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", SAN, "-j4"], check=True, timeout=900)
    return OUT


def _assert_clean(p):
    text = p.stdout + p.stderr
    assert p.returncode == 0, text[-4000:]
    for r in REPORTS:
        assert r not in text, text[-4000:]


def test_ply_readers_under_asan_and_ubsan(built, tmp_path):
    p = subprocess.run([os.path.join(built, "ply_fuzz"), str(tmp_path), "3000"],
                       capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1"))
    _assert_clean(p)
    assert "3000 mutants" in p.stdout


def test_threadpool_under_tsan(built):
    p = subprocess.run([os.path.join(built, "threadpool_stress")], capture_output=True, text=True,
                       timeout=600, env=dict(os.environ, TSAN_OPTIONS="halt_on_error=1"))
    _assert_clean(p)
    assert "0 failures" in p.stdout


def test_oracle_under_asan_and_ubsan(built):
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    assert os.path.isabs(asan), "ASan runtime not found"
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               U3D_ORACLE_SO=os.path.join(built, "liboracle_splat.so"))
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-p", "no:xdist", "-m", "not gpu and not slow",
                        os.path.join(ROOT, "tests", "test_oracle.py")],
                       capture_output=True, text=True, timeout=900, cwd=ROOT, env=env)
    _assert_clean(p)
    assert " passed" in p.stdout
