"""Host C++ under the sanitizers (no device needed).

* The NPY header reader and chunk arithmetic (batchfit_amd/csrc/
  bf_npy_header.hpp) built with g++ under AddressSanitizer and
  UndefinedBehaviorSanitizer (every finding fatal) and fuzzed by
  tests/cpp/npy_fuzz.cpp: NumPy's headers and mutations of them, written to a
  file and read, and the chunk rule over dimensions at the limits of their
  integer types.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_npy_reader_under_address_and_ub_sanitizers(tmp_path):
    exe = tmp_path / "npy_fuzz"
    build = subprocess.run(
        ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         os.path.join(HERE, "cpp", "npy_fuzz.cpp"), "-o", str(exe)],
        capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    run = subprocess.run([str(exe), str(tmp_path), "20000"], capture_output=True, text=True,
                         env=env, timeout=300)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "failures 0" in run.stdout
    accepted = int(run.stdout.split("accepted ")[1].split(",")[0])
    assert accepted > 100  # the mutations leave many valid headers too
