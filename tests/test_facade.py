"""The C++ façade (include/batchfit/fit.hpp) compiled with the host compiler
and run: argument errors without a device (CPU), and with a device the
synchronous fit plus batchfit::submit on a stream of the caller's
(tests/cpp/facade_check.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "batchfit_amd")


def build(tmp_path):
    exe = str(tmp_path / "facade_check")
    subprocess.run(["g++", "-std=c++20", "-O1", "-D__HIP_PLATFORM_AMD__",
                    "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                    os.path.join(ROOT, "tests", "cpp", "facade_check.cpp"), "-o", exe,
                    "-L", LIB, "-lbatchfit_hip", f"-Wl,-rpath,{LIB}", "-L", "/opt/rocm/lib",
                    "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return exe


def test_facade_argument_errors(tmp_path):
    import batchfit_amd
    exe = build(tmp_path)
    if batchfit_amd.hip_available():
        pytest.skip("a device is present: the GPU test runs the whole check")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "facade ok (no device)" in out.stdout, out.stdout + out.stderr


@pytest.mark.gpu
def test_facade_fit_and_submit_on_caller_stream(tmp_path):
    exe = build(tmp_path)
    out = subprocess.run([exe, "device"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and "facade ok (device)" in out.stdout, out.stdout + out.stderr


def test_coordinates_header(tmp_path):
    """include/batchfit/coordinates.hpp compiled and run on the host."""
    exe = str(tmp_path / "coordinates_check")
    subprocess.run(["g++", "-std=c++20", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "coordinates_check.cpp"), "-o", exe,
                    "-L", LIB, "-lbatchfit_hip", f"-Wl,-rpath,{LIB}"], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "coordinates ok" in out.stdout, out.stdout + out.stderr
