"""The C++ binding (include/fastbit_hip.hpp): it compiles and links against
the library with warnings on (CPU), and its self-checking program
(tests/cpp/wrapper_check.cpp: operators, index evaluation, attach,
sumBins, an expression batch, all against brute force) passes on the GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "libfastlib_amd", "lib")
SRC = os.path.join(ROOT, "tests", "cpp", "wrapper_check.cpp")


def build(out):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    if not os.path.exists(os.path.join(LIBDIR, "libfastbit_hip.so")):
        pytest.skip("library not built")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", LIBDIR, "-lfastbit_hip", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,/opt/rocm/lib", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_wrapper_compiles_and_links(tmp_path):
    build(str(tmp_path / "wrapper_check"))


@pytest.mark.gpu
def test_wrapper_program_on_gpu(tmp_path):
    exe = str(tmp_path / "wrapper_check")
    build(exe)
    print("wrapper_check", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and "wrapper ok" in r.stdout, r.stdout + r.stderr
