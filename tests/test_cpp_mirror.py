"""Builds and runs the C++ mirror test (tests/cpp/test_mirror.cpp) against
libndconv.so: plan and error checks on the CPU, convolutions on the GPU."""
import os
import subprocess

import pytest

from conftest import ROOT

LIBDIR = os.path.join(ROOT, "ndarray-conv_amd", "lib")
EXE = os.path.join(ROOT, "ndarray-conv_amd", "_build", "test_mirror")


def build():
    if not os.path.exists(os.path.join(LIBDIR, "libndconv.so")):
        pytest.skip("libndconv.so not built")
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    src = os.path.join(ROOT, "tests", "cpp", "test_mirror.cpp")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", f"-I{ROOT}/include", f"-I{ROOT}/ndarray-conv_amd/cpp", src,
           "-o", EXE, f"-L{LIBDIR}", "-lndconv", f"-Wl,-rpath,{LIBDIR}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return EXE


def test_cpp_mirror_plan_and_errors():
    exe = build()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_mirror_on_gpu(gpu):
    exe = build()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "no device" not in r.stdout, r.stdout + r.stderr
