"""The C++ drop-in (cubismnova_amd/include/Cubism/AMD.h): it compiles with
g++ against the C ABI only (CPU), and the reference's unit tests written
against it pass on the GPU (tests/cpp/test_shim.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cubismnova_amd")


def build(out):
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-Wall", "-Werror", "-I" + os.path.join(PKG, "include"),
           os.path.join(ROOT, "tests", "cpp", "test_shim.cpp"), "-o", out, "-L" + PKG, "-lcubismnova_amd",
           "-Wl,-rpath," + PKG]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_shim_compiles(tmp_path):
    build(str(tmp_path / "test_shim"))


@pytest.mark.gpu
def test_shim_reference_unit_tests_on_gpu(gpu, tmp_path):
    exe = str(tmp_path / "test_shim")
    build(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240, cwd=str(tmp_path))
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert " 0 failed" in r.stdout
