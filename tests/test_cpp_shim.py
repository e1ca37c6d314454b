"""The C++ drop-in (cubismnova_amd/include/Cubism/AMD.h and the headers
under the reference's file names that forward to it): it compiles with g++
against the C ABI only (CPU); the reference's unit tests written against it
pass on the GPU (tests/cpp/test_shim.cpp); and the reference's own Laplacian
benchmark application (test/sandbox/Benchmarks/Kernel/Laplacian/
Laplacian.cpp), compiled unchanged against the drop-in by oracle/Makefile,
runs its per-block host loop over GPU-assembled labs to the same bits as
the device operator."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cubismnova_amd")
REF_APP = os.path.join(ROOT, "oracle", "_ref", "laplacian_dropin")
REF_SRC = "/root/reference/test/sandbox/Benchmarks/Kernel/Laplacian/Laplacian.cpp"


def build(out):
    cmd = ["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-pthread", "-Wall", "-Werror",
           "-I" + os.path.join(PKG, "include"),
           os.path.join(ROOT, "tests", "cpp", "test_shim.cpp"), "-o", out, "-L" + PKG, "-lcubismnova_amd",
           "-Wl,-rpath," + PKG]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_shim_compiles(tmp_path):
    build(str(tmp_path / "test_shim"))


@pytest.mark.skipif(not os.path.exists(REF_SRC), reason="the reference sources are not on this machine")
def test_reference_laplacian_compiles_against_dropin():
    """The reference application, unchanged, against the drop-in headers (the
    recipe of oracle/Makefile)."""
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), "ref"], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(REF_APP)


@pytest.mark.gpu
def test_shim_reference_unit_tests_on_gpu(gpu, tmp_path):
    exe = str(tmp_path / "test_shim")
    build(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240, cwd=str(tmp_path))
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert " 0 failed" in r.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(REF_APP), reason="oracle/_ref/laplacian_dropin is built only where the "
                                                        "reference sources are present (oracle/Makefile)")
def test_reference_laplacian_application_matches_device(gpu, tmp_path):
    """Laplacian.cpp at 2x2x2 blocks of 32^3: its host loop over the labs
    (u.loadLab per block, lab[i +- ix] etc.) writes laplacian.h5 (FP32);
    Device::laplacian on the same grid and initial field writes
    laplacian_device.h5.  The two files hold the same bits."""
    import cubismnova_amd as cnv
    r = subprocess.run([REF_APP, "2"], capture_output=True, text=True, timeout=240, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "compute:" in r.stdout
    exe = str(tmp_path / "test_shim")
    build(exe)
    r = subprocess.run([exe, "laplacian-device", "2"], capture_output=True, text=True, timeout=240,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = []
    for name in ("laplacian", "laplacian_device"):
        assert os.path.exists(tmp_path / (name + ".h5")) and os.path.exists(tmp_path / (name + ".xmf"))
        g = cnv.Cartesian((2, 2, 2), 32, dtype=np.float32)
        cnv.CartesianReadHDF(str(tmp_path / name), g)
        got.append(g.download(0))
    assert np.array_equal(got[0], got[1])
    assert float(np.max(np.abs(got[1]))) > 1.0      # a Laplacian of O(100) magnitude, not an empty file
