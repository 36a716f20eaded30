"""The reference-signature C++ drop-in (libarec3d_dropin.so): a caller
written against the reference's interface compiles and links (CPU), and on
the GPU reproduces the reference's unit-test answers."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "arec3d_amd")
DROPIN = os.path.join(PKG, "libarec3d_dropin.so")
MPI = "/opt/conda"
EXE = os.path.join(ROOT, "tests", "dropin", "dropin_driver")


def build_driver():
    cmd = ["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", f"{MPI}/include",
           "-DMPICH_SKIP_MPICXX", os.path.join(ROOT, "tests", "dropin", "dropin_driver.cpp"), "-o", EXE,
           "-L", PKG, "-larec3d_dropin", "-larec3d_amd", f"-Wl,-rpath,{PKG}",
           f"{MPI}/lib/libmpi.so", f"-Wl,-rpath,/usr/lib/x86_64-linux-gnu:{MPI}/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return EXE


@pytest.mark.skipif(not os.path.exists(DROPIN) or not os.path.exists(f"{MPI}/include/mpi.h"),
                    reason="drop-in library or MPI headers absent")
def test_dropin_driver_links():
    assert os.path.exists(build_driver())
    out = subprocess.run(["nm", "-DC", DROPIN], capture_output=True, text=True).stdout
    for sym in ["arecProject2D_SQ(Image, float*, int, Image*)", "cyl_sirt_SQ(int, Image, float*, Image*, float, int, float)",
                "cyl_cgls(int, Image, float*, Image*, float, int, float)", "arecCCImages(int, Image, Image, float*, float*)",
                "arecRotCCImages(int, Image, Image, int, float*)", "arecRotateImages_skew_safe(Image*, Image*, float*)",
                "arecRotateImages_skew(Image*, float*)", "arecRotateImages(Image*, float*)",
                "arecReadImage(char*, Image*)", "arecWriteImage(char*, Image)",
                "arecReadImageHeader(char*, int*, int*, int*)", "periodicDecomposition(Image)",
                "ccorr2d(int, int, float*, float*, float*)", "fshift2d(int, int, float*, float*, float, float)",
                "quad2dfit(float*, float*, float*, float*)", "speak(int, int, float*, float*, float*)",
                "ImageCB2Cyl(Image, Image, int, int, double, int)", "ImageCyl2CB(Image, Image, int, int, int)"]:
        assert sym in out, sym


@pytest.mark.skipif(not os.path.exists(DROPIN) or not os.path.exists(f"{MPI}/include/mpi.h"),
                    reason="drop-in library or MPI headers absent")
def test_dropin_driver_host_checks():
    """The checks that need no GPU: the cylinder, ImageCyl2CB / ImageCB2Cyl
    with the reference's semantics (AREC3D_test.cpp:326-363; -2 for a
    half-integer radius) and quad2dfit."""
    exe = build_driver()
    r = subprocess.run([exe, "--host"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0 and "all checks passed" in r.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(DROPIN), reason="drop-in library absent")
def test_dropin_driver_runs():
    exe = EXE if os.path.exists(EXE) else build_driver()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0
