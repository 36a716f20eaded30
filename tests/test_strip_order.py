"""Host-only check of the depth-strip (design 2), angle-pair (design 3) and
staged-strip (design 4) table orderings, compiled from the library's
geometry.cpp and run on the CPU for the three projector methods, odd and
even widths and angle counts (one and several angle groups, a partial
last group)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "arec3d_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc absent")
    exe = str(tmp_path_factory.mktemp("order") / "strip_order_check")
    cmd = [HIPCC, "-O1", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "-I", os.path.join(ROOT, "include"),
           "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "strip_order_check.cpp"), os.path.join(CSRC, "geometry.cpp"),
           "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=600)
    return exe


@pytest.mark.parametrize("method", [1, 2, 3])
@pytest.mark.parametrize("nx,ny,na", [(48, 4, 13), (33, 3, 20), (65, 2, 7), (128, 2, 45)])
def test_strip_and_pair_orderings(checker, method, nx, ny, na):
    r = subprocess.run([checker, str(method), str(nx), str(ny), str(na)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
