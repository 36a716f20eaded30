"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls here)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "arec3d_amd", "libarec3d_amd.so")


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(arec_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared(os.path.join(ROOT, "include", "arec3d_amd.h"))
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import arec3d_amd as A

    names = declared(os.path.join(ROOT, "include", "arec3d_amd.h"))
    assert set(names) <= set(A.EXPORTED)


def test_host_only_entry_points():
    import arec3d_amd as A

    assert A.cyl_nrays(63.5) == 12900
    assert A.cyl_cord(1.0).tolist() == [[1, 0], [0, 1], [1, 1], [2, 1], [1, 2]]
    assert A.lib.arec_version().startswith(b"arec3d_amd")


def test_no_gpu_reports_error_not_crash():
    import arec3d_amd as A

    try:
        import torch  # noqa: F401

        have_gpu = torch.cuda.is_available()
    except Exception:
        have_gpu = False
    if have_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(A.ArecError):
        A.Context(0)
