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


def test_quad2dfit_host_entry_matches_oracle():
    """arec_quad2dfit (host code shared with the speak kernel's fit,
    align2dstack.cpp:1125-1166) equals the oracle's fit bit for bit, peak value
    included; a singular fit leaves the peak value unset, as the reference."""
    import numpy as np

    import arec3d_amd as A
    import pyoracle as O

    rng = np.random.default_rng(4)
    for _ in range(200):
        z = rng.standard_normal(9).astype(np.float32)
        assert A.quad2dfit(z) == O.quad2dfit(z)
    z = np.array([2.0 - 0.5 * (u - 0.3) ** 2 - 0.8 * (v + 0.4) ** 2 for v in (-1, 0, 1) for u in (-1, 0, 1)],
                 np.float32)
    xs, ys, pv = A.quad2dfit(z)
    assert abs(xs - 0.3) < 1e-5 and abs(ys + 0.4) < 1e-5 and abs(pv - 2.0) < 1e-5
    xs, ys, pv = A.quad2dfit(np.ones(9, np.float32))
    assert xs == 0.0 and ys == 0.0 and np.isnan(pv)
