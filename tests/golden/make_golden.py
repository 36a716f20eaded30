"""Write golden vectors from the reference's own projector and solver code.

Run in a container that holds the reference sources: it builds
oracle/_ref/libarecref.so (the reference's arecproject.cpp, arecsirt*.cpp and
areccgls*.cpp compiled as they are, see oracle/Makefile) and stores small
input/output pairs in tests/golden/ref_vectors.npz.  The GPU tests compare the
HIP path against these vectors; the CPU tests check the oracle against them.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as O  # noqa: E402


def main():
    O.build()
    assert O.ref_available(), "reference build failed"
    rng = np.random.default_rng(20240611)
    out = {}
    for tag, (nx, ny, na) in {"odd": (33, 6, 24), "even": (32, 5, 20)}.items():
        ang = (np.arange(na) * np.pi / na).astype(np.float32)
        nrays = len(O.ref_cyl_geometry(0.5 * (nx - 1)))
        x = rng.random((nrays, ny)).astype(np.float32)
        x[-2:] = 0.0  # the linear pair's last two columns are undefined in the reference
        p = rng.random((na, ny, nx)).astype(np.float32)
        out[f"{tag}_angles"] = ang
        out[f"{tag}_x"] = x
        out[f"{tag}_p"] = p
        out[f"{tag}_cord"] = O.ref_cyl_geometry(0.5 * (nx - 1)).astype(np.int32)
        for m in ("sq", "kb", "linear"):
            out[f"{tag}_{m}_fwd"] = O.ref_project(m, x, ang, nx, ny)
            out[f"{tag}_{m}_bwd"] = O.ref_backproject(m, p, ang, nx, ny)
        # solvers on a noisy SQ sinogram of a smooth volume
        b = O.ref_project("sq", x, ang, nx, ny)
        b = (b + 0.01 * b.max() * rng.standard_normal(b.shape)).astype(np.float32)
        out[f"{tag}_b"] = b
        for m in ("sq", "kb"):
            for lam in (2e-3, 1.0):
                st, xr = O.ref_solve("sirt", m, b, ang, lam, 15, 1e-4)
                out[f"{tag}_{m}_sirt_{lam:g}"] = xr
                out[f"{tag}_{m}_sirt_{lam:g}_status"] = np.int32(st)
            st, xr = O.ref_solve("cgls", m, b, ang, 0.0, 8, 1e-6)
            out[f"{tag}_{m}_cgls"] = xr
            out[f"{tag}_{m}_cgls_status"] = np.int32(st)
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_vectors.npz"), sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main()
