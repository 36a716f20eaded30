"""Timing calibration of the CPU baseline: the oracle against the reference.

TEST INFRASTRUCTURE ONLY (a measurement recipe, run by hand in a container
that holds the reference sources).  Both are built with the reference's
release flags (AREC3D/Makefile:19-20, -O3 -DNDEBUG -march=native: ``make -C
oracle fast ref-fast``) and run single-threaded on the CPU reference case:
a 128 x 64 detector, 180 angles, the SQ pair (cyl_sirt_SQ,
arecsirt.cpp:193-373), 50 SIRT iterations from zero.  Prints the median wall
time of each over several repetitions and their ratio; bench.py's CPU
baseline times the oracle, so the ratio says how far that stands from the
reference's own speed.

    python oracle/calibrate.py [repetitions]
"""
import ctypes as C
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import pyoracle as O  # noqa: E402


def main(reps=5):
    subprocess.run(["make", "-s", "-C", HERE, "fast", "ref-fast"], check=True)
    nx, ny, na, iters = 128, 64, 180, 50
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    rng = np.random.default_rng(1)
    nr = len(O.cyl_geometry(0.5 * (nx - 1)))
    b = O.project("sq", rng.random((nr, ny)).astype(np.float32), ang, nx, ny)
    lam = np.float32(1.0 / (na * nx))
    R = O.ref(os.path.join(HERE, "_ref", "fast", "libarecref.so"))
    prev = O.use_build(os.path.join(HERE, "_build", "fast", "liboracle.so"))

    def ref_run():
        x = np.zeros(((nr + 8 * nx) * ny,), np.float32)
        nnz = C.c_int(0)
        st = R.ref_solve(1, 3, b, nx, ny, na, ang, lam, iters, 0.0, 0, x, C.byref(nnz))
        assert st == 0
        return x[: nnz.value].reshape(-1, ny)

    def orc_run():
        st, x, _, _ = O.sirt("sq", b, ang, lam, iters, 0.0)
        assert st == 0
        return x

    xr, xo = ref_run(), orc_run()
    dev = float(np.linalg.norm(xr - xo) / np.linalg.norm(xr))
    tr, to = [], []
    for _ in range(reps):
        for fn, acc in ((ref_run, tr), (orc_run, to)):
            t0 = time.perf_counter()
            fn()
            acc.append(time.perf_counter() - t0)
    O.restore_build(prev)
    mr, mo = float(np.median(tr)), float(np.median(to))
    vu = 2.0 * nr * ny * na * iters
    print(f"CPU reference case 128x64/180, 50 SIRT iterations (SQ), 1 thread, {reps} repetitions")
    print(f"reference (-O3 -march=native): {mr:.3f} s ({vu / mr / 1e6:.1f} MVU/s)")
    print(f"oracle    (-O3 -march=native): {mo:.3f} s ({vu / mo / 1e6:.1f} MVU/s)")
    print(f"oracle / reference time: {mo / mr:.3f}   volume difference: {dev:.2e}")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 5)
