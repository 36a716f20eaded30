"""Build libarec3d_amd.so (HIP, gfx950) in-tree.

    python -m arec3d_amd.build          # or arec3d_amd.build.build()

Every source is compiled with hipcc for gfx950 only.  -ffp-contract=off keeps
the geometry and solver arithmetic as written (separate roundings, like the
reference); the hot loops use explicit fmaf.  Packed FP32 instructions are
disabled (-packed-fp32-ops).
"""
import concurrent.futures as cf
import os
import struct
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libarec3d_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["geometry.cpp", "capi.cpp", "projector.hip", "solver.hip", "align.hip", "phantom.hip", "unet.hip",
           "matching.cpp", "mrc.cpp", "rotate.hip"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _write_lut_inc():
    sys.path.insert(0, os.path.join(HERE, "data"))
    try:
        import kb_lut
    finally:
        sys.path.pop(0)
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, "kb_lut_f.inc")
    lines = []
    for b in kb_lut.load_bits():
        v = struct.unpack("<f", struct.pack("<I", b))[0]
        lines.append(f"{v.hex()}f,")
    text = "\n".join(lines) + "\n"
    if not os.path.exists(out) or open(out).read() != text:
        with open(out, "w") as f:
            f.write(text)


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", BUILD]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose=False):
    """Compile every source (always, as the build check requires) and link."""
    _write_lut_inc()
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_dropin(verbose)
    return LIB


MPI_ROOT = os.environ.get("MPI_ROOT", "/opt/conda")
DROPIN = os.path.join(HERE, "libarec3d_dropin.so")


def build_dropin(verbose=False):
    """libarec3d_dropin.so: the reference's C++ symbols over the C ABI (host
    code, g++), built when an MPI installation provides mpi.h."""
    if not os.path.exists(os.path.join(MPI_ROOT, "include", "mpi.h")):
        if verbose:
            print("mpi.h not found: skipping libarec3d_dropin.so")
        return None
    tmp = DROPIN + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-Wall",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(MPI_ROOT, "include"),
           "-DMPICH_SKIP_MPICXX", "-DOMPI_SKIP_MPICXX",
           os.path.join(CSRC, "dropin.cpp"), "-o", tmp,
           "-L", HERE, "-larec3d_amd", "-Wl,-rpath,$ORIGIN",
           # system libraries first: the MPI directory also holds an older libstdc++
           os.path.join(MPI_ROOT, "lib", "libmpi.so"),
           "-Wl,-rpath,/usr/lib/x86_64-linux-gnu:" + os.path.join(MPI_ROOT, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"drop-in build failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, DROPIN)
    if verbose:
        print("built", DROPIN)
    return DROPIN


if __name__ == "__main__":
    build(verbose=True)
