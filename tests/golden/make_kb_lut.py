"""Extract the Kaiser-Bessel footprint table of the KB projector pair as data.

The KB projector (AREC3D/AREC3D/arecproject.cpp:192-198) interpolates a fixed
1000-entry table sampled on [0, 3] pixels.  The table is empirical (rounded
to 5 significant digits, with round-off noise in the tail), so it cannot be
regenerated from a formula; the projector needs the same numbers.  This script
reads the decimal literals from the reference source once and writes them as
float32 bit patterns to ``arec3d_amd/data/kb_lut.txt``.  It is a data recipe,
run by hand in a container that holds the reference; the committed
``kb_lut.txt`` is what the build uses.
"""
import os
import struct
import sys

REF = "/root/reference/AREC3D/AREC3D/arecproject.cpp"


ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main(out=os.path.join(ROOT, "arec3d_amd", "data", "kb_lut.txt")):
    src = open(REF).read()
    start = src.index("float LUT[LUT_SIZE] = {") + len("float LUT[LUT_SIZE] = {")
    stop = src.index("};", start)
    vals = [v.strip() for v in src[start:stop].replace("\n", " ").split(",") if v.strip()]
    if len(vals) != 1000:
        sys.exit(f"expected 1000 entries, found {len(vals)}")
    with open(out, "w") as f:
        f.write("# KB footprint table: 1000 float32 samples on [0,3] (bits, value)\n")
        for v in vals:
            x = struct.unpack("<f", struct.pack("<f", float(v)))[0]
            bits = struct.unpack("<I", struct.pack("<f", x))[0]
            f.write(f"0x{bits:08x} {x!r}\n")


if __name__ == "__main__":
    main()
