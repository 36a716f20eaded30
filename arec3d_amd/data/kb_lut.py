"""Kaiser-Bessel footprint table (1000 float32 samples on [0, 3] pixels).

The numbers are those of the KB projector pair, AREC3D/AREC3D/arecproject.cpp:192-198,
kept as float32 bit patterns in kb_lut.txt.  ``write_inc`` turns them into a C
initialiser list for the builds that compile the table in.
"""
import os
import struct
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def load_bits(path=os.path.join(HERE, "kb_lut.txt")):
    bits = []
    for line in open(path):
        line = line.strip()
        if line and not line.startswith("#"):
            bits.append(int(line.split()[0], 16))
    if len(bits) != 1000:
        raise ValueError(f"{path}: expected 1000 entries, found {len(bits)}")
    return bits


def load_values():
    return [struct.unpack("<f", struct.pack("<I", b))[0] for b in load_bits()]


def write_inc(out):
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    with open(out, "w") as f:
        for i, b in enumerate(load_bits()):
            f.write(f"0x{b:08x}u,{chr(10) if i % 8 == 7 else ' '}")
        f.write("\n")


if __name__ == "__main__":
    write_inc(sys.argv[1])
