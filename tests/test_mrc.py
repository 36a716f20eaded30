"""MRC files (arecImageIO.cpp:10-173): the native reader and writer against
the oracle's restatement, the oracle against the MRC2014 field offsets, and
both against the reference's own reader and writer when that is built.
Host code only (no GPU)."""
import os
import struct
import subprocess

import numpy as np
import pytest

import arec3d_amd as A
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "arec3d_amd")
MPI = "/opt/conda"

SHAPES = [(1, 1, 1), (3, 5, 7), (1, 64, 33), (4, 16, 16)]


def _stack(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) * 10).astype(np.float32)


def _file(tmp_path, mode, nz, ny, nx, payload):
    h = bytearray(O.mrc_header(nx, ny, nz, np.zeros(1, np.float32)))
    h[12:16] = struct.pack("<i", mode)
    p = tmp_path / f"m{mode}.mrc"
    p.write_bytes(bytes(h) + payload)
    return p


def test_oracle_header_offsets():
    # MRC2014 word positions: nx 1, mode 4, mx 8, cella 11, cellb 14, mapc 17,
    # dmin 20, dmax 21, dmean 22, ispg 23, extra 25-49, origin 50-52, map 53,
    # machst 54, rms 55, nlabl 56, labels 57-256.
    d = np.array([[[-2.5, 4.0], [1.0, 0.5]]], np.float32)
    h = O.mrc_header(2, 2, 1, d)
    assert len(h) == 1024
    word = lambda k, c="<i": struct.unpack_from(c, h, 4 * (k - 1))[0]
    assert (word(1), word(2), word(3), word(4)) == (2, 2, 1, 2)
    assert (word(8), word(9), word(10)) == (2, 2, 1)
    assert [word(k, "<f") for k in (11, 12, 13, 14, 15, 16)] == [1.0, 1.0, 1.0, 90.0, 90.0, 90.0]
    assert (word(17), word(18), word(19)) == (1, 2, 3)
    assert (word(20, "<f"), word(21, "<f"), word(22, "<f")) == (-2.5, 4.0, 0.75)
    assert word(25) == 32 and word(48) == 32 and word(49) == 0
    assert h[208:212] == b"MAP\0"
    assert word(55, "<f") == 1.0 and word(56) == 0
    assert h[224:] == bytes(800)


@pytest.mark.parametrize("shape", SHAPES)
def test_write_matches_oracle(tmp_path, shape):
    d = _stack(shape, sum(shape))
    p = tmp_path / "a.mrc"
    A.write_mrc(p, d)
    raw = p.read_bytes()
    assert len(raw) == 1024 + 4 * d.size
    want = bytearray(O.mrc_header(shape[2], shape[1], shape[0], d))
    want[212:216] = struct.pack("<i", 0x4444)  # the stamp the native writer adds
    assert raw[:1024] == bytes(want)
    assert raw[1024:] == d.tobytes()


@pytest.mark.parametrize("shape", SHAPES)
def test_round_trip(tmp_path, shape):
    d = _stack(shape, 7)
    p = tmp_path / "b.mrc"
    A.write_mrc(p, d)
    assert A.read_mrc_header(p) == (shape[2], shape[1], shape[0], 2)
    got = A.read_mrc(p)
    assert got.shape == d.shape and got.tobytes() == d.tobytes()
    assert O.mrc_read(p).tobytes() == d.tobytes()


def test_write_2d_is_one_section(tmp_path):
    d = _stack((6, 9), 1)
    A.write_mrc(tmp_path / "c.mrc", d)
    assert A.read_mrc_header(tmp_path / "c.mrc") == (9, 6, 1, 2)


@pytest.mark.parametrize("mode", [1, 6])
def test_read_16bit_as_signed(tmp_path, mode):
    v = np.array([0, 1, -1, 32767, -32768, 12345, -300, 2], np.int16).reshape(1, 2, 4)
    p = _file(tmp_path, mode, 1, 2, 4, v.astype("<i2").tobytes())
    want = v.astype(np.float32)
    assert A.read_mrc(p).tobytes() == want.tobytes()
    assert O.mrc_read(p).tobytes() == want.tobytes()
    # a mode-6 word above 32767 reads as negative, as the reference's short
    u = np.array([40000], np.uint16)
    p2 = _file(tmp_path, mode, 1, 1, 1, u.tobytes())
    assert A.read_mrc(p2)[0, 0, 0] == 40000 - 65536


@pytest.mark.parametrize("mode", [0, 3, 4, 12])
def test_other_modes_refused(tmp_path, mode):
    p = _file(tmp_path, mode, 1, 1, 2, bytes(16))
    with pytest.raises(A.ArecError) as e:
        A.read_mrc(p)
    assert e.value.rc == -1 and "mode" in str(e.value)
    with pytest.raises(ValueError):
        O.mrc_read(p)


def test_errors(tmp_path):
    with pytest.raises(A.ArecError):
        A.read_mrc_header(tmp_path / "missing.mrc")
    with pytest.raises(A.ArecError):
        A.read_mrc(tmp_path / "missing.mrc")
    (tmp_path / "short.mrc").write_bytes(bytes(100))
    with pytest.raises(A.ArecError):
        A.read_mrc_header(tmp_path / "short.mrc")
    p = _file(tmp_path, 2, 2, 2, 2, bytes(4 * 7))  # one value short
    with pytest.raises(A.ArecError):
        A.read_mrc(p)
    p = _file(tmp_path, 2, 2, 2, 2, bytes(4 * 8))
    buf = np.zeros(7, np.float32)
    assert A.lib.arec_mrc_read(os.fsencode(p), A._ptr(buf), 7) == -103  # AREC_ESIZE
    with pytest.raises(A.ArecError):
        A.write_mrc(tmp_path / "no_such_dir" / "x.mrc", np.zeros((1, 1, 1), np.float32))


@pytest.mark.skipif(not O.ref_available(), reason="reference build not present (needs /root/reference)")
@pytest.mark.parametrize("shape", SHAPES)
def test_reference_cross(tmp_path, shape):
    d = _stack(shape, 11)
    pr, pn = tmp_path / "ref.mrc", tmp_path / "native.mrc"
    assert O.ref_mrc_write(pr, d) == 0
    A.write_mrc(pn, d)
    r, n = pr.read_bytes(), pn.read_bytes()
    # every field the reference sets (it leaves machinestamp and labels unset)
    assert r[:212] == n[:212] and r[216:224] == n[216:224] and r[1024:] == n[1024:]
    assert A.read_mrc(pr).tobytes() == d.tobytes()
    st, back = O.ref_mrc_read(pn, d.size)
    assert st == 0 and back.tobytes() == d.tobytes()
    assert O.mrc_header(shape[2], shape[1], shape[0], d)[:212] == r[:212]


@pytest.mark.skipif(not O.ref_available(), reason="reference build not present (needs /root/reference)")
def test_reference_reads_16bit(tmp_path):
    v = np.array([-7, 300, -32768, 32767], np.int16).reshape(1, 1, 4)
    p = _file(tmp_path, 1, 1, 1, 4, v.tobytes())
    st, got = O.ref_mrc_read(p, 4)
    assert st == 0 and got.tobytes() == A.read_mrc(p).tobytes()


@pytest.mark.skipif(not os.path.exists(os.path.join(PKG, "libarec3d_dropin.so"))
                    or not os.path.exists(f"{MPI}/include/mpi.h"), reason="drop-in library or MPI headers absent")
def test_dropin_mrc(tmp_path):
    exe = str(tmp_path / "mrc_driver")
    cmd = ["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", f"{MPI}/include",
           "-DMPICH_SKIP_MPICXX", os.path.join(ROOT, "tests", "dropin", "mrc_driver.cpp"), "-o", exe,
           "-L", PKG, "-larec3d_dropin", "-larec3d_amd", f"-Wl,-rpath,{PKG}",
           f"{MPI}/lib/libmpi.so", f"-Wl,-rpath,/usr/lib/x86_64-linux-gnu:{MPI}/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
