"""ctypes bindings to the CPU oracle (liboracle.so) and, when it was built, to
the reference's own translation units (_ref/libarecref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py.  The product never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libarecref.so")

METHODS = {"linear": 1, "kb": 2, "sq": 3}

_F = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_D = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class Hist(C.Structure):
    _fields_ = [("iter", C.c_int), ("event", C.c_int), ("relnrm", C.c_double), ("rnorm", C.c_double)]


_lib = None
_ref = None


def build():
    """Compile liboracle.so (and _ref when the reference sources are present)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference/AREC3D/AREC3D"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _bind(path):
    L = C.CDLL(path)
    L.orc_cyl_geometry.argtypes = [C.c_double, C.c_int, C.c_void_p]
    L.orc_cyl_geometry.restype = C.c_int
    for sfx, arr in (("f", _F), ("d", _D), ("g", _F)):
        getattr(L, f"orc_project_{sfx}").argtypes = [C.c_int, _I, C.c_int, C.c_double, C.c_int, C.c_int, arr, _F, C.c_int, arr]
        getattr(L, f"orc_backproject_{sfx}").argtypes = [C.c_int, _I, C.c_int, C.c_double, C.c_int, C.c_int, arr, _F, C.c_int, arr]
    L.orc_sirt_f.argtypes = [C.c_int, _F, C.c_int, C.c_int, C.c_int, _F, C.c_float, C.c_int, C.c_float, _F,
                             C.POINTER(Hist), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_float)]
    L.orc_sirt_d.argtypes = [C.c_int, _D, C.c_int, C.c_int, C.c_int, _F, C.c_double, C.c_int, C.c_double, _D,
                             C.POINTER(Hist), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.orc_cgls_f.argtypes = [C.c_int, _F, C.c_int, C.c_int, C.c_int, _F, C.c_int, C.c_float, C.c_int, _F,
                             C.POINTER(Hist), C.c_int, C.POINTER(C.c_int)]
    L.orc_cgls_d.argtypes = [C.c_int, _D, C.c_int, C.c_int, C.c_int, _F, C.c_int, C.c_double, C.c_int, _D,
                             C.POINTER(Hist), C.c_int, C.POINTER(C.c_int)]
    L.orc_sirt_g.argtypes = L.orc_sirt_f.argtypes
    L.orc_cgls_g.argtypes = L.orc_cgls_f.argtypes
    L.orc_cgls_stat_f.argtypes = L.orc_cgls_f.argtypes
    L.orc_cgls_stat_d.argtypes = L.orc_cgls_d.argtypes
    L.orc_init_weights_f.argtypes = [_F, C.c_int, C.c_int, C.c_int, _F]
    L.orc_ccorr2d.argtypes = [C.c_int, C.c_int, _F, _F, _D]
    L.orc_quad2dfit.argtypes = [_F, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_speak.argtypes = [C.c_int, C.c_int, _F, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_fshift2d.argtypes = [C.c_int, C.c_int, _F, _F, C.c_float, C.c_float]
    L.orc_cc_images.argtypes = [C.c_int, C.c_int, C.c_int, _F, _F, _F, _F]
    L.orc_shift_images.argtypes = [C.c_int, C.c_int, C.c_int, _F, _F, _F]
    L.orc_soft_noneg.argtypes = [_F, C.c_long]
    L.orc_take_log.argtypes = [_F, C.c_long]
    L.orc_periodic_smooth.argtypes = [C.c_int, C.c_int, _F, _F]
    L.orc_periodic_decomposition.argtypes = [C.c_int, C.c_int, C.c_int, _F]
    L.orc_radial_profile.argtypes = [C.c_int, C.c_int, _F, C.c_int, C.c_int, _F]
    L.orc_ccorr1d.argtypes = [C.c_int, _F, _F, _D]
    L.orc_findpeak1d.argtypes = [C.c_int, _F]
    L.orc_findpeak1d.restype = C.c_float
    L.orc_rot_cc_images.argtypes = [C.c_int, C.c_int, C.c_int, _F, _F, C.c_int, _F]
    L.orc_shear_x.argtypes = [_F, C.c_float, C.c_int, C.c_int]
    L.orc_shear_y.argtypes = [_F, C.c_float, C.c_int, C.c_int]
    L.orc_rotate2d.argtypes = [C.c_int, C.c_int, _F, _F, C.c_float]
    L.orc_rotate_skew.argtypes = [C.c_int, C.c_int, C.c_int, _F, _F]
    L.orc_tukey_ncc.argtypes = [_F, _F, C.c_int, C.c_int]
    L.orc_tukey_ncc.restype = C.c_double
    L.orc_rotate_skew_safe.argtypes = [C.c_int, C.c_int, C.c_int, _F, _F, _F, C.POINTER(C.c_double)]
    L.orc_rotate_skew_safe.restype = C.c_int
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(LIB_PATH)
    return _lib


def use_build(path):
    """Route every oracle function through another build of the same source
    (the -O3 -march=native timing build of ``make fast``); returns the
    binding it replaces.  Parity checks use the default build."""
    global _lib
    prev = lib()
    _lib = _bind(path)
    return prev


def restore_build(prev):
    global _lib
    _lib = prev


def ref_available():
    return os.path.exists(REF_PATH)


def ref(path=None):
    """The reference build (default) or, with a path, another build of the
    same reference sources (``make ref-fast``), bound afresh."""
    global _ref
    if path is not None or _ref is None:
        R = C.CDLL(path or REF_PATH)
        R.ref_cyl_geometry.argtypes = [C.c_double, C.c_int, C.POINTER(C.c_int), C.c_void_p]
        R.ref_project.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, _F, _F, C.c_int, _F]
        R.ref_backproject.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, _F, _F, C.c_int, _F]
        R.ref_solve.argtypes = [C.c_int, C.c_int, _F, C.c_int, C.c_int, C.c_int, _F, C.c_float, C.c_int,
                                C.c_float, C.c_int, _F, C.POINTER(C.c_int)]
        if path is not None:
            return R
        _ref = R
    return _ref


def _m(method):
    return METHODS[method] if isinstance(method, str) else int(method)


def cyl_geometry(radius, height=1):
    L = lib()
    n = L.orc_cyl_geometry(float(radius), int(height), None)
    cord = np.zeros(2 * n, np.int32)
    L.orc_cyl_geometry(float(radius), int(height), cord.ctypes.data)
    return cord.reshape(n, 2)


def _sfx(dtype, arith):
    """Instantiation suffix: f (FP32, the reference's arithmetic), d (FP64) or
    g (FP32 with the GPU library's per-tap arithmetic: FP32 weights, FMA)."""
    if arith == "device":
        if dtype != np.float32:
            raise ValueError("device arithmetic is FP32")
        return "g"
    if arith != "reference":
        raise ValueError(f"unknown arithmetic {arith!r}")
    return "f" if dtype == np.float32 else "d"


def project(method, x, angles, nx, ny, dtype=np.float32, arith="reference"):
    """x: [nrays, ny] -> p: [na, ny, nx]."""
    radius = 0.5 * (nx - 1)
    cord = np.ascontiguousarray(cyl_geometry(radius).reshape(-1))
    angles = np.ascontiguousarray(angles, np.float32)
    x = np.ascontiguousarray(x, dtype)
    p = np.zeros((len(angles), ny, nx), dtype)
    f = getattr(lib(), "orc_project_" + _sfx(dtype, arith))
    st = f(_m(method), cord, len(cord) // 2, radius, nx, ny, x, angles, len(angles), p)
    if st:
        raise RuntimeError(f"orc_project status {st}")
    return p


def backproject(method, p, angles, nx, ny, dtype=np.float32, arith="reference"):
    radius = 0.5 * (nx - 1)
    cord = np.ascontiguousarray(cyl_geometry(radius).reshape(-1))
    angles = np.ascontiguousarray(angles, np.float32)
    p = np.ascontiguousarray(p, dtype)
    x = np.zeros((len(cord) // 2, ny), dtype)
    f = getattr(lib(), "orc_backproject_" + _sfx(dtype, arith))
    st = f(_m(method), cord, len(cord) // 2, radius, nx, ny, p, angles, len(angles), x)
    if st:
        raise RuntimeError(f"orc_backproject status {st}")
    return x


def _hist(h, n):
    return [(h[i].iter, h[i].event, h[i].relnrm, h[i].rnorm) for i in range(min(n, len(h)))]


def sirt(method, b, angles, lam, maxit, tol, dtype=np.float32, arith="reference"):
    """Returns (status, x [nrays, ny], history, final lam)."""
    na, ny, nx = b.shape
    _sfx(dtype, arith)
    nrays = len(cyl_geometry(0.5 * (nx - 1)))
    x = np.zeros((nrays, ny), dtype)
    H = (Hist * (maxit + 64))()
    nh = C.c_int(0)
    if dtype == np.float32:
        lo = C.c_float(0)
        st = getattr(lib(), "orc_sirt_" + _sfx(dtype, arith))(_m(method), np.ascontiguousarray(b, np.float32), nx, ny, na,
                              np.ascontiguousarray(angles, np.float32), lam, maxit, tol, x, H, len(H),
                              C.byref(nh), C.byref(lo))
    else:
        lo = C.c_double(0)
        st = lib().orc_sirt_d(_m(method), np.ascontiguousarray(b, np.float64), nx, ny, na,
                              np.ascontiguousarray(angles, np.float32), lam, maxit, tol, x, H, len(H),
                              C.byref(nh), C.byref(lo))
    return st, x, _hist(H, nh.value), lo.value


def cgls(method, b, angles, maxit, tol, x0=None, dtype=np.float32, weighted=False, arith="reference"):
    na, ny, nx = b.shape
    if arith != "reference" and weighted:
        raise ValueError("the weighted solver has the reference's arithmetic only")
    _sfx(dtype, arith)
    radius = (nx // 2) if (method in (2, "kb") or weighted) else 0.5 * (nx - 1)
    nrays = len(cyl_geometry(radius))
    x = np.zeros((nrays, ny), dtype) if x0 is None else np.array(x0, dtype, copy=True)
    H = (Hist * (maxit + 4))()
    nh = C.c_int(0)
    if weighted and dtype == np.float32:
        f, bb = lib().orc_cgls_stat_f, np.ascontiguousarray(b, np.float32)
    elif weighted:
        f, bb = lib().orc_cgls_stat_d, np.ascontiguousarray(b, np.float64)
    elif dtype == np.float32:
        f, bb = getattr(lib(), "orc_cgls_" + _sfx(dtype, arith)), np.ascontiguousarray(b, np.float32)
    else:
        f, bb = lib().orc_cgls_d, np.ascontiguousarray(b, np.float64)
    st = f(_m(method), bb, nx, ny, na, np.ascontiguousarray(angles, np.float32), maxit, tol,
           0 if x0 is None else 1, x, H, len(H), C.byref(nh))
    return st, x, _hist(H, nh.value)


def cyl_to_cube(x, radius):
    """The cylinder x [nrays, ny] scattered into its bounding cube, zero
    outside: ImageCyl2CB's scatter (arecImage.cpp:163-232, cube(ix, iy, iz)
    = cylval(j*ny + iy) over the cord pairs) with the detector rows as the
    slowest axis, [ny][iz][ix], the layout the device conversion writes."""
    cord = cyl_geometry(radius)
    n = int(round(2 * radius + 1))
    cube = np.zeros((x.shape[1], n, n), x.dtype)
    cube[:, cord[:, 1], cord[:, 0]] = x.T
    return cube


def cube_to_cyl(cube, radius):
    """The cylinder's columns gathered from a [ny][iz][ix] cube, [nrays, ny]
    (ImageCB2Cyl's gather over the allocator's cord, arecImage.cpp:88-161)."""
    cord = cyl_geometry(radius)
    return np.ascontiguousarray(cube[:, cord[:, 1], cord[:, 0]].T)


def init_weights(b):
    na, ny, nx = b.shape
    w = np.zeros_like(b, dtype=np.float32)
    lib().orc_init_weights_f(np.ascontiguousarray(b, np.float32), nx, ny, na, w)
    return w


def ccorr2d(x, y):
    ny, nx = x.shape
    c = np.zeros((ny, nx), np.float64)
    lib().orc_ccorr2d(nx, ny, np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32), c)
    return c


def quad2dfit(z3x3):
    xs, ys, pv = C.c_float(0), C.c_float(0), C.c_float(0)
    lib().orc_quad2dfit(np.ascontiguousarray(z3x3, np.float32).reshape(9), C.byref(xs), C.byref(ys), C.byref(pv))
    return xs.value, ys.value, pv.value


def speak(data):
    ny, nx = data.shape
    xt, yt = C.c_float(0), C.c_float(0)
    lib().orc_speak(nx, ny, np.ascontiguousarray(data, np.float32), C.byref(xt), C.byref(yt))
    return xt.value, yt.value


def fshift2d(img, sx, sy):
    ny, nx = img.shape
    out = np.zeros((ny, nx), np.float32)
    lib().orc_fshift2d(nx, ny, np.ascontiguousarray(img, np.float32), out, sx, sy)
    return out


def cc_images(proj, meas):
    nz, ny, nx = proj.shape
    sx = np.zeros(nz, np.float32)
    sy = np.zeros(nz, np.float32)
    lib().orc_cc_images(nx, ny, nz, np.ascontiguousarray(proj, np.float32), np.ascontiguousarray(meas, np.float32), sx, sy)
    return sx, sy


def shift_images(stack, sx, sy):
    nz, ny, nx = stack.shape
    out = np.array(stack, np.float32, copy=True)
    lib().orc_shift_images(nx, ny, nz, out, np.ascontiguousarray(sx, np.float32), np.ascontiguousarray(sy, np.float32))
    return out


def centroid_x_shifts(method, views, angles, nx, ny):
    """x shift of every view from its centroid, restated in FP64 numpy (no
    counterpart in the reference; the device entry is arec_align_centroid):
    d = view centroid along the detector - rotation axis (the centroid of
    the projected uniform cylinder), minus its least-squares fit by
    a cos(t) + b sin(t).  views: [na, ny, nx]."""
    angles = np.asarray(angles, np.float32)
    cord = cyl_geometry(0.5 * (nx - 1))
    axis = project(method, np.ones((len(cord), ny), np.float64), angles, nx, ny, dtype=np.float64)
    u = np.arange(nx, dtype=np.float64)

    def centroid(v):
        v = np.asarray(v, np.float64)
        return (v.sum(axis=1) * u).sum(axis=1) / v.sum(axis=(1, 2))

    d = centroid(views) - centroid(axis)
    t = angles.astype(np.float64)
    M = np.stack([np.cos(t), np.sin(t)], 1)
    return d - M @ np.linalg.lstsq(M, d, rcond=None)[0]


def take_log(x):
    """arecImageTakeLog: x > 0 -> -log(x), else 0 (FP32)."""
    out = np.array(x, np.float32, copy=True)
    lib().orc_take_log(out.reshape(-1), out.size)
    return out


def periodic_smooth(u):
    """The smooth part of Moisan's decomposition of one view [ny, nx]."""
    ny, nx = u.shape
    s = np.zeros((ny, nx), np.float32)
    lib().orc_periodic_smooth(nx, ny, np.ascontiguousarray(u, np.float32), s)
    return s


def periodic_decomposition(stack):
    """periodicDecomposition of a view stack [nz, ny, nx] (a copy)."""
    nz, ny, nx = stack.shape
    out = np.array(stack, np.float32, copy=True)
    lib().orc_periodic_decomposition(nx, ny, nz, out.reshape(-1))
    return out


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def rot_r2(nx, ny):
    """The profile radius of the driver's rotational step (AREC3D.cpp:237-238
    with radius (nx-1)/2 and height ny): min(nx, ny - 10) // 2."""
    return min(nx, ny - 10) // 2


def radial_profile(img, length, R):
    """getRadialProfile of one view [ny, nx]."""
    ny, nx = img.shape
    out = np.zeros(length, np.float32)
    lib().orc_radial_profile(nx, ny, _f32(img), length, R, out)
    return out


def ccorr1d(x, y):
    """ccorr1d: c[k] = n sum_m x[m] y[(m+k) mod n] (FP64)."""
    c = np.zeros(len(x), np.float64)
    lib().orc_ccorr1d(len(x), _f32(x), _f32(y), c)
    return c


def findpeak1d(c):
    return float(lib().orc_findpeak1d(len(c), _f32(c)))


def rot_cc_images(img1, img2, r2):
    """arecRotCCImages: per-view rotation angles (degrees) of img2 against img1."""
    nz, ny, nx = img1.shape
    ang = np.zeros(nz, np.float32)
    lib().orc_rot_cc_images(nx, ny, nz, _f32(img1), _f32(img2), r2, ang)
    return ang


def shear_x(img, shear):
    out = np.array(img, np.float32, copy=True)
    ny, nx = out.shape
    lib().orc_shear_x(out.reshape(-1), shear, nx, ny)
    return out


def shear_y(img, shear):
    out = np.array(img, np.float32, copy=True)
    ny, nx = out.shape
    lib().orc_shear_y(out.reshape(-1), shear, nx, ny)
    return out


def rotate2d(img, deg):
    ny, nx = img.shape
    out = np.zeros((ny, nx), np.float32)
    lib().orc_rotate2d(nx, ny, _f32(img), out, deg)
    return out


def rotate_skew(stack, angles):
    """arecRotateImages_skew on a copy: (rotated stack, updated angles)."""
    out = np.array(stack, np.float32, copy=True)
    ang = np.array(angles, np.float32, copy=True)
    nz, ny, nx = out.shape
    lib().orc_rotate_skew(nx, ny, nz, out.reshape(-1), ang)
    return out, ang


def tukey_ncc(a, b):
    ny, nx = a.shape
    return float(lib().orc_tukey_ncc(_f32(a), _f32(b), nx, ny))


def rotate_skew_safe(stack, ref, angles):
    """arecRotateImages_skew_safe on a copy: (stack, angles, kept, mean cc)."""
    out = np.array(stack, np.float32, copy=True)
    ang = np.array(angles, np.float32, copy=True)
    nz, ny, nx = out.shape
    m = C.c_double(0.0)
    kept = lib().orc_rotate_skew_safe(nx, ny, nz, out.reshape(-1), _f32(ref), ang, C.byref(m))
    return out, ang, kept, m.value


# MRC header: the reference's MRCheader (AREC3D/AREC3D/arecImageIO.h:5-63) as
# (name, struct code) in file order; 1024 bytes, host (little-endian) order.
MRC_FIELDS = [("nx", "i"), ("ny", "i"), ("nz", "i"), ("mode", "i"), ("nxstart", "i"), ("nystart", "i"),
              ("nzstart", "i"), ("mx", "i"), ("my", "i"), ("mz", "i"), ("xlen", "f"), ("ylen", "f"),
              ("zlen", "f"), ("alpha", "f"), ("beta", "f"), ("gamma", "f"), ("mapc", "i"), ("mapr", "i"),
              ("maps", "i"), ("amin", "f"), ("amax", "f"), ("amean", "f"), ("ispg", "i"), ("nsymbt", "i"),
              ("user", "25i"), ("xorigin", "f"), ("yorigin", "f"), ("zorigin", "f"), ("map", "4s"),
              ("machinestamp", "i"), ("rms", "f"), ("nlabels", "i"), ("labels", "800s")]
MRC_FORMAT = "<" + "".join(c for _, c in MRC_FIELDS)


def mrc_header(nx, ny, nz, data):
    """The header arecWriteImage writes (arecImageIO.cpp:84-147): mode 2, cell
    1x1x1 at 90 degrees, axes 1/2/3, min/max/mean of the data accumulated in
    double with the mean divided by the float count, user[] spaces with a zero
    last word, "MAP", rms 1, no labels.  The reference leaves machinestamp and
    labels unset; here they are zero."""
    import struct

    d = np.asarray(data, np.float32).reshape(-1)
    amin, amax, mean = 1.0e20, -1.0e20, 0.0
    for v in d.astype(np.float64):
        amax = v if v > amax else amax
        amin = v if v < amin else amin
        mean += v
    mean = mean / float(np.float32(nx * ny * nz))
    user = [ord(" ")] * 24 + [0]
    return struct.pack(MRC_FORMAT, nx, ny, nz, 2, 0, 0, 0, nx, ny, nz, 1.0, 1.0, 1.0, 90.0, 90.0, 90.0, 1, 2, 3,
                       amin, amax, mean, 0, 0, *user, 0.0, 0.0, 0.0, b"MAP\0", 0, 1.0, 0, b"")


def mrc_parse_header(raw):
    """The 1024-byte header as a dict of the fields above."""
    import struct

    vals = list(struct.unpack(MRC_FORMAT, raw[:1024]))
    out = {}
    for name, code in MRC_FIELDS:
        if code == "25i":
            out[name], vals = vals[:25], vals[25:]
        else:
            out[name], vals = vals[0], vals[1:]
    return out


def mrc_read(path):
    """arecReadImage (arecImageIO.cpp:33-82): modes 1 and 6 as signed 16-bit,
    mode 2 as float32, any other mode an error; [nz, ny, nx] float32."""
    with open(path, "rb") as f:
        raw = f.read()
    h = mrc_parse_header(raw)
    n = h["nx"] * h["ny"] * h["nz"]
    if h["mode"] in (1, 6):
        v = np.frombuffer(raw, "<i2", count=n, offset=1024).astype(np.float32)
    elif h["mode"] == 2:
        v = np.frombuffer(raw, "<f4", count=n, offset=1024).copy()
    else:
        raise ValueError(f"invalid mrc mode {h['mode']}")
    return v.reshape(h["nz"], h["ny"], h["nx"])


def soft_noneg(x):
    out = np.array(x, np.float32, copy=True)
    lib().orc_soft_noneg(out.reshape(-1), out.size)
    return out


# ---- reference build wrappers --------------------------------------------

def ref_cyl_geometry(radius, height=1):
    R = ref()
    n = C.c_int(0)
    R.ref_cyl_geometry(float(radius), height, C.byref(n), None)
    cord = np.zeros(2 * n.value, np.int32)
    R.ref_cyl_geometry(float(radius), height, C.byref(n), cord.ctypes.data)
    return cord.reshape(-1, 2)


def ref_project(method, x, angles, nx, ny):
    na = len(angles)
    p = np.zeros((na, ny, nx), np.float32)
    st = ref().ref_project(_m(method), 0.5 * (nx - 1), nx, ny, np.ascontiguousarray(x, np.float32),
                           np.ascontiguousarray(angles, np.float32), na, p)
    if st:
        raise RuntimeError(f"ref_project status {st}")
    return p


def ref_backproject(method, p, angles, nx, ny):
    na = len(angles)
    nrays = len(ref_cyl_geometry(0.5 * (nx - 1)))
    x = np.zeros((nrays, ny), np.float32)
    st = ref().ref_backproject(_m(method), 0.5 * (nx - 1), nx, ny, np.ascontiguousarray(p, np.float32),
                               np.ascontiguousarray(angles, np.float32), na, x)
    if st:
        raise RuntimeError(f"ref_backproject status {st}")
    return x


def ref_solve(solver, method, b, angles, lam, maxit, tol, x0=None):
    na, ny, nx = b.shape
    nrays = len(ref_cyl_geometry(0.5 * (nx - 1)))
    x = np.zeros((nrays + 8 * nx, ny), np.float32) if x0 is None else np.array(x0, np.float32, copy=True)
    nnz = C.c_int(0)
    st = ref().ref_solve(1 if solver == "sirt" else 2, _m(method), np.ascontiguousarray(b, np.float32), nx, ny, na,
                         np.ascontiguousarray(angles, np.float32), lam, maxit, tol, 0 if x0 is None else 1, x,
                         C.byref(nnz))
    return st, x.reshape(-1)[: nnz.value].reshape(-1, ny)


def ref_mrc_write(path, data):
    """arecWriteImage of a [nz, ny, nx] float32 array through the reference build."""
    R = ref()
    a = np.ascontiguousarray(data, np.float32)
    nz, ny, nx = a.shape
    R.ref_mrc_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _F]
    return R.ref_mrc_write(os.fsencode(path), nx, ny, nz, a)


def ref_mrc_read(path, capacity):
    """arecReadImage through the reference build: (status, [nz, ny, nx] array)."""
    R = ref()
    R.ref_mrc_read.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _F,
                               C.c_long]
    nx, ny, nz = C.c_int(0), C.c_int(0), C.c_int(0)
    buf = np.zeros(capacity, np.float32)
    st = R.ref_mrc_read(os.fsencode(path), C.byref(nx), C.byref(ny), C.byref(nz), buf, capacity)
    return st, buf[: nx.value * ny.value * nz.value].reshape(nz.value, ny.value, nx.value)
