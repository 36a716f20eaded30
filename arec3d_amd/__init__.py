"""arec3d_amd: MI355X-native reconstruction hot path of AREC3D.

Python view of libarec3d_amd.so (ctypes).  Two layers, mirroring the C ABI in
include/arec3d_amd.h:

* reference-style functions on host arrays with the reference's names,
  arguments and status codes: ``arecProject2D``, ``arecProject2D_KB``,
  ``arecProject2D_SQ``, the matching ``arecBackProject2D*``, ``cyl_sirt*``,
  ``cyl_cgls*`` and ``cyl_cgls*_stat`` (AREC3D/AREC3D/arecproject.h:55-62,
  arecsirt.h, areccgls.h, areccgls_stat.h), over :class:`ArecImage`;
* a device-resident API (:class:`Context`, :class:`Geometry`,
  :class:`DeviceArray`) used by the solvers, tests and benchmarks.

There is no CPU fallback: every call runs the HIP kernels, and importing the
package raises if the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libarec3d_amd.so")

LINEAR, KB, SQ = 1, 2, 3
METHODS = {"linear": LINEAR, "kb": KB, "sq": SQ}


class ArecError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"arec3d_amd error {rc}: {msg}")
        self.rc = rc


class Hist(C.Structure):
    _fields_ = [("iter", C.c_int32), ("event", C.c_int32), ("relnrm", C.c_double), ("rnorm", C.c_double)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m arec3d_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl, flt = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_float
    P = C.POINTER
    sig = {
        "arec_last_error": ([], C.c_char_p),
        "arec_version": ([], C.c_char_p),
        "arec_create": ([P(vp), C.c_int], C.c_int),
        "arec_destroy": ([vp], C.c_int),
        "arec_synchronize": ([vp], C.c_int),
        "arec_dev_alloc": ([vp, P(vp), C.c_size_t], C.c_int),
        "arec_dev_free": ([vp, vp], C.c_int),
        "arec_memcpy_h2d": ([vp, vp, vp, C.c_size_t], C.c_int),
        "arec_memcpy_d2h": ([vp, vp, vp, C.c_size_t], C.c_int),
        "arec_memset": ([vp, vp, C.c_int, C.c_size_t], C.c_int),
        "arec_cyl_nrays": ([dbl, vp], i64),
        "arec_geom_create": ([vp, P(vp), C.c_int, i64, i64, dbl, vp, i64], C.c_int),
        "arec_geom_destroy": ([vp], C.c_int),
        "arec_geom_info": ([vp, P(i64), P(i64), P(i64)], C.c_int),
        "arec_geom_device_bytes": ([vp], i64),
        "arec_project": ([vp, vp, vp, vp], C.c_int),
        "arec_backproject": ([vp, vp, vp, vp], C.c_int),
        "arec_sirt": ([vp, vp, vp, vp, flt, i32, flt, P(Hist), i32, P(i32), P(flt)], C.c_int),
        "arec_cgls": ([vp, vp, vp, vp, vp, i32, i32, flt, P(Hist), i32, P(i32)], C.c_int),
        "arec_init_weights": ([vp, vp, vp, i64, i64, i64], C.c_int),
        "arec_soft_noneg": ([vp, vp, i64], C.c_int),
        "arec_dot": ([vp, vp, vp, i64, P(dbl)], C.c_int),
        "arec_max": ([vp, vp, i64, P(flt)], C.c_int),
        "arec_copy": ([vp, vp, vp, i64], C.c_int),
        "arec_scale": ([vp, vp, i64, flt], C.c_int),
        "arec_time_operator": ([vp, vp, C.c_int, vp, vp, i32, P(flt)], C.c_int),
        "arec_phantom": ([vp, dbl, i64, dbl, vp, i32, vp], C.c_int),
        "arec_add_noise": ([vp, vp, i64, C.c_uint64, i32, dbl, dbl, dbl], C.c_int),
        "arec_align_ccorr": ([vp, vp, vp, i32, i32, i32, vp], C.c_int),
        "arec_align_cc": ([vp, vp, vp, i32, i32, i32, vp, vp], C.c_int),
        "arec_align_cc_noalloc": ([vp, vp, vp, i32, i32, i32, vp, vp], C.c_int),
        "arec_h_ccorr2d": ([i32, i32, vp, vp, vp], C.c_int),
        "arec_h_fshift2d": ([i32, i32, vp, vp, C.c_float, C.c_float], C.c_int),
        "arec_h_speak": ([i32, i32, vp, vp, vp], C.c_int),
        "arec_quad2dfit": ([vp, vp, vp, vp], C.c_int),
        "arec_align_peak": ([vp, vp, i32, i32, i32, vp, vp], C.c_int),
        "arec_align_shift": ([vp, vp, i32, i32, i32, vp, vp, flt], C.c_int),
        "arec_h_cc_images": ([i32, i32, i32, vp, i32, i32, i32, vp, vp, vp], C.c_int),
        "arec_h_shift_images": ([i32, i32, i32, vp, vp, vp], C.c_int),
        "arec_sirt_warm": ([vp, vp, vp, vp, flt, i32, flt, dbl, i32, P(Hist), i32, P(i32), P(flt)], C.c_int),
        "arec_sirt_batch": ([vp, i32, P(vp), P(vp), P(vp), P(flt), i32, flt, P(Hist), i32, P(i32), P(flt)],
                            C.c_int),
        "arec_geom_plan": ([vp, vp, i32, i32], C.c_int),
        "arec_geom_set_forward": ([vp, vp, i32, i32, i32], C.c_int),
        "arec_geom_forward_design": ([vp], i32),
        "arec_geom_set_back": ([vp, vp, i32, i32], C.c_int),
        "arec_geom_get_back": ([vp, P(i32), P(i32)], C.c_int),
        "arec_match": ([vp, vp, vp, vp, flt, i32, i32, flt, i32, i32, vp, vp, P(Hist)], C.c_int),
        "arec_match_ex": ([vp, vp, vp, vp, flt, i32, i32, flt, i32, i32, i32, vp, vp, vp, P(Hist)], C.c_int),
        "arec_rot_cc": ([vp, vp, vp, i32, i32, i32, i32, vp], C.c_int),
        "arec_rotate_views": ([vp, vp, i32, i32, i32, vp, i32], C.c_int),
        "arec_rotate_views_safe": ([vp, vp, vp, i32, i32, i32, vp, P(i32), P(dbl)], C.c_int),
        "arec_h_rot_cc_images": ([i32, i32, i32, vp, i32, i32, i32, vp, i32, vp], C.c_int),
        "arec_h_rotate_images": ([i32, i32, i32, vp, vp, i32], C.c_int),
        "arec_h_rotate_images_safe": ([i32, i32, i32, vp, vp, vp, P(i32), P(dbl)], C.c_int),
        "arec_align_centroid": ([vp, vp, vp, vp], C.c_int),
        "arec_take_log": ([vp, vp, i64], C.c_int),
        "arec_periodic_decomposition": ([vp, vp, i32, i32, i32], C.c_int),
        "arec_h_take_log": ([vp, i64], C.c_int),
        "arec_h_periodic_decomposition": ([i32, i32, i32, vp], C.c_int),
        "arec_mrc_read_header": ([C.c_char_p, P(i32), P(i32), P(i32), P(i32)], C.c_int),
        "arec_mrc_read": ([C.c_char_p, vp, i64], C.c_int),
        "arec_mrc_write": ([C.c_char_p, i32, i32, i32, vp], C.c_int),
        "arec_geom_get_plan": ([vp, P(i32), P(i32), P(i32)], C.c_int),
        "arec_unet_create": ([vp, P(vp), i32, i32, i32, i32, C.c_uint64], C.c_int),
        "arec_unet_destroy": ([vp], C.c_int),
        "arec_unet_nparams": ([vp], i64),
        "arec_unet_get_params": ([vp, vp], C.c_int),
        "arec_unet_set_params": ([vp, vp], C.c_int),
        "arec_unet_set_scale": ([vp, flt], C.c_int),
        "arec_unet_forward": ([vp, vp, vp], C.c_int),
        "arec_unet_train_step": ([vp, vp, vp, flt, P(dbl)], C.c_int),
        "arec_unet_gradient": ([vp, vp, vp, P(dbl), vp], C.c_int),
        "arec_conv3_check": ([vp, vp, i32, vp, i32, vp, vp, vp, i32, i32, i32, i32, i32], C.c_int),
        "arec_conv3_wgrad_check": ([vp, vp, i32, vp, i32, vp, i32, i32, i32, i32, vp], C.c_int),
        "arec_cyl_to_cube": ([vp, vp, dbl, i64, vp], C.c_int),
        "arec_cube_to_cyl": ([vp, vp, dbl, i64, vp], C.c_int),
        "arec_h_project": ([C.c_int, dbl, i32, i32, vp, vp, i32, vp], C.c_int),
        "arec_h_backproject": ([C.c_int, dbl, i32, i32, i32, vp, vp, i32, vp], C.c_int),
        "arec_h_solve": ([C.c_int, C.c_int, vp, i32, i32, i32, vp, flt, i32, flt, i32, vp, P(Hist), i32, P(i32), P(flt)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


lib = _load()

EXPORTED = [
    "arec_last_error", "arec_version", "arec_create", "arec_destroy", "arec_synchronize",
    "arec_dev_alloc", "arec_dev_free", "arec_memcpy_h2d", "arec_memcpy_d2h", "arec_memset",
    "arec_cyl_nrays", "arec_geom_create", "arec_geom_destroy", "arec_geom_info",
    "arec_geom_device_bytes", "arec_project", "arec_backproject", "arec_sirt", "arec_cgls",
    "arec_init_weights", "arec_soft_noneg", "arec_dot", "arec_max", "arec_copy", "arec_time_operator", "arec_phantom",
    "arec_add_noise", "arec_align_ccorr", "arec_align_cc", "arec_align_cc_noalloc", "arec_align_peak",
    "arec_align_shift", "arec_h_ccorr2d", "arec_h_fshift2d", "arec_h_speak", "arec_quad2dfit",
    "arec_h_cc_images", "arec_h_shift_images", "arec_sirt_warm", "arec_unet_create",
    "arec_unet_destroy", "arec_unet_nparams", "arec_unet_get_params", "arec_unet_set_params",
    "arec_unet_set_scale", "arec_unet_forward", "arec_unet_train_step", "arec_unet_gradient",
    "arec_conv3_check", "arec_conv3_wgrad_check", "arec_cyl_to_cube", "arec_cube_to_cyl", "arec_h_project", "arec_h_backproject",
    "arec_h_solve", "arec_sirt_batch", "arec_geom_plan", "arec_geom_get_plan", "arec_match", "arec_match_ex", "arec_align_centroid", "arec_take_log", "arec_periodic_decomposition",
    "arec_h_take_log", "arec_h_periodic_decomposition", "arec_scale", "arec_geom_set_forward",
    "arec_geom_forward_design", "arec_geom_set_back", "arec_geom_get_back", "arec_mrc_read_header", "arec_mrc_read", "arec_mrc_write",
    "arec_rot_cc", "arec_rotate_views", "arec_rotate_views_safe", "arec_h_rot_cc_images", "arec_h_rotate_images",
    "arec_h_rotate_images_safe",
]


def _check(rc):
    if rc != 0:
        raise ArecError(rc, lib.arec_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _method(m):
    return METHODS[m] if isinstance(m, str) else int(m)


def cyl_cord(radius):
    """(ix, iz) of every voxel column of arecAllocateCylImage (arecImage.cpp:29-86)."""
    n = lib.arec_cyl_nrays(float(radius), None)
    cord = np.zeros(2 * n, np.int32)
    lib.arec_cyl_nrays(float(radius), _ptr(cord))
    return cord.reshape(n, 2)


def cyl_nrays(radius):
    return int(lib.arec_cyl_nrays(float(radius), None))


# --------------------------------------------------------------------------
# Device-resident API
# --------------------------------------------------------------------------

class Context:
    """A device and a HIP stream (one per process and GPU is the usual use)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib.arec_create(C.byref(self.h), device))

    def synchronize(self):
        _check(lib.arec_synchronize(self.h))

    def close(self):
        if self.h:
            lib.arec_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def empty(self, shape, dtype=np.float32):
        return DeviceArray(self, shape, dtype)

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = DeviceArray(self, a.shape, a.dtype)
        d.upload(a)
        return d


class DeviceArray:
    """A device buffer owned by the library's allocator."""

    def __init__(self, ctx, shape, dtype=np.float32):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = C.c_void_p()
        _check(lib.arec_dev_alloc(ctx.h, C.byref(self.ptr), self.nbytes))

    @property
    def size(self):
        return int(np.prod(self.shape))

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _check(lib.arec_memcpy_h2d(self.ctx.h, self.ptr, _ptr(a), self.nbytes))
        return self

    def download(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib.arec_memcpy_d2h(self.ctx.h, _ptr(out), self.ptr, self.nbytes))
        return out

    def zero(self):
        _check(lib.arec_memset(self.ctx.h, self.ptr, 0, self.nbytes))
        return self

    def free(self):
        if self.ptr:
            lib.arec_dev_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Geometry:
    """Projector geometry: detector nx x ny, cylinder radius, FP32 angles.

    ``method`` is 'sq', 'kb' or 'linear' (reference pmethod 3, 2, 1).
    """

    def __init__(self, ctx, method, nx, ny, angles, radius=None):
        self.ctx = ctx
        self.method = _method(method)
        self.nx, self.ny = int(nx), int(ny)
        self.angles = np.ascontiguousarray(angles, np.float32)
        self.na = len(self.angles)
        self.radius = 0.5 * (self.nx - 1) if radius is None else float(radius)
        self.h = C.c_void_p()
        _check(lib.arec_geom_create(ctx.h, C.byref(self.h), self.method, self.nx, self.ny, self.radius,
                                    _ptr(self.angles), self.na))
        nr, nnz, npj = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib.arec_geom_info(self.h, C.byref(nr), C.byref(nnz), C.byref(npj)))
        self.nrays, self.nnz, self.nproj = nr.value, nnz.value, npj.value

    @property
    def vol_shape(self):
        return (self.nrays, self.ny)

    @property
    def proj_shape(self):
        return (self.na, self.ny, self.nx)

    def device_bytes(self):
        return int(lib.arec_geom_device_bytes(self.h))

    def project(self, x, p):
        _check(lib.arec_project(self.ctx.h, self.h, x.ptr, p.ptr))

    def backproject(self, p, x):
        _check(lib.arec_backproject(self.ctx.h, self.h, p.ptr, x.ptr))

    def time_operator(self, which, src, dst, reps):
        """ms per launch of the forward (which=0) or back (1) projector."""
        ms = C.c_float()
        _check(lib.arec_time_operator(self.ctx.h, self.h, which, src.ptr, dst.ptr, reps, C.byref(ms)))
        return ms.value

    def sirt(self, b, x, lam, maxit, tol, hist_cap=None):
        cap = hist_cap or (maxit + 64)
        H = (Hist * cap)()
        nh, lo = C.c_int32(), C.c_float()
        _check(lib.arec_sirt(self.ctx.h, self.h, b.ptr, x.ptr, lam, maxit, tol, H, cap, C.byref(nh), C.byref(lo)))
        return _hist(H, nh.value), lo.value

    def sirt_warm(self, b, x, lam, maxit, tol, hist_cap=None, reject_above=0.0, check_pass=1):
        """SIRT continued from the volume in x (fixed ||A^T b|| denominator).
        With reject_above > 0 the start is rejected (last history line has
        event 3) if the relative residual printed at pass check_pass is not
        below reject_above or a restart comes first."""
        cap = hist_cap or (maxit + 64)
        H = (Hist * cap)()
        nh, lo = C.c_int32(), C.c_float()
        _check(lib.arec_sirt_warm(self.ctx.h, self.h, b.ptr, x.ptr, lam, maxit, tol, float(reject_above),
                                  int(check_pass), H, cap, C.byref(nh), C.byref(lo)))
        return _hist(H, nh.value), lo.value

    def plan(self, rows_per_lane=0, min_groups=0):
        """Set the forward projector's launch plan (results do not depend on it)."""
        _check(lib.arec_geom_plan(self.ctx.h, self.h, int(rows_per_lane), int(min_groups)))

    def set_forward(self, design, rows_per_lane=0, min_groups=0):
        """Forward design 1 (detector sweep), 2 (depth strips), 3 (depth
        strips walked by angle pairs) or 4 (staged strips) and its plan."""
        _check(lib.arec_geom_set_forward(self.ctx.h, self.h, int(design), int(rows_per_lane), int(min_groups)))

    @property
    def forward_design(self):
        return int(lib.arec_geom_forward_design(self.h))

    def set_back(self, rows, angles):
        """Back projector tile shape: rows per workgroup x angles per LDS
        pass, one of (32, 2) (the default), (16, 4), (64, 2)."""
        _check(lib.arec_geom_set_back(self.ctx.h, self.h, int(rows), int(angles)))

    @property
    def back_tile(self):
        r, a = C.c_int32(), C.c_int32()
        _check(lib.arec_geom_get_back(self.h, C.byref(r), C.byref(a)))
        return r.value, a.value

    def get_plan(self):
        r, m, s = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib.arec_geom_get_plan(self.h, C.byref(r), C.byref(m), C.byref(s)))
        return r.value, m.value, s.value

    def cgls(self, b, x, maxit, tol, warm=False, w=None, hist_cap=None):
        cap = hist_cap or (maxit + 4)
        H = (Hist * cap)()
        nh = C.c_int32()
        _check(lib.arec_cgls(self.ctx.h, self.h, b.ptr, None if w is None else w.ptr, x.ptr, int(bool(warm)),
                             maxit, tol, H, cap, C.byref(nh)))
        return _hist(H, nh.value)

    def close(self):
        if self.h:
            lib.arec_geom_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _hist(H, n):
    return [(H[i].iter, H[i].event, H[i].relnrm, H[i].rnorm) for i in range(min(n, len(H)))]


def dot(ctx, a, b):
    out = C.c_double()
    _check(lib.arec_dot(ctx.h, a.ptr, b.ptr, a.size, C.byref(out)))
    return out.value


def sirt_batch(ctx, geoms, bs, xs, lams, maxit, tol, hist_cap=None):
    """Batched SIRT (arec_sirt_batch): problem k = (geoms[k], bs[k], xs[k],
    lams[k]); returns per problem (history, final lam)."""
    n = len(geoms)
    cap = hist_cap or (maxit + 64)
    H = (Hist * (cap * max(1, n)))()
    nh = (C.c_int32 * max(1, n))()
    lo = (C.c_float * max(1, n))()
    G = (C.c_void_p * max(1, n))(*[g.h for g in geoms])
    B = (C.c_void_p * max(1, n))(*[b.ptr for b in bs])
    X = (C.c_void_p * max(1, n))(*[x.ptr for x in xs])
    LM = (C.c_float * max(1, n))(*[float(v) for v in lams])
    _check(lib.arec_sirt_batch(ctx.h, n, G, B, X, LM, maxit, tol, H, cap, nh, lo))
    out = []
    for k in range(n):
        h = [(H[k * cap + i].iter, H[k * cap + i].event, H[k * cap + i].relnrm, H[k * cap + i].rnorm)
             for i in range(min(nh[k], cap))]
        out.append((h, lo[k]))
    return out


def match(ctx, g, views, x, lam, outer=5, maxit=20, tol=0.0, cc_passes=3, method="sirt", x_from="cc",
          rotate=False):
    """Projection matching (arec_match / arec_match_ex): returns (sx, sy,
    last history line of each outer loop's reconstruction), and the summed
    rotation angles (degrees) as a fourth item when rotate=True (the
    driver's align = 2).  x_from="cc" takes every shift from the
    cross-correlation passes (the reference's loop); x_from="centroid" takes
    the x shifts from the views' centroids."""
    sx = np.zeros(g.na, np.float32)
    sy = np.zeros(g.na, np.float32)
    rot = np.zeros(g.na, np.float32)
    H = (Hist * max(1, outer))()
    m = {"sirt": 1, "cgls": 2, "sirt_warm": 3}[method]
    flags = {"cc": 0, "centroid": 1}[x_from] | (2 if rotate else 0)
    _check(lib.arec_match_ex(ctx.h, g.h, views.ptr, x.ptr, lam, outer, maxit, tol, cc_passes, m, flags, _ptr(sx),
                             _ptr(sy), _ptr(rot), H))
    return (sx, sy, _hist(H, outer), rot) if rotate else (sx, sy, _hist(H, outer))


ROT_RULES = {"skew": 0, "bilinear": 1}


def rot_r2(nx, ny):
    """The driver's profile radius for the rotational step: min(nx, ny - 10) // 2."""
    return min(nx, ny - 10) // 2


def rot_cc(ctx, images1, images2, r2):
    """arecRotCCImages on [nv, ny, nx] DeviceArrays: angles (degrees) of images2 against images1."""
    nv, ny, nx = images1.shape
    ang = np.zeros(nv, np.float32)
    _check(lib.arec_rot_cc(ctx.h, images1.ptr, images2.ptr, nx, ny, nv, r2, _ptr(ang)))
    return ang


def rotate_views(ctx, views, angles, rule="skew"):
    """arecRotateImages_skew (rule "skew") or arecRotateImages ("bilinear")
    on a DeviceArray, in place; returns the angles as the reference leaves them."""
    nv, ny, nx = views.shape
    ang = np.array(angles, np.float32, copy=True)
    _check(lib.arec_rotate_views(ctx.h, views.ptr, nx, ny, nv, _ptr(ang), ROT_RULES[rule]))
    return ang


def rotate_views_safe(ctx, views, ref, angles):
    """arecRotateImages_skew_safe on DeviceArrays (views in place): (angles, kept, mean cc)."""
    nv, ny, nx = views.shape
    ang = np.array(angles, np.float32, copy=True)
    kept, mcc = C.c_int32(0), C.c_double(0.0)
    _check(lib.arec_rotate_views_safe(ctx.h, views.ptr, ref.ptr, nx, ny, nv, _ptr(ang), C.byref(kept),
                                      C.byref(mcc)))
    return ang, kept.value, mcc.value


def take_log(ctx, data):
    """arecImageTakeLog on a DeviceArray, in place: -log(x) for x > 0, else 0."""
    _check(lib.arec_take_log(ctx.h, data.ptr, int(np.prod(data.shape))))


def periodic_decomposition(ctx, views):
    """periodicDecomposition of a view stack ([nv, ny, nx] DeviceArray), in place."""
    nv, ny, nx = views.shape
    _check(lib.arec_periodic_decomposition(ctx.h, views.ptr, nx, ny, nv))


def read_mrc_header(path):
    """(nx, ny, nz, mode) of an MRC file (arecReadImageHeader)."""
    v = [C.c_int32() for _ in range(4)]
    _check(lib.arec_mrc_read_header(os.fsencode(path), *[C.byref(x) for x in v]))
    return tuple(x.value for x in v)


def read_mrc(path):
    """An MRC file as a float32 array [nz, ny, nx] (arecReadImage: modes 1, 2, 6)."""
    nx, ny, nz, _ = read_mrc_header(path)
    out = np.empty((max(nz, 0), max(ny, 0), max(nx, 0)), np.float32)
    _check(lib.arec_mrc_read(os.fsencode(path), _ptr(out), out.size))
    return out


def write_mrc(path, data):
    """A float32 array [nz, ny, nx] (or [ny, nx]) as a mode-2 MRC file (arecWriteImage)."""
    a = np.ascontiguousarray(data, np.float32)
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3:
        raise ValueError("write_mrc takes a 2-D or 3-D array")
    nz, ny, nx = a.shape
    _check(lib.arec_mrc_write(os.fsencode(path), nx, ny, nz, _ptr(a)))


def align_centroid(ctx, g, views):
    """x shift of every view from its centroid (arec_align_centroid)."""
    sx = np.zeros(g.na, np.float32)
    _check(lib.arec_align_centroid(ctx.h, g.h, views.ptr, _ptr(sx)))
    return sx


def align_ccorr(ctx, x, y, c):
    """Correlation maps of view stacks x, y ([nv, ny, nx] DeviceArrays) into c."""
    nv, ny, nx = x.shape
    _check(lib.arec_align_ccorr(ctx.h, x.ptr, y.ptr, nx, ny, nv, c.ptr))


def align_cc(ctx, proj, meas, sx, sy):
    """Sub-pixel shifts of meas against proj (arecCCImages) into sx, sy."""
    nv, ny, nx = proj.shape
    _check(lib.arec_align_cc(ctx.h, proj.ptr, meas.ptr, nx, ny, nv, sx.ptr, sy.ptr))


def align_cc_noalloc(ctx, proj, meas, sx, sy):
    """align_cc through the entry point that allocates nothing (its work
    buffers live in a fixed device array; calls overlapping on one GPU race)."""
    nv, ny, nx = proj.shape
    _check(lib.arec_align_cc_noalloc(ctx.h, proj.ptr, meas.ptr, nx, ny, nv, sx.ptr, sy.ptr))


def align_peak(ctx, cmap, sx, sy):
    """speak's sub-pixel peak of each map ([nv, ny, nx] DeviceArray)."""
    nv, ny, nx = cmap.shape
    _check(lib.arec_align_peak(ctx.h, cmap.ptr, nx, ny, nv, sx.ptr, sy.ptr))


def align_shift(ctx, data, sx, sy, sign=-1.0):
    """Fourier shift of every view by (sign*sx, sign*sy) (arecShiftImages: sign -1)."""
    nv, ny, nx = data.shape
    _check(lib.arec_align_shift(ctx.h, data.ptr, nx, ny, nv, sx.ptr, sy.ptr, sign))


def ccorr2d(nx, ny, x, y, c):
    """align2dstack.cpp:480-521 on host arrays [ny][nx]: c = IDFT(conj(DFT x) DFT y)."""
    _check(lib.arec_h_ccorr2d(nx, ny, _ptr(_f32c(x)), _ptr(_f32c(y)), _ptr(c)))


def fshift2d(nx, ny, imagein, imageout, sx, sy):
    """align2dstack.cpp:566-630: imageout = imagein shifted by (+sx, +sy)."""
    _check(lib.arec_h_fshift2d(nx, ny, _ptr(_f32c(imagein)), _ptr(imageout), float(sx), float(sy)))


def speak(nx, ny, data):
    """align2dstack.cpp:1059-1105: the sub-pixel peak (xt, yt) of one map."""
    xt, yt = C.c_float(), C.c_float()
    _check(lib.arec_h_speak(nx, ny, _ptr(_f32c(data)), C.byref(xt), C.byref(yt)))
    return xt.value, yt.value


def quad2dfit(z):
    """align2dstack.cpp:1125-1166 on 9 values (host): (xsh, ysh, peakv); peakv
    is NaN for a singular fit (the reference leaves it unset)."""
    xs, ys, pv = C.c_float(), C.c_float(), C.c_float(float("nan"))
    _check(lib.arec_quad2dfit(_ptr(_f32c(z)), C.byref(xs), C.byref(ys), C.byref(pv)))
    return xs.value, ys.value, pv.value


def _f32c(a):
    return np.ascontiguousarray(a, np.float32)


def arecCCImages(comm, images1, images2, sx, sy):
    """arecCCImages (arecImage.cpp:384-434) on host ArecImage stacks."""
    a = np.ascontiguousarray(images1.data, np.float32)
    b = np.ascontiguousarray(images2.data, np.float32)
    ox = np.zeros(images1.nz, np.float32)
    oy = np.zeros(images1.nz, np.float32)
    rc = lib.arec_h_cc_images(images1.nx, images1.ny, images1.nz, _ptr(a), images2.nx, images2.ny, images2.nz,
                              _ptr(b), _ptr(ox), _ptr(oy))
    if rc <= -100:
        raise ArecError(rc, lib.arec_last_error().decode())
    if rc == 0:
        sx[: images1.nz] = ox
        sy[: images1.nz] = oy
    return rc


def arecShiftImages(images, sx, sy):
    """arecShiftImages (arecImage.cpp:653-671): views shifted by (-sx, -sy)."""
    d = np.ascontiguousarray(images.data, np.float32)
    ax = np.ascontiguousarray(sx, np.float32)
    ay = np.ascontiguousarray(sy, np.float32)
    _check(lib.arec_h_shift_images(images.nx, images.ny, images.nz, _ptr(d), _ptr(ax), _ptr(ay)))
    images.data[...] = d


def copy(ctx, dst, src):
    _check(lib.arec_copy(ctx.h, dst.ptr, src.ptr, src.size))


def scale(ctx, x, a):
    """x *= a on the device."""
    _check(lib.arec_scale(ctx.h, x.ptr, x.size, float(a)))


def reduce_max(ctx, a):
    out = C.c_float()
    _check(lib.arec_max(ctx.h, a.ptr, a.size, C.byref(out)))
    return out.value


def init_weights(ctx, b, w, nx, ny, nz):
    _check(lib.arec_init_weights(ctx.h, b.ptr, w.ptr, nx, ny, nz))


def soft_noneg(ctx, x):
    _check(lib.arec_soft_noneg(ctx.h, x.ptr, x.size))


# --------------------------------------------------------------------------
# Reference-style interface on host arrays
# --------------------------------------------------------------------------

class ArecImage:
    """Host mirror of the reference's arecImage (arecImage.h:5-22): a cube
    ``data[nz][ny][nx]`` or a compact cylinder ``data[nrays][height]``."""

    def __init__(self, data=None, nx=0, ny=0, nz=0, radius=0.0, height=0, cord=None, is_cyl=False):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.radius, self.height = radius, height
        self.cord = cord
        self.is_cyl = is_cyl
        self.data = data

    @property
    def nrays(self):
        return 0 if self.cord is None else len(self.cord)

    @property
    def nnz(self):
        return self.nrays * self.height


def arecAllocateCBImage(nx, ny, nz):
    """arecAllocateCBImage (arecImage.cpp:15-27)."""
    return ArecImage(np.zeros((nz, ny, nx), np.float32), nx, ny, nz)


def arecAllocateCylImage(radius, height):
    """arecAllocateCylImage (arecImage.cpp:29-86)."""
    cord = cyl_cord(radius)
    n = int(round(2 * radius + 1))
    return ArecImage(np.zeros((len(cord), height), np.float32), n, height, n, float(radius), height, cord, True)


def _project(method, cylvol, angles, projstack):
    if not cylvol.is_cyl:
        return 1
    if projstack.nx != cylvol.nx or projstack.ny != cylvol.ny:
        return 2
    ang = np.ascontiguousarray(angles, np.float32)
    out = np.zeros((len(ang), cylvol.ny, cylvol.nx), np.float32)
    x = np.ascontiguousarray(cylvol.data, np.float32)
    rc = lib.arec_h_project(method, cylvol.radius, cylvol.nx, cylvol.ny, _ptr(x), _ptr(ang), len(ang), _ptr(out))
    if rc < 0 and rc <= -100:
        raise ArecError(rc, lib.arec_last_error().decode())
    if rc == 0:
        projstack.data[: len(ang)] = out
    return rc


def _backproject(method, images, angles, bvol):
    ang = np.ascontiguousarray(angles, np.float32)
    if images.nz != len(ang):
        return 1
    if bvol.nx != images.nx or bvol.ny != images.ny or not bvol.is_cyl:
        return -1
    out = np.zeros((bvol.nrays, images.ny), np.float32)
    p = np.ascontiguousarray(images.data, np.float32)
    rc = lib.arec_h_backproject(method, bvol.radius, images.nx, images.ny, images.nz, _ptr(p), _ptr(ang),
                                len(ang), _ptr(out))
    if rc <= -100:
        raise ArecError(rc, lib.arec_last_error().decode())
    if rc == 0:
        bvol.data[...] = out
    return rc


def arecProject2D(cylvol, angles, nangles, projstack):
    return _project(LINEAR, cylvol, np.asarray(angles)[:nangles], projstack)


def arecProject2D_KB(cylvol, angles, nangles, projstack):
    return _project(KB, cylvol, np.asarray(angles)[:nangles], projstack)


def arecProject2D_SQ(cylvol, angles, nangles, projstack):
    return _project(SQ, cylvol, np.asarray(angles)[:nangles], projstack)


def arecBackProject2D(images, angles, nangles, bvol):
    return _backproject(LINEAR, images, np.asarray(angles)[:nangles], bvol)


def arecBackProject2D_KB(images, angles, nangles, bvol):
    return _backproject(KB, images, np.asarray(angles)[:nangles], bvol)


def arecBackProject2D_SQ(images, angles, nangles, bvol):
    return _backproject(SQ, images, np.asarray(angles)[:nangles], bvol)


def _solve(solver, method, images, angles, xvol, lam, maxit, tol):
    """Shared body of cyl_sirt*/cyl_cgls*: returns the status; fills xvol
    (allocating it like the reference when xvol.data is None) and stores the
    printed history in xvol.history."""
    nx, ny, nz = images.nx, images.ny, images.nz
    radius = 0.5 * (nx - 1)
    if (solver == 2 and method == KB) or solver == 3:
        radius = float(nx // 2)
    warm = xvol.data is not None and solver != 1
    if xvol.data is None:
        fresh = arecAllocateCylImage(radius, ny)
        xvol.__dict__.update(fresh.__dict__)
    b = np.ascontiguousarray(images.data, np.float32)
    ang = np.ascontiguousarray(angles, np.float32)[:nz]
    x = np.ascontiguousarray(xvol.data, np.float32).copy()
    cap = maxit * 22 + 64
    H = (Hist * cap)()
    nh, lo = C.c_int32(), C.c_float()
    rc = lib.arec_h_solve(solver, method, _ptr(b), nx, ny, nz, _ptr(ang), lam, maxit, tol, int(warm), _ptr(x),
                          H, cap, C.byref(nh), C.byref(lo))
    if rc <= -100:
        raise ArecError(rc, lib.arec_last_error().decode())
    if rc == 0:
        xvol.data[...] = x
    xvol.history = _hist(H, nh.value)
    xvol.lam = lo.value
    return rc


def cyl_sirt(comm, images, angles, xvol, lam, maxit, tol):
    return _solve(1, LINEAR, images, angles, xvol, lam, maxit, tol)


def cyl_sirt_KB(comm, images, angles, xvol, lam, maxit, tol):
    return _solve(1, KB, images, angles, xvol, lam, maxit, tol)


def cyl_sirt_SQ(comm, images, angles, xvol, lam, maxit, tol):
    return _solve(1, SQ, images, angles, xvol, lam, maxit, tol)


def cyl_cgls(comm, images, angles, xvol, lam, maxit, tol):
    return _solve(2, LINEAR, images, angles, xvol, lam, maxit, tol)


def cyl_cgls_KB(comm, images, angles, xvol, lam, maxit, tol):
    return _solve(2, KB, images, angles, xvol, lam, maxit, tol)


def cyl_cgls_SQ(comm, images, angles, xvol, lam, maxit, tol):
    return _solve(2, SQ, images, angles, xvol, lam, maxit, tol)


def cyl_cgls_stat(comm, images, angles, xvol, lam, maxit, tol):
    return _solve(3, LINEAR, images, angles, xvol, lam, maxit, tol)


def cyl_cgls_KB_stat(comm, images, angles, xvol, lam, maxit, tol):
    return _solve(3, KB, images, angles, xvol, lam, maxit, tol)


def cyl_cgls_SQ_stat(comm, images, angles, xvol, lam, maxit, tol):
    return _solve(3, SQ, images, angles, xvol, lam, maxit, tol)
