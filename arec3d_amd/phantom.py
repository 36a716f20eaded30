"""Seeded ellipsoid phantoms and noisy sinograms for synthetic workloads.

Parameters come from SplitMix64 (so CPU tests and GPU runs agree on the
phantom); evaluation runs on the device (arec_phantom) or, for small test
sizes, in numpy (``evaluate_host``, same FP64 formula).  Axes follow the
reference's layout: the cylinder disk is (ix, iz), the rows y run along the
rotation axis (AREC3D/AREC3D/arecImage.h:5-22).
"""
import ctypes as C
import math

import numpy as np

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / 9007199254740992.0))


def ellipsoid_params(seed, n_ell, support, height, axes=(4.0, 30.0), dens=(0.1, 1.0)):
    """n_ell ellipsoids: centres uniform in the support cylinder (rejection in
    the disk, uniform along the rows), semi-axes U[axes], a uniform random
    rotation (Shoemake's unit quaternion), density U[dens].  Returns an
    (n_ell, 16) float64 array in the layout of arec_phantom."""
    rng = SplitMix64(seed)
    out = np.zeros((n_ell, 16), np.float64)
    for e in range(n_ell):
        while True:
            cx, cz = rng.uniform(-support, support), rng.uniform(-support, support)
            if cx * cx + cz * cz <= support * support:
                break
        cy = rng.uniform(-0.5 * height, 0.5 * height)
        u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
        qw = math.sqrt(1 - u1) * math.sin(2 * math.pi * u2)
        qx = math.sqrt(1 - u1) * math.cos(2 * math.pi * u2)
        qy = math.sqrt(u1) * math.sin(2 * math.pi * u3)
        qz = math.sqrt(u1) * math.cos(2 * math.pi * u3)
        R = np.array([
            [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
            [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
            [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)],
        ])
        a = [rng.uniform(*axes) for _ in range(3)]
        d = rng.uniform(*dens)
        out[e, 0:3] = (cx, cy, cz)
        out[e, 3:12] = R.T.reshape(-1)  # world -> body
        out[e, 12:15] = a
        out[e, 15] = d
    return out


def evaluate_host(params, cord, ny, radius, support):
    """numpy evaluation with the formula of phantom_kernel (FP64)."""
    px = (cord[:, 0] - radius)[:, None]
    pz = (cord[:, 1] - radius)[:, None]
    py = (np.arange(ny) - 0.5 * (ny - 1))[None, :]
    v = np.zeros((len(cord), ny))
    inside = (px * px + pz * pz) <= support * support
    for e in params:
        dx, dy, dz = px - e[0], py - e[1], pz - e[2]
        r = e[3:12]
        u = r[0] * dx + r[1] * dy + r[2] * dz
        w = r[3] * dx + r[4] * dy + r[5] * dz
        t = r[6] * dx + r[7] * dy + r[8] * dz
        q = (u / e[12]) ** 2 + (w / e[13]) ** 2 + (t / e[14]) ** 2
        v += np.where(q <= 1.0, e[15], 0.0)
    return np.where(inside, v, 0.0).astype(np.float32)


def phantom_device(ctx, x, radius, ny, support, params):
    """Evaluate into the DeviceArray x ([nrays, ny])."""
    from . import _check, lib

    p = np.ascontiguousarray(params, np.float64)
    _check(lib.arec_phantom(ctx.h, float(radius), int(ny), float(support), p.ctypes.data_as(C.c_void_p),
                            len(p), x.ptr))


def add_noise(ctx, b, seed, sigma=None, rel_sigma=None, i0=None, scale=1.0):
    """Gaussian (sigma or rel_sigma * max(b)) or Beer-Lambert counts (i0)."""
    from . import _check, lib, reduce_max

    if i0 is not None:
        _check(lib.arec_add_noise(ctx.h, b.ptr, b.size, seed & MASK, 1, 0.0, float(i0), float(scale)))
        return
    if sigma is None:
        sigma = rel_sigma * reduce_max(ctx, b)
    _check(lib.arec_add_noise(ctx.h, b.ptr, b.size, seed & MASK, 0, float(sigma), 0.0, 0.0))
