"""Learned warm start for SIRT: a 3-level 3-D U-Net run by the HIP kernels of
unet.hip (no counterpart in the reference).

* :class:`UNet` wraps one network instance for a fixed volume size; the
  parameters are a flat FP32 vector shared between instances of different
  sizes (train at 128^3, run at 512^3).
* :func:`make_pairs` generates training pairs with the method itself: seeded
  phantoms, their sinograms, and SIRT after ``k_in`` and ``k_out``
  iterations.  Detector rows are independent, so ``stack`` phantoms are
  stacked along the rows and solved as one problem (SIRT with a fixed lam
  and no early termination is row-separable).
* :func:`warm_sirt` runs SIRT for ``k_in`` iterations, applies the network,
  keeps its output only if its residual is lower (SIRT remains the fallback
  and decides whether the network is used), then continues SIRT to the tolerance.
"""
import ctypes as C
import json
import os
import time

import numpy as np

from . import (ArecError, Context, DeviceArray, Geometry, Hist, _check, _hist, _ptr, dot, lib, scale,
               reduce_max)
from . import phantom as P


class UNet:
    def __init__(self, ctx, D, H, W, training=False, seed=0, params=None, scale=1.0):
        self.ctx = ctx
        self.shape = (D, H, W)
        self.h = C.c_void_p()
        _check(lib.arec_unet_create(ctx.h, C.byref(self.h), D, H, W, int(training), seed))
        self.nparams = int(lib.arec_unet_nparams(self.h))
        if params is not None:
            self.set_params(params)
        self.set_scale(scale)

    def set_scale(self, scale):
        self.scale = float(scale)
        _check(lib.arec_unet_set_scale(self.h, self.scale))

    def get_params(self):
        out = np.zeros(self.nparams, np.float32)
        _check(lib.arec_unet_get_params(self.h, _ptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        _check(lib.arec_unet_set_params(self.h, _ptr(p)))

    def forward(self, x_cube, y_cube):
        _check(lib.arec_unet_forward(self.h, x_cube.ptr, y_cube.ptr))

    def train_step(self, x_cube, t_cube, lr):
        loss = C.c_double()
        _check(lib.arec_unet_train_step(self.h, x_cube.ptr, t_cube.ptr, lr, C.byref(loss)))
        return loss.value

    def gradient(self, x_cube, t_cube):
        loss = C.c_double()
        g = np.zeros(self.nparams, np.float32)
        _check(lib.arec_unet_gradient(self.h, x_cube.ptr, t_cube.ptr, C.byref(loss), _ptr(g)))
        return loss.value, g

    def close(self):
        if self.h:
            lib.arec_unet_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cyl_to_cube(ctx, x, radius, ny, cube):
    _check(lib.arec_cyl_to_cube(ctx.h, x.ptr, float(radius), int(ny), cube.ptr))


def cube_to_cyl(ctx, cube, radius, ny, x):
    _check(lib.arec_cube_to_cyl(ctx.h, cube.ptr, float(radius), int(ny), x.ptr))


def sigma_max2(ctx, g, iters=20, seed=0):
    """Largest eigenvalue of A^T A by power iteration (FP64 dots)."""
    # random start on the device (SplitMix64 Gaussian), no host-sized array
    v = ctx.empty(g.vol_shape).zero()
    P.add_noise(ctx, v, 777 + seed, sigma=1.0)
    w = ctx.empty(g.proj_shape)
    v2 = ctx.empty(g.vol_shape)
    lam = 0.0
    for _ in range(iters):
        n = np.sqrt(dot(ctx, v, v))
        scale(ctx, v, 1.0 / n)
        g.project(v, w)
        g.backproject(w, v2)
        lam = dot(ctx, v, v2)
        v, v2 = v2, v
    return lam


def phantom_params(seed, n, height):
    """The eval generator of the 512^3 workload scaled to width n: support
    radius 250/512 n, 40 ellipsoids, semi-axes U[8,120]/512 n, density
    U[0.05, 1]."""
    f = n / 512.0
    return P.ellipsoid_params(seed, 40, 250.0 * f, height, axes=(8.0 * f, 120.0 * f), dens=(0.05, 1.0)), 250.0 * f


def make_pairs(ctx, n_samples, n=128, na=360, k_in=5, k_out=200, stack=16, seed0=40000, noise=0.01,
               log=print):
    """Training pairs (SIRT after k_in and after k_out iterations) as FP32
    cubes [n][n][n] on the host.  noise: Gaussian sigma relative to the
    sinogram's maximum, or a range (lo, hi) from which each stack draws its
    sigma log-uniformly (seeded by seed0)."""
    noise_rng = np.random.default_rng(seed0 + 7)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    ny = n * stack
    g = Geometry(ctx, "sq", n, ny, ang)
    lam = 0.9 / sigma_max2(ctx, g)
    xs = ctx.empty(g.vol_shape)
    x = ctx.empty(g.vol_shape)
    b = ctx.empty(g.proj_shape)
    cube = ctx.empty((ny, n, n))
    xin = np.zeros((n_samples, n, n, n), np.float32)
    xout = np.zeros((n_samples, n, n, n), np.float32)
    tpl = np.zeros(g.vol_shape, np.float32)
    t0 = time.time()
    for s0 in range(0, n_samples, stack):
        cnt = min(stack, n_samples - s0)
        # phantoms for the stack, one height-n block each
        for k in range(stack):
            prm, sup = phantom_params(seed0 + s0 + k, n, n)
            tmp = ctx.empty((g.nrays, n))
            P.phantom_device(ctx, tmp, g.radius, n, sup, prm)
            tpl[:, k * n:(k + 1) * n] = tmp.download()
        xs.upload(tpl)
        g.project(xs, b)
        if np.isscalar(noise):
            sig = noise
        else:
            sig = float(np.exp(noise_rng.uniform(np.log(noise[0]), np.log(noise[1]))))
        P.add_noise(ctx, b, seed0 + s0, rel_sigma=sig)
        g.sirt(b, x, lam, k_in, 0.0)
        cyl_to_cube(ctx, x, g.radius, ny, cube)
        c_in = cube.download().reshape(stack, n, n, n)
        g.sirt_warm(b, x, lam, k_out - k_in, 0.0)
        cyl_to_cube(ctx, x, g.radius, ny, cube)
        c_out = cube.download().reshape(stack, n, n, n)
        xin[s0:s0 + cnt] = c_in[:cnt]
        xout[s0:s0 + cnt] = c_out[:cnt]
        if log and (s0 // stack) % 10 == 0:
            log(f"pairs {s0 + cnt}/{n_samples} ({time.time() - t0:.1f}s)")
    return xin, xout, lam


def train(ctx, xin, xout, epochs=2, lr=1e-3, seed=0, log=print, scale=None):
    """Adam on the MSE between net(x_in) and x_out; returns (params, scale, losses)."""
    n = xin.shape[1]
    net = UNet(ctx, n, n, n, training=True, seed=seed)
    if scale is None:
        scale = 1.0 / max(1e-12, float(np.std(xin)))
    net.set_scale(scale)
    dx, dt = ctx.empty(xin.shape[1:]), ctx.empty(xin.shape[1:])
    rng = np.random.default_rng(seed)
    losses = []
    steps = epochs * len(xin)
    t0 = time.time()
    for e in range(epochs):
        for i, k in enumerate(rng.permutation(len(xin))):
            step = e * len(xin) + i
            cur = lr * 0.5 * (1 + np.cos(np.pi * step / steps))  # cosine decay
            dx.upload(xin[k])
            dt.upload(xout[k])
            losses.append(net.train_step(dx, dt, cur))
            if log and step % 200 == 0:
                log(f"step {step}/{steps} loss {np.mean(losses[-50:]):.4e} ({time.time() - t0:.1f}s)")
    return net.get_params(), scale, losses


def normal_residual(ctx, g, b, x, bnorm, tmp_p=None, tmp_x=None):
    """||A^T (b - A x)|| / bnorm with FP64 dots."""
    tmp_p = tmp_p or ctx.empty(g.proj_shape)
    tmp_x = tmp_x or ctx.empty(g.vol_shape)
    g.project(x, tmp_p)
    r = tmp_p.download()
    r = np.ascontiguousarray(b.download() - r)
    tmp_p.upload(r)
    g.backproject(tmp_p, tmp_x)
    return np.sqrt(dot(ctx, tmp_x, tmp_x)) / bnorm


DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_sq128.npz")


def load_weights(path=None):
    """(params, scale, meta) of a trained network (numpy, no pickles)."""
    with np.load(path or DEFAULT_WEIGHTS, allow_pickle=False) as z:
        return z["params"].astype(np.float32), float(z["scale"]), json.loads(str(z["meta"]))


class WarmStart:
    """The trained network and its buffers for one geometry's volume.

    Scale rule.  k SIRT iterations resolve about k cycles per field of view
    whatever the grid, so at width n the network's input is SIRT after
    k_in = k_train * n / n_train iterations: that has the training input's
    resolution in pixel units (5 iterations at 128 -> 20 at 512).

    Check.  SIRT's own residual decides whether the network's output is
    used: after check_pass SIRT passes from it, the printed relative
    residual must be below that of the SIRT iterate it replaced.  Its
    excess at the first pass is mostly low-frequency error, which SIRT
    removes in a few passes, so the check waits for them."""

    def __init__(self, ctx, g, params=None, scale=None, meta=None, check_pass=5):
        if params is None:
            params, scale, meta = load_weights()
        meta = meta or {"n": g.nx, "k_in": 5}
        n = g.nx
        self.ctx, self.g = ctx, g
        self.k_in = max(1, int(round(meta["k_in"] * n / meta["n"])))
        self.check_pass = check_pass
        self.cube_in = ctx.empty((g.ny, n, n))
        self.cube_out = ctx.empty((g.ny, n, n))
        self.x_keep = ctx.empty(g.vol_shape)
        self.net = UNet(ctx, g.ny, n, n, training=False, params=params, scale=scale)

    def apply(self, x):
        """x <- network(x) (in place, cylinder layout)."""
        cyl_to_cube(self.ctx, x, self.g.radius, self.g.ny, self.cube_in)
        self.net.forward(self.cube_in, self.cube_out)
        cube_to_cyl(self.ctx, self.cube_out, self.g.radius, self.g.ny, x)

    def close(self):
        self.net.close()


def warm_sirt(ctx, g, b, x, lam, tol, ws, k_in=None, maxit=3000, timing=False):
    """Warm-started SIRT: k_in SIRT iterations from zero (default: the
    WarmStart's scale rule), the network, then SIRT from its output to
    relnrm <= tol.  SIRT's residual decides on the network: if the residual
    printed ws.check_pass passes after the network is not below the last
    one of the k_in run, SIRT continues from the k_in iterate instead (the
    fallback).  Residuals use the fixed denominator ||A^T b|| throughout.
    Returns ((hist_in, hist_out), phase seconds, 'network' | 'sirt')."""
    from . import copy as dcopy

    t = {}
    k_in = ws.k_in if k_in is None else k_in
    t0 = time.perf_counter()
    h1, _ = g.sirt(b, x, lam, k_in, 0.0)
    if timing:
        ctx.synchronize()
        t["sirt_in"] = time.perf_counter() - t0
        t0 = time.perf_counter()
    dcopy(ctx, ws.x_keep, x)
    ws.apply(x)
    if timing:
        ctx.synchronize()
        t["net"] = time.perf_counter() - t0
        t0 = time.perf_counter()
    h2, _ = g.sirt_warm(b, x, lam, maxit, tol, hist_cap=maxit + 64, reject_above=h1[-1][2],
                        check_pass=ws.check_pass)
    chosen = "network"
    if h2 and h2[-1][1] == 3:
        chosen = "sirt"
        dcopy(ctx, x, ws.x_keep)
        h3, _ = g.sirt_warm(b, x, lam, maxit, tol, hist_cap=maxit + 64)
        h2 = h2 + h3
    ctx.synchronize()
    if timing:
        t["sirt_out"] = time.perf_counter() - t0
    return (h1, h2), t, chosen
