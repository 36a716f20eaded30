"""The benchmark workloads at their real sizes, on the GPU, against the oracle.

* CGLS 512 x 512 x 256, 720 angles, Beer-Lambert data at 1e4 photons
  (areccgls.cpp:245-478): the full problem runs 30 iterations; a one-row
  instance of the same problem (CGLS couples rows through its scalars, so
  the instance is solved as its own problem on both sides) follows the FP64
  oracle's history and volume within 4x the FP32 oracle's distance from it
  (FP32 CGLS drifts on this noisy system, whatever the summation order).
* Batched SIRT, 64 samples of uneven width, height and angle count, 100
  iterations: 4 sampled problems equal their single solves bit for bit, and
  the cheapest problem's rows equal the oracle's row solves.
* The learned warm start at 512^3/720 on a phantom seed disjoint from the
  seeds its input iteration count was tuned on (49000-49002): the network is
  kept and reaches relnrm 1e-3 in at least 3x fewer passes than plain SIRT.
* The 2048-wide slab (radius 1023.5, 1800 angles): forward design 4 (the
  full slab's plan: staged strips at 2 rows per lane) against the oracle's
  rows with the adjoint identity on that geometry; forward design 1 at 4
  rows per lane and the back projector's computed taps match the oracle's
  rows; the whole 2048^2 x 1024 slab satisfies
  <Ax, y> = <x, A^T y> with element offsets past 2^31.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import pyoracle as O
from test_gpu_full_rows import column_blocks, oracle_back_rows, oracle_forward_rows, rel_views, workers
from test_gpu_solver import relnorm

pytestmark = pytest.mark.gpu


def test_cgls_512x512x256_720(ctx):
    import arec3d_amd as A
    from arec3d_amd import phantom as P

    nx, ny, na, iters = 512, 256, 720, 30
    print(f"\n[case] CGLS {nx}x{nx}x{ny}/{na}, Beer-Lambert 1e4 photons", flush=True)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    g.set_forward(2, 4)
    prm = P.ellipsoid_params(2, 40, 250.0, ny, axes=(8.0, 120.0), dens=(0.05, 1.0))
    xt = ctx.empty(g.vol_shape)
    P.phantom_device(ctx, xt, g.radius, ny, 250.0, prm)
    b = ctx.empty(g.proj_shape)
    g.project(xt, b)
    P.add_noise(ctx, b, 2, i0=1e4, scale=3.0 / A.reduce_max(ctx, b))
    x = ctx.empty(g.vol_shape)
    hist = g.cgls(b, x, iters, 0.0)
    assert len(hist) == iters + 1 and all(np.isfinite(h[2]) for h in hist)  # iteration 0, then 1..30
    assert min(h[2] for h in hist[1:]) < hist[0][2]
    r0 = ny // 2
    bsub = np.ascontiguousarray(b.download()[:, r0:r0 + 1, :])
    gs = A.Geometry(ctx, "sq", nx, 1, ang)
    xs = ctx.empty(gs.vol_shape)
    hs = gs.cgls(ctx.to_device(bsub), xs, iters, 0.0)
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        f32 = ex.submit(O.cgls, "sq", bsub, ang, iters, 0.0)
        f64 = ex.submit(O.cgls, "sq", bsub, ang, iters, 0.0, dtype=np.float64)
        (_, x32, h32), (_, x64, h64) = f32.result(), f64.result()
    xg = xs.download()
    spread = relnorm(x32, x64)
    print(f"row {r0}: GPU vs FP64 {relnorm(xg, x64):.3e}, FP32 oracle vs FP64 {spread:.3e}")
    # FP32 CGLS on this noisy system drifts from FP64 and the drift does not
    # shrink again: each printed relnrm of the GPU lies within 4x the largest
    # FP32-oracle-vs-FP64 deviation so far of the FP64 value (iteration
    # numbers and events exact)
    assert len(hs) == len(h64) == len(h32)
    env = 0.0
    for k, (g, o, d) in enumerate(zip(hs, h32, h64)):
        assert g[0] == d[0] and g[1] == d[1], (k, g, d)
        env = max(env, abs(o[2] - d[2]) / d[2])
        assert abs(g[2] - d[2]) / d[2] <= max(1e-5, 4 * env), (k, g, o, d, env)
    assert relnorm(xg, x64) <= max(1e-4, 4 * spread)


def _batch_problems(ctx):
    """The 64 seeded problems of the batched-SIRT workload (seed 3000)."""
    import arec3d_amd as A
    from arec3d_amd import phantom as P
    from arec3d_amd import unet as U

    rng = np.random.default_rng(3000)
    out = []
    for k in range(64):
        nx = int(rng.integers(96, 385))
        ny = int(rng.integers(32, 257))
        na = int(rng.choice([90, 180, 360, 720]))
        ne = int(rng.integers(8, 31))
        ang = (np.arange(na) * np.pi / na).astype(np.float32)
        g = A.Geometry(ctx, "sq", nx, ny, ang)
        sup = 0.47 * nx
        f = nx / 128.0
        prm = P.ellipsoid_params(3000 + k, ne, sup, ny, axes=(4.0 * f, 30.0 * f), dens=(0.1, 1.0))
        xt = ctx.empty(g.vol_shape)
        P.phantom_device(ctx, xt, g.radius, ny, sup, prm)
        b = ctx.empty(g.proj_shape)
        g.project(xt, b)
        P.add_noise(ctx, b, 3100 + k, rel_sigma=0.01)
        xt.free()
        lam = float(np.float32(0.9 / U.sigma_max2(ctx, g, iters=10)))
        out.append((g, b, lam))
    return out


def test_batched_sirt_64_samples(ctx):
    import arec3d_amd as A

    iters = 100
    print(f"\n[case] batched SIRT, 64 samples x {iters} iterations", flush=True)
    probs = _batch_problems(ctx)
    geoms, bs, lams = [p[0] for p in probs], [p[1] for p in probs], [p[2] for p in probs]
    xs = [ctx.empty(g.vol_shape) for g in geoms]
    res = A.sirt_batch(ctx, geoms, bs, xs, lams, iters, 0.0)
    assert all(len(r[0]) == iters and r[0][-1][2] < r[0][0][2] for r in res)
    for k in [int(i) for i in np.random.default_rng(7).choice(64, 4, replace=False)]:
        g, b, lam = probs[k]
        x1 = ctx.empty(g.vol_shape)
        h1, l1 = g.sirt(b, x1, lam, iters, 0.0)
        assert res[k][0] == h1 and res[k][1] == l1, k
        assert np.array_equal(xs[k].download(), x1.download()), k
    # the cheapest problem's rows against the oracle (no restart: rows are separable)
    k = int(np.argmin([g.nrays * g.na for g in geoms]))
    g, b, lam = probs[k]
    assert all(h[1] == 0 for h in res[k][0])
    rows = [0, g.ny // 2, g.ny - 1]
    bh = b.download()

    def one(r):
        st, xo, ho, _ = O.sirt("sq", np.ascontiguousarray(bh[:, r:r + 1, :]), g.angles, lam, iters, 0.0)
        assert st == 0 and all(h[1] == 0 for h in ho)
        return xo[:, 0]

    with cf.ThreadPoolExecutor(max_workers=workers(len(rows))) as ex:
        xo = np.stack(list(ex.map(one, rows)), axis=1)
    print(f"problem {k}: {g.nx}x{g.ny}x{g.na}")
    assert relnorm(xs[k].download()[:, rows], xo) <= 1e-4


def test_warm_start_512_held_out_seed(ctx):
    import arec3d_amd as A
    from arec3d_amd import phantom as P
    from arec3d_amd import unet as U

    n, na, seed, tol = 512, 720, 49010, 1e-3
    print(f"\n[case] warm start {n}^3/{na}, seed {seed}", flush=True)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, "sq", n, n, ang)
    g.set_forward(2, 4)
    xt = ctx.empty(g.vol_shape)
    prm = P.ellipsoid_params(seed, 40, 250.0, n, axes=(8.0, 120.0), dens=(0.05, 1.0))
    P.phantom_device(ctx, xt, g.radius, n, 250.0, prm)
    b = ctx.empty(g.proj_shape)
    g.project(xt, b)
    P.add_noise(ctx, b, seed, rel_sigma=0.01)
    xt.free()
    lam = 0.9 / U.sigma_max2(ctx, g)
    x = ctx.empty(g.vol_shape)
    hp, _ = g.sirt(b, x, lam, 3000, tol, hist_cap=3100)
    assert hp[-1][1] == 2 and hp[-1][2] < tol
    ws = U.WarmStart(ctx, g)
    (h1, h2), _, chosen = U.warm_sirt(ctx, g, b, x, lam, tol, ws)
    ws.close()
    warm = len(h1) + len(h2)
    print(f"passes to {tol:g}: plain {len(hp)}, warm {len(h1)} + {len(h2)} ({chosen})")
    assert chosen == "network"
    assert h2[-1][1] == 2 and h2[-1][2] < tol
    assert 3 * warm <= len(hp)


def test_slab_2048_rows_vs_oracle(ctx):
    """2048 wide, 256 rows, 1800 angles: design 1 at 4 rows per lane with
    whole chunks; the back projector computes its taps (its table would
    exceed its 16 GiB limit).  Two rows against the oracle."""
    import arec3d_amd as A

    nx, ny, na = 2048, 256, 1800
    print(f"\n[case] slab rows {nx}x{ny}x{na}", flush=True)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    g.set_forward(1, 4, 2048)
    assert g.forward_design == 1 and g.get_plan()[0] == 4
    rows = [0, 129]
    rng = np.random.default_rng(2048)
    x = rng.random(g.vol_shape, dtype=np.float32)
    p = rng.random(g.proj_shape, dtype=np.float32)
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        fo = ex.submit(oracle_forward_rows, "sq", x, ang, nx, rows)
        bo = ex.submit(oracle_back_rows, "sq", p, ang, nx, rows)
        dp, dx = ctx.empty(g.proj_shape), ctx.empty(g.vol_shape)
        g.project(ctx.to_device(x), dp)
        g.backproject(ctx.to_device(p), dx)
        gp, gx = dp.download()[:, rows, :], dx.download()[:, rows]
        op, ox = fo.result(), bo.result()
    assert rel_views(gp, op).max() <= 1e-5
    assert rel_views(column_blocks(gx), column_blocks(ox)).max() <= 1e-5


def test_slab_2048_rows_vs_oracle_staged(ctx):
    """2048 wide, 256 rows, 1800 angles on the forward the full slab's
    projection matching runs: design 4 (staged strips) at 2 rows per lane,
    its table of 7.2e9 entries (116 GB, the same at any row count) built by
    the parallel host ordering.  Two rows of the forward against the oracle,
    and the back projection's adjoint identity on that geometry."""
    import arec3d_amd as A

    nx, ny, na = 2048, 256, 1800
    print(f"\n[case] slab rows, staged strips {nx}x{ny}x{na}", flush=True)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    g.set_forward(4, 2, 0)
    assert g.forward_design == 4 and g.get_plan()[0] == 2
    rows = [0, 129]
    rng = np.random.default_rng(2049)
    x = rng.random(g.vol_shape, dtype=np.float32)
    with cf.ThreadPoolExecutor(max_workers=1) as ex:
        fo = ex.submit(oracle_forward_rows, "sq", x, ang, nx, rows)
        dx = ctx.to_device(x)
        dp = ctx.empty(g.proj_shape)
        g.project(dx, dp)
        gp = dp.download()[:, rows, :]
        y = ctx.empty(g.proj_shape)
        y.upload(rng.random(g.proj_shape, dtype=np.float32))
        aty = ctx.empty(g.vol_shape)
        g.backproject(y, aty)
        lhs, rhs = A.dot(ctx, dp, y), A.dot(ctx, dx, aty)
        scale = np.sqrt(A.dot(ctx, aty, aty) * A.dot(ctx, dx, dx))
        op = fo.result()
    assert rel_views(gp, op).max() <= 1e-5
    assert abs(lhs - rhs) <= 1e-5 * scale


def test_slab_2048x2048x1024_1800_adjoint(ctx):
    """The whole slab (13.5 GB volume, 15.1 GB of projections: element
    offsets past 2^31) satisfies <Ax, y> = <x, A^T y> to FP32 summation
    accuracy, with x and y drawn on the device."""
    import arec3d_amd as A
    from arec3d_amd import phantom as P

    nx, ny, na = 2048, 1024, 1800
    print(f"\n[case] slab adjoint {nx}x{nx}x{ny}/{na}", flush=True)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    assert g.nnz > 2 ** 31 and g.nproj > 2 ** 31
    x, y = ctx.empty(g.vol_shape).zero(), ctx.empty(g.proj_shape).zero()
    P.add_noise(ctx, x, 1, sigma=1.0)
    P.add_noise(ctx, y, 2, sigma=1.0)
    ax = ctx.empty(g.proj_shape)
    g.project(x, ax)
    lhs = A.dot(ctx, ax, y)
    ax.free()
    aty = ctx.empty(g.vol_shape)
    g.backproject(y, aty)
    rhs = A.dot(ctx, x, aty)
    scale = np.sqrt(A.dot(ctx, aty, aty) * A.dot(ctx, x, x))
    print(f"<Ax,y> {lhs:.9e}  <x,A^T y> {rhs:.9e}  scale {scale:.3e}")
    assert abs(lhs - rhs) <= 1e-5 * scale
    # the last rows of the last view and column are written (offsets past 2^31)
    tail = np.empty(4096, np.float32)
    from arec3d_amd import _check, lib
    import ctypes as C
    _check(lib.arec_memcpy_d2h(ctx.h, tail.ctypes.data_as(C.c_void_p),
                               C.c_void_p(aty.ptr.value + (g.nnz - 4096) * 4), 4096 * 4))
    assert np.all(np.isfinite(tail)) and np.any(tail != 0)
