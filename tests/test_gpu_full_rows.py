"""GPU parity of the projector instantiations that whole row chunks select.

Every forward design has a branch for a wave whose 64 x rows-per-lane rows
all lie inside the volume (vector loads of 2 or 4 rows per lane): design 1's
RPL == 4 / RPL == 2 loads, design 2's strip_walk<RPL, true>, design 3's
pair_walk<RPL, true> and design 4's staged_walk<RPL, true>.  They run
wherever the row count is a multiple of the chunk height -- the 512^3/720 solve runs design 2 at 4 rows per lane, the
2048^2 x 1024/1800 slab design 1 at 4 -- so these tests use 256 rows.

The oracle is the reference's loop order in C (angle -> column -> row,
arecproject.cpp:540-765).  Rows of the volume are independent 2-D problems
with the same weights, so the oracle is run on single rows (one thread
each) and compared with the same rows of the GPU's full-height result.
Tolerance: per projection view 1e-5 relative norm (FP32 summation order and
FP32-rounded weights against FP64-then-round per tap).
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5
ROWS = [0, 1, 63, 64, 130, 191, 254, 255]  # lane and chunk boundaries of every rows-per-lane plan
PLANS = [(2, 4), (2, 2), (2, 1), (1, 4), (1, 2), (1, 1), (3, 2), (3, 1), (4, 4), (4, 2), (4, 1)]


def workers(n):
    return max(1, min(n, os.cpu_count() or 1, 16))


def rel_views(a, b):
    a = a.reshape(a.shape[0], -1).astype(np.float64)
    b = b.reshape(b.shape[0], -1).astype(np.float64)
    nb = np.linalg.norm(b, axis=1)
    return np.linalg.norm(a - b, axis=1) / np.where(nb > 0, nb, 1)


def oracle_forward_rows(method, x, ang, nx, rows):
    """p[:, rows, :] of the oracle, one single-row problem per thread."""
    def one(r):
        return O.project(method, np.ascontiguousarray(x[:, r:r + 1]), ang, nx, 1)[:, 0, :]

    with cf.ThreadPoolExecutor(max_workers=workers(len(rows))) as ex:
        return np.stack(list(ex.map(one, rows)), axis=1)


def oracle_back_rows(method, p, ang, nx, rows):
    """x[:, rows] of the oracle's back projection, one row per thread."""
    def one(r):
        return O.backproject(method, np.ascontiguousarray(p[:, r:r + 1, :]), ang, nx, 1)[:, 0]

    with cf.ThreadPoolExecutor(max_workers=workers(len(rows))) as ex:
        return np.stack(list(ex.map(one, rows)), axis=1)


def column_blocks(x, width=64):
    """[nrays, k] -> blocks of `width` columns (zero padded), for per-block norms."""
    nb = (x.shape[0] + width - 1) // width
    pad = np.zeros((nb * width - x.shape[0], x.shape[1]), x.dtype)
    return np.concatenate([x, pad]).reshape(nb, -1)


@pytest.mark.parametrize("method", ["sq", "kb", "linear"])
def test_forward_full_row_chunks_vs_oracle(ctx, method):
    """512 wide, 256 rows, 720 angles: every design and rows-per-lane plan
    takes its whole-chunk branch; rows at lane and chunk boundaries match
    the oracle, and every plan matches design 2 at 4 rows per lane on all
    rows."""
    import arec3d_amd as A

    nx, ny, na = 512, 256, 720
    print(f"\n[case] full-row forward {method} {nx}x{ny}x{na}", flush=True)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, method, nx, ny, ang)
    rng = np.random.default_rng(512256)
    x = rng.random(g.vol_shape).astype(np.float32)
    if method == "linear":
        x[-2:] = 0.0  # the reference reads past cord for these two columns
    op = oracle_forward_rows(method, x, ang, nx, ROWS)
    dx, dp = ctx.to_device(x), ctx.empty(g.proj_shape)
    first = None
    for design, rpl in PLANS:
        print(f"[case]   design {design}, {rpl} rows/lane", flush=True)
        g.set_forward(design, rpl, 2048 if design == 1 else 0)
        assert g.forward_design == design and g.get_plan()[0] == rpl
        g.project(dx, dp)
        gp = dp.download()
        err = rel_views(gp[:, ROWS, :], op).max()
        assert err <= TOL, (design, rpl, err)
        if first is None:
            first = gp
        else:
            assert rel_views(gp, first).max() <= TOL, (design, rpl)


@pytest.mark.parametrize("method", ["sq", "kb", "linear"])
def test_back_full_row_tiles_vs_oracle(ctx, method):
    """The back projector at 512 wide, 256 rows, 720 angles (8 row blocks of
    32, its tap table): rows at block boundaries match the oracle per block
    of 64 columns."""
    import arec3d_amd as A

    nx, ny, na = 512, 256, 720
    print(f"\n[case] full-row back {method} {nx}x{ny}x{na}", flush=True)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, method, nx, ny, ang)
    rng = np.random.default_rng(256512)
    p = rng.random(g.proj_shape, dtype=np.float32)
    ox = oracle_back_rows(method, p, ang, nx, ROWS)
    dx = ctx.empty(g.vol_shape)
    g.backproject(ctx.to_device(p), dx)
    gx = dx.download()[:, ROWS]
    if method == "linear":  # columns the reference's linear pair reads past cord for
        gx, ox = gx[:-2], ox[:-2]
    assert rel_views(column_blocks(gx), column_blocks(ox)).max() <= TOL


@pytest.fixture(scope="module")
def sirt_256(ctx):
    """SIRT problem 256 wide, 256 rows, 360 angles (SQ): 12 seeded
    ellipsoids, 1% noise, lam = 0.9 / sigma_max^2; and the oracle's 30
    iterations on three single rows."""
    import arec3d_amd as A
    from arec3d_amd import phantom as P
    from arec3d_amd import unet as U

    nx, ny, na = 256, 256, 360
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    prm = P.ellipsoid_params(11, 12, 120.0, ny, axes=(8.0, 60.0), dens=(0.1, 1.0))
    xt = ctx.empty(g.vol_shape)
    P.phantom_device(ctx, xt, g.radius, ny, 120.0, prm)
    b = ctx.empty(g.proj_shape)
    g.project(xt, b)
    P.add_noise(ctx, b, 11, rel_sigma=0.01)
    lam = float(np.float32(0.9 / U.sigma_max2(ctx, g)))
    bh = b.download()
    rows = [5, 128, 250]

    def one(r):
        st, xo, ho, lo = O.sirt("sq", np.ascontiguousarray(bh[:, r:r + 1, :]), ang, lam, 30, 0.0)
        assert st == 0 and all(h[1] == 0 for h in ho)  # no restart: SIRT is separable by rows
        return xo[:, 0]

    with cf.ThreadPoolExecutor(max_workers=workers(len(rows))) as ex:
        xo = np.stack(list(ex.map(one, rows)), axis=1)
    return g, b, lam, rows, xo


@pytest.mark.parametrize("design,rpl", [(2, 4), (1, 4), (3, 2), (4, 4)])
def test_sirt_full_row_chunks_rows_vs_oracle(ctx, sirt_256, design, rpl):
    """30 SIRT iterations with the forward design's whole-chunk branch: with
    a fixed lam, no restart and no tolerance stop every row is its own
    problem, so the oracle's single-row solves give the GPU volume's rows
    (1e-4 relative: FP32 summation order over 30 iterations)."""
    g, b, lam, rows, xo = sirt_256
    print(f"\n[case] full-row SIRT design {design}, {rpl} rows/lane", flush=True)
    g.set_forward(design, rpl, 2048 if design == 1 else 0)
    assert g.forward_design == design and g.get_plan()[0] == rpl
    x = g.ctx.empty(g.vol_shape)
    hg, _ = g.sirt(b, x, lam, 30, 0.0)
    assert len(hg) == 30 and all(h[1] == 0 for h in hg)
    xg = x.download()[:, rows].astype(np.float64)
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) <= 1e-4
