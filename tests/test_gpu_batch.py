"""GPU checks of the batched SIRT (many independent problems per launch) and
of the forward projector's launch plans.

* Forward projections are bit-identical under every plan (rows per lane
  1/2/4, any detector segmentation): each pixel sums the same table entries
  in the same order.
* Batched SIRT over problems of uneven width, height and angle count gives
  each problem exactly (bit for bit) the history and volume of its own
  single-problem SIRT, including problems in different rows-per-lane
  classes, on any forward design (design-4 problems sharing one forward
  launch per rows-per-lane class), and one that restarts.
* Each batched problem's history agrees with the FP32 oracle under the
  solver tests' tolerance (1e-5 relative, or 4x the FP32/FP64 oracle spread).
"""
import numpy as np
import pytest

import pyoracle as O
from test_gpu_solver import check_hist, problem

pytestmark = pytest.mark.gpu


def test_forward_plans_bit_identical(ctx):
    import arec3d_amd as A

    print("\n[case] forward plans", flush=True)
    nx, ny, na = 97, 24, 50
    rng = np.random.default_rng(5)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    x = ctx.to_device(rng.random(g.vol_shape).astype(np.float32))
    p = ctx.empty(g.proj_shape)
    g.project(x, p)
    ref = p.download()
    for rpl in (1, 2, 4):
        for groups in (1, 64, 4096):
            g.plan(rpl, groups)
            assert g.get_plan()[0] == rpl
            g.project(x, p)
            assert np.array_equal(p.download(), ref), (rpl, groups)
    with pytest.raises(A.ArecError):
        A.Geometry(ctx, "sq", nx, 6, ang).plan(4, 0)  # 6 rows: 4 per lane impossible


SHAPES = [  # (nx, ny, na, rows per lane)
    (64, 16, 45, 4),
    (48, 12, 30, 2),
    (80, 8, 90, 1),
    (41, 20, 20, 0),
    (33, 8, 30, 0),  # restarts (large lam below)
]


def _problems(ctx):
    import arec3d_amd as A

    out = []
    for k, (nx, ny, na, rpl) in enumerate(SHAPES):
        b, ang = problem("sq", nx, ny, na, 100 + k)
        g = A.Geometry(ctx, "sq", nx, ny, ang)
        if rpl:
            g.plan(rpl, 0)
        lam = 0.5 if nx == 33 else 1.0 / (na * nx)
        out.append((g, b, ang, lam))
    return out


def test_batched_sirt_equals_single(ctx):
    import arec3d_amd as A

    print("\n[case] batched sirt vs single", flush=True)
    probs = _problems(ctx)
    maxit, tol = 15, 1e-6
    bs = [ctx.to_device(b) for _, b, _, _ in probs]
    xs = [ctx.empty(g.vol_shape) for g, _, _, _ in probs]
    res = A.sirt_batch(ctx, [p[0] for p in probs], bs, xs, [p[3] for p in probs], maxit, tol)
    assert any(h[1] == 1 for h in res[-1][0])  # the last problem restarts
    for k, (g, b, ang, lam) in enumerate(probs):
        xk = ctx.empty(g.vol_shape)
        h1, l1 = g.sirt(bs[k], xk, lam, maxit, tol)
        hb, lb = res[k]
        assert hb == h1, k
        assert lb == l1
        assert np.array_equal(xs[k].download(), xk.download()), k


def test_batched_sirt_with_depth_strips_equals_single(ctx):
    """Problems on forward design 2 (depth strips) mixed with design-1
    problems: each batched problem still equals its single SIRT bit for
    bit (the batch launches every forward with the problem's own design)."""
    import arec3d_amd as A

    print("\n[case] batched sirt (depth strips) vs single", flush=True)
    probs = _problems(ctx)
    for k in (0, 2, 4):
        probs[k][0].set_forward(2, 0)
        assert probs[k][0].forward_design == 2
    maxit, tol = 10, 1e-6
    bs = [ctx.to_device(b) for _, b, _, _ in probs]
    xs = [ctx.empty(g.vol_shape) for g, _, _, _ in probs]
    res = A.sirt_batch(ctx, [p[0] for p in probs], bs, xs, [p[3] for p in probs], maxit, tol)
    for k, (g, b, ang, lam) in enumerate(probs):
        xk = ctx.empty(g.vol_shape)
        h1, l1 = g.sirt(bs[k], xk, lam, maxit, tol)
        assert res[k][0] == h1, k
        assert res[k][1] == l1
        assert np.array_equal(xs[k].download(), xk.download()), k


STAGED = [  # (nx, ny, na, design, rows per lane): two rows-per-lane classes of design 4
    (64, 140, 45, 4, 2),   # 140 rows: one whole 128-row chunk and a partial one
    (48, 12, 30, 4, 2),
    (97, 40, 37, 4, 2),    # angle count leaves a partial last group of 8
    (80, 260, 90, 4, 4),   # one whole 256-row chunk and a partial one
    (41, 20, 20, 4, 4),
    (56, 16, 24, 4, 1),    # the only problem of its class: launched on its own
    (33, 8, 30, 2, 2),     # depth strips: launched on its own
    (45, 9, 17, 1, 0),     # detector sweep: launched on its own
]


BACK_TILES = [(32, 2), (16, 4), (64, 2)]


@pytest.mark.parametrize("batched_forward", [True, False])
def test_batched_sirt_staged_forward_equals_single(ctx, monkeypatch, batched_forward):
    """Problems on forward design 4 share one forward launch per rows-per-lane
    class (classes of 2 and 3 problems here, whole and partial row chunks);
    a class of one, and problems on designs 1 and 2, get launches of their
    own; the back projections run as one persistent launch per back tile
    shape.  Each batched problem equals its single-problem SIRT bit for bit,
    with the shared forward launches and with one launch per problem
    (AREC_NO_FWD_BATCH)."""
    import arec3d_amd as A

    print(f"\n[case] batched sirt staged forward, shared launch {batched_forward}", flush=True)
    if not batched_forward:
        monkeypatch.setenv("AREC_NO_FWD_BATCH", "1")
    probs = []
    for k, (nx, ny, na, design, rpl) in enumerate(STAGED):
        b, ang = problem("sq", nx, ny, na, 200 + k)
        g = A.Geometry(ctx, "sq", nx, ny, ang)
        if design != 1:
            g.set_forward(design, rpl)
        assert g.forward_design == design
        g.set_back(*BACK_TILES[k % 3])  # three back classes: one persistent launch each
        probs.append((g, b, ang, 1.0 / (na * nx)))
    maxit, tol = 12, 1e-6
    bs = [ctx.to_device(b) for _, b, _, _ in probs]
    xs = [ctx.empty(g.vol_shape) for g, _, _, _ in probs]
    res = A.sirt_batch(ctx, [p[0] for p in probs], bs, xs, [p[3] for p in probs], maxit, tol)
    for k, (g, b, ang, lam) in enumerate(probs):
        xk = ctx.empty(g.vol_shape)
        h1, l1 = g.sirt(bs[k], xk, lam, maxit, tol)
        assert res[k][0] == h1, k
        assert res[k][1] == l1
        assert np.array_equal(xs[k].download(), xk.download()), k
    # and against the oracle on one staged problem of each shared class
    for k in (0, 4):
        g, b, ang, lam = probs[k]
        _, xo, ho, _ = O.sirt("sq", b, ang, lam, maxit, tol)
        _, _, ho64, _ = O.sirt("sq", b, ang, lam, maxit, tol, dtype=np.float64)
        check_hist(res[k][0], ho, ho64)
        x = xs[k].download().astype(np.float64)
        assert np.linalg.norm(x - xo) <= 1e-4 * np.linalg.norm(xo)


def test_batched_sirt_vs_oracle(ctx):
    import arec3d_amd as A

    print("\n[case] batched sirt vs oracle", flush=True)
    probs = _problems(ctx)[:4]
    maxit, tol = 12, 1e-6
    bs = [ctx.to_device(b) for _, b, _, _ in probs]
    xs = [ctx.empty(g.vol_shape) for g, _, _, _ in probs]
    res = A.sirt_batch(ctx, [p[0] for p in probs], bs, xs, [p[3] for p in probs], maxit, tol)
    for k, (g, b, ang, lam) in enumerate(probs):
        _, xo, ho, _ = O.sirt("sq", b, ang, lam, maxit, tol)
        _, _, ho64, _ = O.sirt("sq", b, ang, lam, maxit, tol, dtype=np.float64)
        check_hist(res[k][0], ho, ho64)
        x = xs[k].download().astype(np.float64)
        assert np.linalg.norm(x - xo) <= 1e-4 * np.linalg.norm(xo)


def test_batched_sirt_mixed_methods_rejected(ctx):
    import arec3d_amd as A

    ang = (np.arange(10) * np.pi / 10).astype(np.float32)
    g1, g2 = A.Geometry(ctx, "sq", 17, 4, ang), A.Geometry(ctx, "kb", 17, 4, ang)
    b = [ctx.empty(g1.proj_shape), ctx.empty(g2.proj_shape)]
    x = [ctx.empty(g1.vol_shape), ctx.empty(g2.vol_shape)]
    with pytest.raises(A.ArecError):
        A.sirt_batch(ctx, [g1, g2], b, x, [0.01, 0.01], 3, 0.0)
