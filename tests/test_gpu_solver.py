"""GPU parity of the device-resident SIRT and CGLS against the oracle
(FP32 reference semantics) and the reference's golden vectors.

History tolerance: at every printed iteration the GPU's relnrm and rnorm
agree with the FP32 oracle to 1e-5 relative, or -- once the residual has
dropped so far that FP32 rounding dominates it -- to within 4x the distance
between the FP32 oracle and the FP64 oracle (the noise floor of computing
the reference's recurrence in FP32, whatever the summation order).
Iteration numbers, events (restarts, termination), restart counts and final
lam are exact; volumes match to 1e-4 relative norm.
"""
import os

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")


def problem(method, nx, ny, na, seed, noise=0.01):
    rng = np.random.default_rng(seed)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    nr = len(O.cyl_geometry(0.5 * (nx - 1)))
    xt = np.zeros((nr, ny), np.float32)
    cord = O.cyl_geometry(0.5 * (nx - 1))
    r = 0.5 * (nx - 1)
    d2 = (cord[:, 0] - r) ** 2 + (cord[:, 1] - r) ** 2
    xt[:] = ((d2 < (0.6 * r) ** 2)[:, None] * (0.5 + 0.5 * rng.random((1, ny)))).astype(np.float32)
    b = O.project(method, xt, ang, nx, ny)
    b = (b + noise * b.max() * rng.standard_normal(b.shape)).astype(np.float32)
    return b, ang


def check_hist(gh, oh, oh64=None):
    assert len(gh) == len(oh)
    for k, (g, o) in enumerate(zip(gh, oh)):
        assert g[0] == o[0] and g[1] == o[1], (k, g, o)
        for c in (2, 3):
            tol = 1e-5 * abs(o[c])
            if oh64 is not None and k < len(oh64):
                tol = max(tol, 4 * abs(o[c] - oh64[k][c]))
            assert abs(g[c] - o[c]) <= tol, (k, c, g, o, None if oh64 is None else oh64[k])


def relnorm(a, b):
    return float(np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("method", ["sq", "kb", "linear"])
def test_sirt_history_vs_oracle(ctx, method):
    import arec3d_amd as A

    nx, ny, na = 65, 40, 60
    print(f"\n[case] sirt {method}", flush=True)
    b, ang = problem(method, nx, ny, na, 1)
    lam = 1.0 / (na * nx)  # stable for these footprints
    st, xo, ho, lo = O.sirt(method, b, ang, lam, 30, 1e-6)
    _, _, ho64, _ = O.sirt(method, b, ang, lam, 30, 1e-6, dtype=np.float64)
    g = A.Geometry(ctx, method, nx, ny, ang)
    db, dx = ctx.to_device(b), ctx.empty(g.vol_shape)
    hg, lg = g.sirt(db, dx, lam, 30, 1e-6)
    check_hist(hg, ho, ho64)
    assert lg == lo
    assert relnorm(dx.download(), xo) <= 1e-4


def test_sirt_restart_chain_vs_oracle(ctx):
    import arec3d_amd as A

    nx, ny, na = 33, 8, 30
    print("\n[case] sirt restarts", flush=True)
    b, ang = problem("sq", nx, ny, na, 2)
    st, xo, ho, lo = O.sirt("sq", b, ang, 0.5, 12, 1e-6)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    db, dx = ctx.to_device(b), ctx.empty(g.vol_shape)
    hg, lg = g.sirt(db, dx, 0.5, 12, 1e-6)
    assert sum(h[1] == 1 for h in ho) >= 2
    check_hist(hg, ho)
    assert lg == lo
    assert relnorm(dx.download(), xo) <= 1e-4


def test_sirt_tolerance_termination(ctx):
    import arec3d_amd as A

    nx, ny, na = 33, 8, 30
    b, ang = problem("sq", nx, ny, na, 4, noise=0.0)
    lam = 1.0 / (na * nx)
    st, xo, ho, lo = O.sirt("sq", b, ang, lam, 500, 0.05)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    db, dx = ctx.to_device(b), ctx.empty(g.vol_shape)
    hg, lg = g.sirt(db, dx, lam, 500, 0.05)
    assert ho[-1][1] == 2 and hg[-1][1] == 2 and hg[-1][0] == ho[-1][0]
    check_hist(hg, ho)


@pytest.mark.parametrize("method", ["sq", "kb", "linear"])
def test_cgls_history_vs_oracle(ctx, method):
    import arec3d_amd as A

    nx, ny, na = 65, 40, 60
    print(f"\n[case] cgls {method}", flush=True)
    b, ang = problem(method, nx, ny, na, 3)
    st, xo, ho = O.cgls(method, b, ang, 15, 1e-6)
    _, x64, ho64 = O.cgls(method, b, ang, 15, 1e-6, dtype=np.float64)
    g = A.Geometry(ctx, method, nx, ny, ang)
    db, dx = ctx.to_device(b), ctx.empty(g.vol_shape)
    hg = g.cgls(db, dx, 15, 1e-6)
    check_hist(hg, ho, ho64)
    assert relnorm(dx.download(), xo) <= max(1e-4, 4 * relnorm(xo, x64))


def test_cgls_warm_start_vs_oracle(ctx):
    import arec3d_amd as A

    nx, ny, na = 65, 16, 45
    b, ang = problem("sq", nx, ny, na, 5)
    _, x0, _ = O.cgls("sq", b, ang, 3, 1e-9)
    st, xo, ho = O.cgls("sq", b, ang, 6, 1e-9, x0=x0)
    _, x64, ho64 = O.cgls("sq", b, ang, 6, 1e-9, x0=x0, dtype=np.float64)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    db, dx = ctx.to_device(b), ctx.to_device(x0)
    hg = g.cgls(db, dx, 6, 1e-9, warm=True)
    check_hist(hg, ho, ho64)
    assert relnorm(dx.download(), xo) <= max(1e-4, 4 * relnorm(xo, x64))


def test_weighted_cgls_vs_oracle(ctx):
    import arec3d_amd as A

    nx, ny, na = 65, 12, 45  # odd width: the weighted solvers' nx/2 radius works
    b, ang = problem("sq", nx, ny, na, 6)
    b = (b / b.max()).astype(np.float32)
    st, xo, ho = O.cgls("sq", b, ang, 8, 1e-9, weighted=True)
    _, x64, ho64 = O.cgls("sq", b, ang, 8, 1e-9, weighted=True, dtype=np.float64)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    db, dx, dw = ctx.to_device(b), ctx.empty(g.vol_shape), ctx.empty(b.shape)
    A.init_weights(ctx, db, dw, nx, ny, na)
    np.testing.assert_allclose(dw.download(), O.init_weights(b), rtol=2e-7, atol=1e-7)
    hg = g.cgls(db, dx, 8, 1e-9, w=dw)
    check_hist(hg, ho, ho64)
    assert relnorm(dx.download(), xo) <= max(1e-4, 4 * relnorm(xo, x64))


@pytest.mark.parametrize("tag,nx,ny", [("odd", 33, 6), ("even", 32, 5)])
def test_solvers_vs_reference_golden(ctx, tag, nx, ny):
    """Drop-in host entry points against vectors from the reference build."""
    import arec3d_amd as A

    G = np.load(GOLD)
    ang, b = G[f"{tag}_angles"], G[f"{tag}_b"]
    img = A.ArecImage(b, nx, ny, len(ang))
    for m, sirt, cgls in (("sq", A.cyl_sirt_SQ, A.cyl_cgls_SQ), ("kb", A.cyl_sirt_KB, A.cyl_cgls_KB)):
        for lam in (2e-3, 1.0):
            xv = A.ArecImage()
            st = sirt(None, img, ang, xv, lam, 15, 1e-4)
            assert st == G[f"{tag}_{m}_sirt_{lam:g}_status"]
            assert relnorm(xv.data, G[f"{tag}_{m}_sirt_{lam:g}"]) <= 1e-4
        xv = A.ArecImage()
        st = cgls(None, img, ang, xv, 0.0, 8, 1e-6)
        assert st == G[f"{tag}_{m}_cgls_status"]  # KB with even nx: -1 (F8)
        if st == 0:  # CGLS amplifies FP32 rounding: bound by the FP32-vs-FP64 spread
            _, x64, _ = O.cgls(m, b, ang, 8, 1e-6, dtype=np.float64)
            ref = G[f"{tag}_{m}_cgls"]
            assert relnorm(xv.data, ref) <= max(1e-4, 4 * relnorm(ref, x64))


def test_graph_and_eager_agree(ctx, monkeypatch):
    """Each pass is replayed from a captured hipGraph; replays match eager
    launches exactly (same kernels, same order)."""
    import arec3d_amd as A

    nx, ny, na = 65, 20, 40
    b, ang = problem("sq", nx, ny, na, 8)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    db = ctx.to_device(b)
    x1, x2 = ctx.empty(g.vol_shape), ctx.empty(g.vol_shape)
    h1, _ = g.sirt(db, x1, 1.0 / (na * nx), 25, 1e-9)
    monkeypatch.setenv("AREC_NO_GRAPH", "1")
    h2, _ = g.sirt(db, x2, 1.0 / (na * nx), 25, 1e-9)
    assert h1 == h2
    assert np.array_equal(x1.download(), x2.download())


def test_sirt_config1_shape_rows_vs_oracle(ctx):
    """The CPU reference case at its own shape: 128 wide, 64 rows, 180 angles,
    12 seeded ellipsoids (support 60, semi-axes U[4,30], density
    U[0.1,1]), 1% noise, 50 SIRT iterations from zero.  With a fixed lam, no
    restart and no tolerance stop, SIRT is separable by rows, so the oracle
    solved on 4 of the rows must give the GPU volume's same 4 rows (1e-4
    relative: FP32 summation order over 50 iterations)."""
    import arec3d_amd as A
    from arec3d_amd import phantom as P
    from arec3d_amd import unet as U

    print("\n[case] sirt 128x64x180, 50 iterations, rows vs oracle", flush=True)
    nx, ny, na = 128, 64, 180
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    prm = P.ellipsoid_params(1, 12, 60.0, ny, axes=(4.0, 30.0), dens=(0.1, 1.0))
    xt = ctx.empty(g.vol_shape)
    P.phantom_device(ctx, xt, g.radius, ny, 60.0, prm)
    b = ctx.empty(g.proj_shape)
    g.project(xt, b)
    P.add_noise(ctx, b, 1, rel_sigma=0.01)
    lam = 0.9 / U.sigma_max2(ctx, g)
    x = ctx.empty(g.vol_shape)
    hg, _ = g.sirt(b, x, lam, 50, 0.0)
    assert len(hg) == 50 and all(h[1] == 0 for h in hg)
    r0 = 30
    bsub = np.ascontiguousarray(b.download()[:, r0:r0 + 4, :])
    st, xo, ho, lo = O.sirt("sq", bsub, ang, lam, 50, 0.0)
    assert st == 0 and all(h[1] == 0 for h in ho)
    assert relnorm(x.download()[:, r0:r0 + 4], xo) <= 1e-4


@pytest.mark.parametrize("design", [2, 3, 4])
@pytest.mark.parametrize("method", ["sq", "kb", "linear"])
def test_solvers_with_depth_strip_forward_vs_oracle(ctx, method, design):
    """SIRT and CGLS with forward designs 2, 3 and 4 (depth strips, angle
    pairs, staged strips; their kernels also deliver the residual-norm
    partials): the same bar against the oracle as with the sweep."""
    import arec3d_amd as A

    nx, ny, na = 65, 40, 61 if design == 3 else 60
    print(f"\n[case] solvers with depth strips design {design} {method}", flush=True)
    b, ang = problem(method, nx, ny, na, 1)
    lam = 1.0 / (na * nx)
    _, xo, ho, lo = O.sirt(method, b, ang, lam, 30, 1e-6)
    _, _, ho64, _ = O.sirt(method, b, ang, lam, 30, 1e-6, dtype=np.float64)
    g = A.Geometry(ctx, method, nx, ny, ang)
    g.set_forward(design, 0)
    assert g.forward_design == design
    db, dx = ctx.to_device(b), ctx.empty(g.vol_shape)
    hg, lg = g.sirt(db, dx, lam, 30, 1e-6)
    check_hist(hg, ho, ho64)
    assert lg == lo
    assert relnorm(dx.download(), xo) <= 1e-4
    _, xc, hc = O.cgls(method, b, ang, 15, 1e-6)
    _, xc64, hc64 = O.cgls(method, b, ang, 15, 1e-6, dtype=np.float64)
    dx = ctx.empty(g.vol_shape)
    hg = g.cgls(db, dx, 15, 1e-6)
    check_hist(hg, hc, hc64)
    assert relnorm(dx.download(), xc) <= max(1e-4, 4 * relnorm(xc, xc64))


def test_autotune_designs(ctx, tmp_path):
    """The autotuner times every forward design and back-projector tile,
    sets the fastest, caches them and applies the cached plan on the next
    geometry of the same shape; selecting design 1 releases the depth-strip
    table."""
    import arec3d_amd as A
    from arec3d_amd.autotune import autotune, candidates

    print("\n[case] autotune designs", flush=True)
    nx, ny, na = 96, 64, 90
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, "sq", nx, ny, ang)
    g.set_forward(1, 4, 0)
    base = g.device_bytes()
    assert {c[0] for c in candidates(g)} == {1, 2, 3, 4}
    cache = str(tmp_path / "plans.json")
    d, r, m, ms, tile, bms = autotune(ctx, g, cache=cache, reps=1)
    assert g.forward_design == d and g.get_plan()[0] == r and ms > 0
    assert g.back_tile == tile and bms > 0
    g2 = A.Geometry(ctx, "sq", nx, ny, ang)
    assert g2.back_tile == (32, 2)
    assert autotune(ctx, g2, cache=cache)[:5] == (d, r, m, ms, tile)
    assert g2.forward_design == d and g2.back_tile == tile
    g2.set_forward(2, 4)
    assert g2.device_bytes() > base
    g2.set_forward(1, 4, 0)
    assert g2.device_bytes() == base
    rng = np.random.default_rng(3)
    x = rng.random(g.vol_shape).astype(np.float32)
    p = ctx.empty(g.proj_shape)
    g.project(ctx.to_device(x), p)
    op = O.project("sq", x, ang, nx, ny)
    assert np.abs(p.download() - op).max() <= 1e-5 * np.abs(op).max()
