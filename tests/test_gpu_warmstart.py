"""What the trained warm-start network gains over SIRT without it.

On held-out seeded phantoms at the training size (128^3 cylinder, 360
angles, SQ, 1% noise; phantom seeds disjoint from the training seeds), SIRT
from zero and the warm start (5 SIRT iterations, the network, SIRT from its
output) both run to relnrm <= 1e-3 with the same fixed denominator
||A^T b||.  SIRT's residual check must accept the network, which must cut the
projector passes to the tolerance by at least 2x.  A start whose residual is
not below the bound is rejected with x untouched, and the solver
falls back to the SIRT iterate.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(ctx):
    import arec3d_amd as A
    from arec3d_amd import phantom as P
    from arec3d_amd import unet as U

    n, na = 128, 360
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, "sq", n, n, ang)
    lam = 0.9 / U.sigma_max2(ctx, g)
    ws = U.WarmStart(ctx, g)
    return A, P, U, g, lam, ws


@pytest.mark.parametrize("seed", [49500, 49501])
def test_warm_start_beats_plain_sirt(ctx, setup, seed):
    A, P, U, g, lam, ws = setup
    print(f"\n[case] warm start 128^3/360 seed {seed}", flush=True)
    prm, sup = U.phantom_params(seed, g.nx, g.ny)
    xt = ctx.empty(g.vol_shape)
    P.phantom_device(ctx, xt, g.radius, g.ny, sup, prm)
    b = ctx.empty(g.proj_shape)
    g.project(xt, b)
    P.add_noise(ctx, b, seed, rel_sigma=0.01)
    tol = 1e-3
    x = ctx.empty(g.vol_shape)
    hp, _ = g.sirt(b, x, lam, 3000, tol, hist_cap=3100)
    assert hp[-1][1] == 2 and hp[-1][2] < tol
    plain_passes = len(hp)
    (h1, h2), _, chosen = U.warm_sirt(ctx, g, b, x, lam, tol, ws)
    assert chosen == "network"
    assert h2[-1][1] == 2 and h2[-1][2] < tol
    warm_passes = len(h1) + len(h2)
    print(f"passes to {tol:g}: plain {plain_passes}, warm {warm_passes} "
          f"(network output relnrm {h2[0][2]:.3e} vs 4-iteration SIRT {h1[-1][2]:.3e})")
    assert warm_passes * 2 <= plain_passes


def test_residual_check_rejects_a_bad_start(ctx, setup):
    A, P, U, g, lam, ws = setup
    print("\n[case] warm start residual check", flush=True)
    rng = np.random.default_rng(1)
    b = ctx.to_device(rng.random(g.proj_shape).astype(np.float32))
    bad = (10.0 * rng.standard_normal(g.vol_shape)).astype(np.float32)
    x = ctx.to_device(bad)
    h, _ = g.sirt_warm(b, x, lam, 50, 1e-3, reject_above=0.5)
    assert len(h) == 1 and h[0][1] == 3 and h[0][2] >= 0.5
    assert np.array_equal(x.download(), bad)
    # checked after 5 passes: rejected at the latest there
    h, _ = g.sirt_warm(b, x, lam, 50, 1e-3, reject_above=1e-9, check_pass=5)
    assert 1 <= len(h) <= 5 and h[-1][1] == 3 and all(e[1] == 0 for e in h[:-1])
