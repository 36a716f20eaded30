"""Rotational alignment on the GPU (arecRotCCImages, arecRotateImages,
arecRotateImages_skew, arecRotateImages_skew_safe; arecImage.cpp:555-794)
against the oracle's restatement.

The device takes every transcendental value from the host in the
reference's precision and repeats the reference's float and double
operations in order, so rotations and angles equal the oracle's bit for bit;
the windowed correlation coefficients (parallel double sums) agree to 1e-12.
Recovery: views of a phantom turned by known angles come back within 0.2
degrees from the radial-profile correlation against unturned reprojections.
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def _scene(nx, ny, th, seed=0):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    xc, yc = (nx - 1) / 2, (ny - 1) / 2
    im = np.zeros((ny, nx))
    for _ in range(6):
        px, py = rng.uniform(-0.3, 0.3) * nx, rng.uniform(-0.3, 0.3) * ny
        s, w = rng.uniform(2.0, 5.0), rng.uniform(0.3, 1.0)
        qx, qy = xc + px * np.cos(th) - py * np.sin(th), yc + px * np.sin(th) + py * np.cos(th)
        im += w * np.exp(-((xx - qx) ** 2 + (yy - qy) ** 2) / (2 * s * s))
    return im.astype(np.float32)


def _stack(nx, ny, degs, seed=0, noise=0.0):
    rng = np.random.default_rng(seed + 100)
    st = np.stack([_scene(nx, ny, np.deg2rad(d), seed) for d in degs])
    return (st + noise * rng.standard_normal(st.shape)).astype(np.float32)


@pytest.mark.parametrize("nx,ny", [(80, 96), (65, 40)])
def test_rot_cc_vs_oracle(ctx, nx, ny):
    import arec3d_amd as A

    print(f"\n[case] rot_cc {nx}x{ny}", flush=True)
    degs = [0.0, 3.0, -6.5, 12.0, -30.0, 1.25]
    ref = _stack(nx, ny, [0.0] * len(degs), seed=1)
    mov = _stack(nx, ny, degs, seed=1, noise=0.02)
    r2 = A.rot_r2(nx, ny)
    got = A.rot_cc(ctx, ctx.to_device(ref), ctx.to_device(mov), r2)
    want = O.rot_cc_images(ref, mov, r2)
    np.testing.assert_array_equal(got, want)
    # larger r2 than the driver's, up to min(nx, ny)/2
    r2b = min(nx, ny) // 2
    np.testing.assert_array_equal(A.rot_cc(ctx, ctx.to_device(ref), ctx.to_device(mov), r2b),
                                  O.rot_cc_images(ref, mov, r2b))


def test_rot_cc_recovers_angles(ctx):
    import arec3d_amd as A

    print("\n[case] rot_cc recovery", flush=True)
    nx, ny = 128, 120
    degs = np.array([2.0, -4.0, 7.5, -1.0, 0.0, 5.0, -9.0, 3.5], np.float32)
    ref = _stack(nx, ny, np.zeros(len(degs)), seed=4)
    mov = _stack(nx, ny, degs, seed=4, noise=0.01)
    got = A.rot_cc(ctx, ctx.to_device(ref), ctx.to_device(mov), A.rot_r2(nx, ny))
    assert np.abs(got - degs).max() < 0.2


@pytest.mark.parametrize("rule", ["skew", "bilinear"])
def test_rotate_views_vs_oracle(ctx, rule):
    import arec3d_amd as A

    print(f"\n[case] rotate_views {rule}", flush=True)
    nx, ny = 70, 52
    rng = np.random.default_rng(3)
    angles = np.float32([350.0, 200.0, 30.0, -60.0, 4.0, -7.5, 0.0, 12.0, -185.0, 10.0])
    st = rng.standard_normal((len(angles), ny, nx)).astype(np.float32)
    d = ctx.to_device(st)
    got_ang = A.rotate_views(ctx, d, angles, rule)
    got = d.download()
    if rule == "skew":
        want, want_ang = O.rotate_skew(st, angles)
        np.testing.assert_array_equal(got_ang, want_ang)
    else:
        want = np.stack([O.rotate2d(st[v], angles[v]) for v in range(len(angles))])
        np.testing.assert_array_equal(got_ang, angles)
    np.testing.assert_array_equal(got, want)


def test_rotate_views_safe_vs_oracle(ctx):
    import arec3d_amd as A

    print("\n[case] rotate_views_safe", flush=True)
    nx, ny = 80, 96
    truth = [5.0, -3.0, 2.0, 0.0, -8.0, 6.0]
    ref = _stack(nx, ny, [0.0] * len(truth), seed=2)
    view = _stack(nx, ny, truth, seed=2, noise=0.01)
    angles = np.float32([5.0, 3.0, 2.0, 0.0, -8.0, -20.0])  # views 1 and 5 turn the wrong way
    d = ctx.to_device(view)
    ang, kept, mcc = A.rotate_views_safe(ctx, d, ctx.to_device(ref), angles)
    w, want_ang, want_kept, want_mcc = O.rotate_skew_safe(view, ref, angles)
    np.testing.assert_array_equal(ang, want_ang)
    assert kept == want_kept
    assert abs(mcc - want_mcc) <= 1e-12 * abs(want_mcc)
    np.testing.assert_array_equal(d.download(), w)
    assert ang[1] == 0.0 and ang[5] == 0.0 and ang[0] == 2.5 and ang[4] == -4.0


def test_host_versions(ctx):
    import arec3d_amd as A

    print("\n[case] rotation host entry points", flush=True)
    nx, ny = 40, 36
    ref = _stack(nx, ny, [0.0, 0.0, 0.0], seed=5)
    mov = _stack(nx, ny, [4.0, -2.0, 9.0], seed=5)
    r2 = A.rot_r2(nx, ny)
    ang = np.zeros(3, np.float32)
    assert A.lib.arec_h_rot_cc_images(nx, ny, 3, A._ptr(ref), nx, ny, 3, A._ptr(mov), r2, A._ptr(ang)) == 0
    np.testing.assert_array_equal(ang, O.rot_cc_images(ref, mov, r2))
    assert A.lib.arec_h_rot_cc_images(nx, ny, 3, A._ptr(ref), nx, ny, 2, A._ptr(mov), r2, A._ptr(ang)) == -1
    data = mov.copy()
    a = np.float32([15.0, -3.0, 50.0])
    assert A.lib.arec_h_rotate_images(nx, ny, 3, A._ptr(data), A._ptr(a), 0) == 0
    want, want_a = O.rotate_skew(mov, [15.0, -3.0, 50.0])
    np.testing.assert_array_equal(data, want)
    np.testing.assert_array_equal(a, want_a)
    data = mov.copy()
    a = np.float32([4.0, 2.0, 9.0])
    kept, mcc = A.C.c_int32(0), A.C.c_double(0.0)
    assert A.lib.arec_h_rotate_images_safe(nx, ny, 3, A._ptr(data), A._ptr(ref), A._ptr(a), A.C.byref(kept),
                                           A.C.byref(mcc)) == 0
    w, wa, wk, wm = O.rotate_skew_safe(mov, ref, [4.0, 2.0, 9.0])
    np.testing.assert_array_equal(data, w)
    np.testing.assert_array_equal(a, wa)
    assert kept.value == wk


def test_argument_checks(ctx):
    import arec3d_amd as A

    d = ctx.to_device(np.zeros((2, 30, 40), np.float32))
    with pytest.raises(A.ArecError):
        A.rot_cc(ctx, d, d, 16)  # 2 r2 > min(nx, ny)
    with pytest.raises(A.ArecError):
        A.rot_cc(ctx, d, d, 0)
    with pytest.raises(A.ArecError):
        A.rotate_views_safe(ctx, d, d, [120.0, 0.0])
