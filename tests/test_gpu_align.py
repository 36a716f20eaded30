"""GPU parity of the alignment kernels (FFT cross-correlation, sub-pixel peak,
Fourier shift) against the FP64 oracle.

FFTW, which the reference uses, is not present; the oracle is pinned by the
cross-correlation theorem and brute force (tests/test_oracle.py).
Tolerances: correlation maps 1e-5 relative to their max (FP32 FFTs);
recovered shifts 1e-3 px of the oracle's; shifted views 1e-5 relative to
their max.
"""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def smooth_views(rng, nv, ny, nx):
    yy, xx = np.mgrid[0:ny, 0:nx]
    out = np.zeros((nv, ny, nx), np.float32)
    for v in range(nv):
        for _ in range(4):
            cx, cy = rng.uniform(0.2 * nx, 0.8 * nx), rng.uniform(0.2 * ny, 0.8 * ny)
            s = rng.uniform(2, 0.15 * min(nx, ny) + 2)
            out[v] += np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s)).astype(np.float32)
    return out


@pytest.mark.parametrize("nx,ny", [(64, 32), (33, 20), (128, 128)])
def test_ccorr_maps_vs_oracle(ctx, nx, ny):
    import arec3d_amd as A

    print(f"\n[case] ccorr {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx + ny)
    nv = 3
    x = rng.random((nv, ny, nx)).astype(np.float32)
    y = rng.random((nv, ny, nx)).astype(np.float32)
    dx, dy, dc = ctx.to_device(x), ctx.to_device(y), ctx.empty((nv, ny, nx))
    A.align_ccorr(ctx, dx, dy, dc)
    c = dc.download()
    for v in range(nv):
        co = O.ccorr2d(x[v], y[v])
        assert np.abs(c[v] - co).max() <= 1e-5 * np.abs(co).max()


@pytest.mark.parametrize("nx,ny", [(64, 64), (128, 64), (45, 30)])
def test_cc_and_shift_vs_oracle(ctx, nx, ny):
    import arec3d_amd as A

    print(f"\n[case] cc+shift {nx}x{ny}", flush=True)
    rng = np.random.default_rng(nx * ny)
    nv = 5
    proj = smooth_views(rng, nv, ny, nx)
    true = rng.uniform(-5, 5, size=(nv, 2)).astype(np.float32)
    meas = np.stack([O.fshift2d(proj[v], true[v, 0], true[v, 1]) for v in range(nv)])
    dp, dm = ctx.to_device(proj), ctx.to_device(meas)
    sx, sy = ctx.empty(nv), ctx.empty(nv)
    A.align_cc(ctx, dp, dm, sx, sy)
    gx, gy = sx.download(), sy.download()
    ox, oy = O.cc_images(proj, meas)
    np.testing.assert_allclose(gx, ox, atol=1e-3)
    np.testing.assert_allclose(gy, oy, atol=1e-3)
    np.testing.assert_allclose(gx, true[:, 0], atol=0.1)
    np.testing.assert_allclose(gy, true[:, 1], atol=0.1)
    # shift back: arecShiftImages semantics (by -s)
    A.align_shift(ctx, dm, sx, sy, -1.0)
    got = dm.download()
    ref = O.shift_images(meas, gx, gy)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max() + 1e-6


def test_speak_on_gpu(ctx):
    """Scan order, tie-break, signed-pmax takeover and wrap, map by map."""
    import arec3d_amd as A

    rng = np.random.default_rng(9)
    maps = []
    d = np.zeros((8, 16), np.float32)
    d[2, 5] = d[1, 7] = 1.0  # tie: smaller x wins
    maps.append(d)
    d = np.zeros((8, 16), np.float32)
    d[3, 2] = -4.0  # negative takeover
    maps.append(d)
    d = np.zeros((8, 16), np.float32)
    d[7, 15] = 3.0  # wraps to (-1, -1)
    d[6, 15] = 1.0
    maps.append(d)
    maps.append(rng.standard_normal((8, 16)).astype(np.float32))
    maps.append(np.maximum(0, 50 - (np.mgrid[0:8, 0:16][1] - 11.3) ** 2 - 2 * (np.mgrid[0:8, 0:16][0] - 5.6) ** 2).astype(np.float32))
    m = np.stack(maps)
    dm = ctx.to_device(m)
    sx, sy = ctx.empty(len(maps)), ctx.empty(len(maps))
    A.align_peak(ctx, dm, sx, sy)
    gx, gy = sx.download(), sy.download()
    for v in range(len(maps)):
        ox, oy = O.speak(m[v])
        assert gx[v] == pytest.approx(ox, abs=1e-6), v
        assert gy[v] == pytest.approx(oy, abs=1e-6), v


def test_host_dropin_cc_shift(ctx):
    import arec3d_amd as A

    rng = np.random.default_rng(4)
    nv, ny, nx = 3, 32, 64
    proj = smooth_views(rng, nv, ny, nx)
    meas = np.stack([O.fshift2d(proj[v], 1.5, -2.25) for v in range(nv)])
    i1 = A.ArecImage(proj.copy(), nx, ny, nv)
    i2 = A.ArecImage(meas.copy(), nx, ny, nv)
    sx, sy = np.zeros(nv, np.float32), np.zeros(nv, np.float32)
    assert A.arecCCImages(None, i1, i2, sx, sy) == 0
    np.testing.assert_allclose(sx, 1.5, atol=0.1)
    np.testing.assert_allclose(sy, -2.25, atol=0.1)
    A.arecShiftImages(i2, sx, sy)
    assert np.abs(i2.data - proj).max() < 0.05
    bad = A.ArecImage(meas[:2].copy(), nx, ny, 2)
    assert A.arecCCImages(None, i1, bad, sx, sy) == -1


@pytest.mark.parametrize("method,nx,ny,na", [("sq", 64, 9, 45), ("kb", 33, 5, 30), ("sq", 200, 3, 7)])
def test_centroid_x_shifts_vs_oracle(ctx, method, nx, ny, na):
    """arec_align_centroid (FP64 device moments, host fit) against the FP64
    numpy restatement on the same seeded, Fourier-shifted views."""
    import arec3d_amd as A

    print(f"\n[case] centroid x shifts {method} {nx}x{ny}x{na}", flush=True)
    rng = np.random.default_rng(nx + ny + na)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    g = A.Geometry(ctx, method, nx, ny, ang)
    r = 0.5 * (nx - 1)
    cord = O.cyl_geometry(r)
    inner = ((cord[:, 0] - r) ** 2 + (cord[:, 1] - r) ** 2) < (0.6 * r) ** 2
    x = (rng.random(g.vol_shape) * inner[:, None]).astype(np.float32)
    p = ctx.empty(g.proj_shape)
    g.project(ctx.to_device(x), p)
    dx = rng.uniform(-2, 2, na).astype(np.float32)
    A.align_shift(ctx, p, ctx.to_device(dx), ctx.to_device(np.zeros(na, np.float32)), sign=1.0)
    views = p.download()
    s = A.align_centroid(ctx, g, p)
    so = O.centroid_x_shifts(method, views, ang, nx, ny)
    np.testing.assert_allclose(s, so, atol=1e-4)  # FP32 device projection of the axis cylinder


def test_take_log_vs_oracle(ctx):
    """arec_take_log against the oracle's logf restatement: within 2 ulp
    (the device and host logf differ by at most that), zeros and negatives
    exact."""
    import arec3d_amd as A

    print("\n[case] take log", flush=True)
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(1e-6, 10.0, 4000), [0.0, -1.0, 1.0, -0.0]]).astype(np.float32)
    d = ctx.to_device(x)
    A.take_log(ctx, d)
    g, o = d.download(), O.take_log(x)
    ulp = np.spacing(np.abs(o).astype(np.float32))
    assert np.all(np.abs(g - o) <= 2 * ulp + 1e-12)
    assert np.array_equal(g[-4:], o[-4:])
    # host-pointer version (the drop-in's arecImageTakeLog)
    h = x.copy()
    assert A.lib.arec_h_take_log(h.ctypes.data, h.size) == 0
    np.testing.assert_array_equal(h, g)


@pytest.mark.parametrize("nz,ny,nx", [(3, 32, 64), (2, 17, 30), (1, 1, 9), (2, 8, 1)])
def test_periodic_decomposition_vs_oracle(ctx, nz, ny, nx):
    """arec_periodic_decomposition (FP32 FFTs) against the oracle (FP64
    DFTs) on seeded views: within 2e-5 of the views' scale."""
    import arec3d_amd as A

    print(f"\n[case] periodic decomposition {nz}x{ny}x{nx}", flush=True)
    rng = np.random.default_rng(nx + 7 * ny + nz)
    u = (rng.random((nz, ny, nx)) * 3.0 + np.linspace(0, 2, nx)[None, None, :]).astype(np.float32)
    d = ctx.to_device(u)
    A.periodic_decomposition(ctx, d)
    g, o = d.download(), O.periodic_decomposition(u)
    assert np.abs(g - o).max() <= 2e-5 * np.abs(u).max()
    h = u.copy()
    assert A.lib.arec_h_periodic_decomposition(nx, ny, nz, h.ctypes.data) == 0
    np.testing.assert_array_equal(h, g)
