"""Pins of the oracle's rotational alignment (arecRotCCImages,
arecRotateImages_skew, arecRotateImages_skew_safe and their helpers).  The
reference's sources for these need FFTW (absent here), so no reference vector
exists; each function is held to what the mathematics fixes: the correlation
theorem, the least-squares parabola, exact interpolation of linear images,
circular shifts for integer shears, rotation of a point about the stated
centre, the Pearson coefficient and SciPy's Tukey window."""
import numpy as np
import pytest
from scipy.signal.windows import tukey

import pyoracle as O


def _blob(nx, ny, cx, cy, s=4.0):
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    return np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s)).astype(np.float32)


def _centroid(im):
    ny, nx = im.shape
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    w = im.astype(np.float64)
    return (w * xx).sum() / w.sum(), (w * yy).sum() / w.sum()


@pytest.mark.parametrize("n", [7, 12, 251])
def test_ccorr1d_is_the_correlation_theorem(n):
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n).astype(np.float32), rng.standard_normal(n).astype(np.float32)
    want = n * np.fft.ifft(np.conj(np.fft.fft(x.astype(np.float64))) * np.fft.fft(y.astype(np.float64))).real
    np.testing.assert_allclose(O.ccorr1d(x, y), want, rtol=1e-10, atol=1e-9 * n)
    # y = x circularly delayed by 5: the peak sits at lag 5
    assert int(np.argmax(O.ccorr1d(x, np.roll(x, 5)))) == 5


@pytest.mark.parametrize("peak,frac", [(10, 0.0), (10, 0.37), (10, -0.8), (0, 0.25), (49, -0.45)])
def test_findpeak1d_parabola_vertex(peak, frac):
    n = 50
    i = np.arange(n, dtype=np.float64)
    d = (i - peak - frac + n / 2) % n - n / 2  # circular distance
    c = (100.0 - 0.5 * d * d).astype(np.float32)
    got = O.findpeak1d(c)
    assert abs(((got - (peak + frac)) + n / 2) % n - n / 2) < 2e-5


def test_findpeak1d_is_least_squares_fit():
    rng = np.random.default_rng(3)
    c = rng.standard_normal(30).astype(np.float32)
    c[12] = 10.0
    k = np.arange(-3, 4)
    a2, a1, _ = np.polyfit(k, c[12 + k].astype(np.float64), 2)
    assert abs(O.findpeak1d(c) - (12 - a1 / (2 * a2))) < 1e-5
    # first of equal maxima
    e = np.zeros(20, np.float32)
    e[[4, 9]] = 1.0
    assert abs(O.findpeak1d(e) - 4.0) < 1e-6


def test_radial_profile_closed_forms():
    nx, ny, R, L = 40, 34, 12, 75
    # constant image: sum of r
    const = np.full((ny, nx), 2.5, np.float32)
    np.testing.assert_allclose(O.radial_profile(const, L, R), 2.5 * R * (R - 1) / 2, rtol=1e-6)
    # linear images are interpolated exactly: x uses cos, y uses sin, about ((nx-1)/2, (ny-1)/2)
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float32)
    a = 2 * np.pi * np.arange(L) / L
    r = np.arange(R)
    xs = ((nx - 1) / 2 + np.cos(a)[:, None] * r).astype(np.float32)
    ys = ((ny - 1) / 2 + np.sin(a)[:, None] * r).astype(np.float32)
    np.testing.assert_allclose(O.radial_profile(xx, L, R), (xs * r).sum(1), rtol=1e-5)
    np.testing.assert_allclose(O.radial_profile(yy, L, R), (ys * r).sum(1), rtol=1e-5)


@pytest.mark.parametrize("deg", [6.0, -4.5, 2.0])
def test_rot_cc_recovers_rotation(deg):
    nx, ny = 80, 96
    xc, yc = (nx - 1) / 2, (ny - 1) / 2
    t = np.deg2rad(deg)

    def scene(th):
        pts = [(20.0, 0.0, 4.0, 1.0), (-5.0, 15.0, 3.0, 0.7), (-12.0, -20.0, 3.5, 0.5)]
        im = np.zeros((ny, nx), np.float32)
        for px, py, s, w in pts:
            im += w * _blob(nx, ny, xc + px * np.cos(th) - py * np.sin(th), yc + px * np.sin(th) + py * np.cos(th), s)
        return im

    # the second stack is the first turned by +deg (x towards y): the angle is +deg
    ang = O.rot_cc_images(scene(0.0)[None], scene(t)[None], O.rot_r2(nx, ny))
    assert abs(ang[0] - deg) < 0.15


def test_rot_cc_drops_large_angles():
    nx, ny = 64, 80
    xc, yc = (nx - 1) / 2, (ny - 1) / 2
    a = _blob(nx, ny, xc + 18, yc)
    b = _blob(nx, ny, xc, yc + 18)  # a quarter turn
    assert O.rot_cc_images(a[None], b[None], O.rot_r2(nx, ny))[0] == 0.0


@pytest.mark.parametrize("nx,ny", [(16, 11), (9, 14)])
def test_integer_shears_are_circular_shifts(nx, ny):
    rng = np.random.default_rng(nx)
    im = rng.standard_normal((ny, nx)).astype(np.float32)
    cy, cx = (ny - 1) / 2, (nx - 1) / 2
    if ny % 2:  # integer skews for every row: row y moves by y - cy along x
        out = O.shear_x(im, 1.0)
        for y in range(ny):
            np.testing.assert_array_equal(out[y], np.roll(im[y], int(y - cy)))
    if nx % 2:
        out = O.shear_y(im, -2.0)
        for x in range(nx):
            np.testing.assert_array_equal(out[:, x], np.roll(im[:, x], int(-2 * (x - cx))))


def test_fractional_shear_is_linear_interpolation():
    nx, ny = 32, 9
    im = np.tile(np.arange(nx, dtype=np.float32), (ny, 1))
    s = 0.3
    out = O.shear_x(im, s)
    for y in range(ny):
        k = (y - (ny - 1) / 2) * s
        x = np.arange(nx)
        src = x - k  # the row moves by +k: out(x) = in(x - k), linear between samples
        inner = (src >= 1) & (src <= nx - 2)
        np.testing.assert_allclose(out[y, inner], src[inner], atol=2e-5)
    # constant along the shear direction: unchanged
    const = np.tile(np.linspace(-1, 1, ny, dtype=np.float32)[:, None], (1, nx))
    np.testing.assert_allclose(O.shear_x(const, 0.41), const, atol=1e-6)


def test_rotate2d_point_rotation_and_identity():
    nx, ny = 48, 40
    rng = np.random.default_rng(5)
    im = rng.standard_normal((ny, nx)).astype(np.float32)
    out = O.rotate2d(im, 0.0)
    np.testing.assert_array_equal(out[:-1, :-1], im[:-1, :-1])
    assert not out[-1].any() and not out[:, -1].any()
    # a blob at (+12, 0) from the centre (nx/2, ny/2) turns by -deg (x towards -y)
    deg = 30.0
    b = _blob(nx, ny, nx / 2 + 12, ny / 2, 2.0)
    cx, cy = _centroid(O.rotate2d(b, deg))
    t = np.deg2rad(deg)
    assert abs(cx - (nx / 2 + 12 * np.cos(t))) < 0.05 and abs(cy - (ny / 2 - 12 * np.sin(t))) < 0.05


@pytest.mark.parametrize("deg", [7.0, -3.0])
def test_three_shears_rotate_like_rotate2d(deg):
    nx, ny = 80, 96
    xc, yc = (nx - 1) / 2, (ny - 1) / 2
    b = _blob(nx, ny, xc + 20, yc + 6, 3.0)
    out, ang = O.rotate_skew(b[None], [deg])
    assert ang[0] == np.float32(deg)
    cx, cy = _centroid(out[0])
    t = -np.deg2rad(deg)  # x towards -y for a positive angle, as rotate2d
    assert abs(cx - (xc + 20 * np.cos(t) - 6 * np.sin(t))) < 0.05
    assert abs(cy - (yc + 20 * np.sin(t) + 6 * np.cos(t))) < 0.05


def test_rotate_skew_branches():
    nx, ny = 24, 20
    rng = np.random.default_rng(9)
    st = rng.standard_normal((5, ny, nx)).astype(np.float32)
    out, ang = O.rotate_skew(st, [350.0, 200.0, 30.0, -60.0, 4.0])
    # 350 wraps to -10 (three shears); 200 wraps to -160 and, like -60, is dropped; 30 is bilinear
    np.testing.assert_array_equal(ang, np.float32([-10.0, 0.0, 30.0, 0.0, 4.0]))
    np.testing.assert_array_equal(out[1], st[1])
    np.testing.assert_array_equal(out[3], st[3])
    np.testing.assert_array_equal(out[2], O.rotate2d(st[2], 30.0))
    a, b = np.float32(-np.tan(-0.5 * np.float32(np.deg2rad(4.0)))), np.sin(-np.float32(4.0 * np.pi / 180))
    np.testing.assert_array_equal(out[4], O.shear_x(O.shear_y(O.shear_x(st[4], a), b), a))


def test_tukey_ncc_is_windowed_pearson():
    nx, ny = 37, 29
    rng = np.random.default_rng(11)
    a = rng.standard_normal((ny, nx)).astype(np.float32)
    b = (0.6 * a + rng.standard_normal((ny, nx))).astype(np.float32)
    w = (np.outer(tukey(ny, 0.3), tukey(nx, 0.3))).astype(np.float32)
    want = np.corrcoef((a * w).ravel().astype(np.float64), (b * w).ravel().astype(np.float64))[0, 1]
    assert abs(O.tukey_ncc(a, b) - want) < 1e-6
    assert abs(O.tukey_ncc(a, a) - 1.0) < 1e-12 and abs(O.tukey_ncc(a, -3 * a) + 1.0) < 1e-12


def test_rotate_skew_safe_keeps_improvements_only():
    nx, ny = 80, 96
    xc, yc = (nx - 1) / 2, (ny - 1) / 2
    ref = _blob(nx, ny, xc + 20, yc + 4, 3.0) + 0.5 * _blob(nx, ny, xc - 8, yc - 18, 3.0)
    t = np.deg2rad(5.0)  # the measured view, turned by +5 degrees (x towards y)
    view = (_blob(nx, ny, xc + 20 * np.cos(t) - 4 * np.sin(t), yc + 20 * np.sin(t) + 4 * np.cos(t), 3.0)
            + 0.5 * _blob(nx, ny, xc - 8 * np.cos(t) + 18 * np.sin(t), yc - 8 * np.sin(t) - 18 * np.cos(t), 3.0))
    st = np.stack([view, view])
    refs = np.stack([ref, ref])
    out, ang, kept, mean = O.rotate_skew_safe(st, refs, [5.0, -5.0])
    # +5 turns it back (kept, angle halved); -5 turns it further (restored, angle 0)
    assert kept == 1 and ang[0] == np.float32(2.5) and ang[1] == 0.0
    np.testing.assert_array_equal(out[1], view)
    assert O.tukey_ncc(out[0], ref) > O.tukey_ncc(view, ref) + 1e-3
    assert abs(mean - 0.5 * (O.tukey_ncc(out[0], ref) + O.tukey_ncc(view, ref))) < 1e-12
