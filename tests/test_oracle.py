"""CPU tests that pin the oracle (no GPU).

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a wrong index, a transposed operand) fails: the reference's own
unit tests (AREC3D/AREC3D_test/AREC3D_test.cpp), golden vectors written by the
reference's own compiled code (tests/golden/ref_vectors.npz), and closed
forms / brute force where the reference holds no pin.
"""
import os

import numpy as np
import pytest

import pyoracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


# ---- geometry: AREC3D_test.cpp:8-70 ---------------------------------------

def test_cylvol_odd():
    c = O.cyl_geometry(1, 2)
    assert len(c) == 5  # nrays, nnz = 10 at height 2
    assert c.tolist() == [[1, 0], [0, 1], [1, 1], [2, 1], [1, 2]]


def test_cylvol_even():
    c = O.cyl_geometry(0.5, 3)
    assert len(c) == 4  # nnz = 12: the truncated mask keeps all 4 voxels
    assert c.tolist() == [[0, 0], [1, 0], [0, 1], [1, 1]]


@pytest.mark.parametrize("r", [1.0, 1.5, 7.0, 63.5, 255.5])
def test_cylvol_brute_force(r):
    n = int(round(2 * r + 1))
    ref = [(ix, iz) for iz in range(n) for ix in range(n)
           if int(ix - r) ** 2 + int(iz - r) ** 2 <= int(r * r)]
    assert O.cyl_geometry(r).tolist() == [list(t) for t in ref]


def test_cylvol_counts_128_and_512():
    # column counts of the 128- and 512-wide cylinders (radius (n-1)/2)
    assert len(O.cyl_geometry(63.5)) == 12900
    assert len(O.cyl_geometry(255.5)) == 206124


def test_cube_round_trip_reference_case():
    """AREC3D_test.cpp:326-363: a radius-1, height-2 cylinder holding 0..9
    scattered into its cube keeps its sum, and gathers back unchanged; the
    rows are the cube's slowest axis and every other cube element is 0."""
    x = np.arange(10, dtype=np.float32).reshape(5, 2)
    cube = O.cyl_to_cube(x, 1.0)
    assert cube.shape == (2, 3, 3) and cube.sum() == x.sum()
    assert np.array_equal(O.cube_to_cyl(cube, 1.0), x)
    # cord (1,0),(0,1),(1,1),(2,1),(1,2): column j at [row][iz][ix]
    assert cube[1, 0, 1] == x[0, 1] and cube[0, 1, 0] == x[1, 0] and cube[1, 2, 1] == x[4, 1]
    assert np.count_nonzero(cube) == 9  # x[0, 0] == 0 is the tenth value
    assert cube[:, 0, 0].tolist() == [0, 0] and cube[:, 2, 2].tolist() == [0, 0]


# ---- SQ known answers: AREC3D_test.cpp:97-189 ------------------------------

def test_sq_projection_odd():
    x = np.zeros((5, 2), np.float32)
    x[2] = [1.0, 2.0]  # data[4], data[5]: centre column
    p = O.project("sq", x, np.array([0.0], np.float32), 3, 2)
    assert p.sum() == pytest.approx(3.0, rel=1e-6)
    assert p.reshape(-1)[1] == pytest.approx(1.0, rel=1e-5)
    assert p.reshape(-1)[4] == pytest.approx(2.0, rel=1e-5)


def test_sq_backprojection_odd():
    p = np.zeros((1, 2, 3), np.float32)
    p.reshape(-1)[1] = 1.0
    p.reshape(-1)[4] = 2.0
    x = O.backproject("sq", p, np.array([0.0], np.float32), 3, 2)
    assert x.sum() == pytest.approx(9.0, rel=1e-5)


def test_sq_weights_closed_form_theta0():
    # theta = 0: a = 0.5, b = 0.5 + 1e-6, the footprint is the unit box and
    # a voxel centred on a pixel gives (~5e-7, ~1-1e-6, ~5e-7).
    x = np.zeros((5, 1), np.float32)
    x[2, 0] = 1.0
    p = O.project("sq", x, np.array([0.0], np.float32), 3, 1, dtype=np.float64)
    assert p[0, 0, 0] == pytest.approx(5e-7, rel=0.01)
    assert p[0, 0, 1] == pytest.approx(1 - 1e-6, abs=1e-9)
    assert p[0, 0, 2] == pytest.approx(5e-7, rel=0.01)


@pytest.mark.parametrize("method", ["sq", "kb"])
def test_sq_kb_weights_sum(method):
    # SQ footprints integrate to 1 at every angle (b <= 0.7072 < 1.5 keeps the
    # footprint inside the 3 taps); KB sums to ~0.978 at zero offset.
    r = 7.0
    n = int(2 * r + 1)
    nr = len(O.cyl_geometry(r))
    ang = np.linspace(0, np.pi, 13, endpoint=False).astype(np.float32)
    cord = O.cyl_geometry(r)
    inner = ((cord[:, 0] - r) ** 2 + (cord[:, 1] - r) ** 2) <= (r - 2) ** 2  # taps stay on the detector
    x = inner.astype(np.float64)[:, None]
    p = O.project(method, x, ang, n, 1, dtype=np.float64)
    sums = p.reshape(len(ang), -1).sum(1)
    if method == "sq":
        np.testing.assert_allclose(sums, inner.sum(), rtol=1e-9)
    else:
        assert np.all(sums > 0.95 * inner.sum()) and np.all(sums < 1.0 * inner.sum())


def _dense(method, nx, ny, ang, dtype=np.float32):
    nr = len(O.cyl_geometry(0.5 * (nx - 1)))
    A = np.zeros((len(ang) * ny * nx, nr * ny), dtype)
    AT = np.zeros_like(A)
    for a in range(nr * ny):
        e = np.zeros(nr * ny, dtype)
        e[a] = 1
        A[:, a] = O.project(method, e.reshape(nr, ny), ang, nx, ny, dtype=dtype).reshape(-1)
    for b in range(A.shape[0]):
        e = np.zeros(A.shape[0], dtype)
        e[b] = 1
        AT[b, :] = O.backproject(method, e.reshape(len(ang), ny, nx), ang, nx, ny, dtype=dtype).reshape(-1)
    return A, AT


# ---- exact adjointness: AREC3D_test.cpp:216-324 ----------------------------

@pytest.mark.parametrize("method,r,ang", [
    ("sq", 1.0, [1.475, 21.123, 2.123]),
    ("sq", 1.5, [0.698132, 1.22173]),
    ("kb", 0.5, [0.698132, 1.22173]),
    ("kb", 1.0, [1.475, 21.123, 2.123]),
    ("linear", 2.0, [0.3, 1.1, 2.5]),
])
def test_impulse_adjoint_exact(method, r, ang):
    nx = int(round(2 * r + 1))
    A, AT = _dense(method, nx, 2, np.array(ang, np.float32))
    assert np.array_equal(A, AT)


# ---- bit-exact against the reference's compiled code -----------------------

@pytest.mark.parametrize("tag,nx,ny,na", [("odd", 33, 6, 24), ("even", 32, 5, 20)])
@pytest.mark.parametrize("method", ["sq", "kb", "linear"])
def test_projectors_vs_golden(gold, tag, nx, ny, na, method):
    ang, x, p = gold[f"{tag}_angles"], gold[f"{tag}_x"], gold[f"{tag}_p"]
    assert np.array_equal(O.cyl_geometry(0.5 * (nx - 1)), gold[f"{tag}_cord"])
    assert np.array_equal(O.project(method, x, ang, nx, ny), gold[f"{tag}_{method}_fwd"])
    bo, br = O.backproject(method, p, ang, nx, ny), gold[f"{tag}_{method}_bwd"]
    if method == "linear":  # last two columns read past the cord table in the reference
        bo, br = bo[:-2], br[:-2]
    assert np.array_equal(bo, br)


@pytest.mark.parametrize("tag", ["odd", "even"])
@pytest.mark.parametrize("method", ["sq", "kb"])
def test_solvers_vs_golden(gold, tag, method):
    ang, b = gold[f"{tag}_angles"], gold[f"{tag}_b"]
    for lam in (2e-3, 1.0):
        st, x, hist, _ = O.sirt(method, b, ang, lam, 15, 1e-4)
        assert st == gold[f"{tag}_{method}_sirt_{lam:g}_status"]
        assert np.array_equal(x, gold[f"{tag}_{method}_sirt_{lam:g}"])
    st, x, hist = O.cgls(method, b, ang, 8, 1e-6)
    assert st == gold[f"{tag}_{method}_cgls_status"]
    if st == 0:
        assert np.array_equal(x, gold[f"{tag}_{method}_cgls"])


@pytest.mark.skipif(not O.ref_available(), reason="reference build not present (needs /root/reference)")
def test_solvers_vs_reference_build():
    rng = np.random.default_rng(7)
    nx, ny, na = 41, 3, 30
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    x0 = rng.random((len(O.cyl_geometry(20.0)), ny)).astype(np.float32)
    b = O.project("sq", x0, ang, nx, ny)
    st, x, _ = O.cgls("sq", b, ang, 5, 1e-9, x0=0.5 * x0)
    sr, xr = O.ref_solve("cgls", "sq", b, ang, 0.0, 5, 1e-9, x0=0.5 * x0)
    assert st == sr == 0 and np.array_equal(x, xr)


# ---- solver invariants where no reference pin exists -----------------------

def _small_problem(method="sq", nx=15, ny=2, na=12, seed=3):
    rng = np.random.default_rng(seed)
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    nr = len(O.cyl_geometry(0.5 * (nx - 1)))
    xt = rng.random((nr, ny))
    A, _ = _dense(method, nx, ny, ang, np.float64)
    b = A @ xt.reshape(-1) + 0.01 * rng.standard_normal(A.shape[0])
    return A, b.reshape(na, ny, nx), ang, nr


def test_landweber_matches_dense_iteration():
    A, b, ang, nr = _small_problem()
    lam = 0.9 / np.linalg.norm(A, 2) ** 2
    st, x, hist, _ = O.sirt("sq", b, ang, lam, 30, 1e-12, dtype=np.float64)
    xd = np.zeros(A.shape[1])
    atb = A.T @ b.reshape(-1)
    r = None
    for _ in range(30):  # iterations 1..30 each print a residual, then update
        g = atb - A.T @ (A @ xd)
        r = np.linalg.norm(g) / np.linalg.norm(atb)
        xd = xd + lam * g
    np.testing.assert_allclose(x.reshape(-1), xd, rtol=1e-5, atol=1e-9)
    # the residual printed at iteration k is ||A^T(b - A x_{k-1})|| / ||A^T b||
    assert hist[-1][2] == pytest.approx(r, rel=1e-6)


def test_cgls_matches_textbook_and_lstsq():
    A, b, ang, nr = _small_problem(nx=11, ny=1, na=30)
    bb = b.reshape(-1)
    # Bjorck's CGLS in FP64 on the dense matrix: first 4 iterates
    x = np.zeros(A.shape[1])
    s = bb.copy()
    g = A.T @ s
    p = g.copy()
    gam = g @ g
    for k in range(4):
        q = A @ p
        alpha = gam / (q @ q)
        x = x + alpha * p
        s = s - alpha * q
        g = A.T @ s
        gam_new = g @ g
        p = g + (gam_new / gam) * p
        gam = gam_new
    st, xo, hist = O.cgls("sq", b, ang, 4, 1e-13, dtype=np.float64)
    np.testing.assert_allclose(xo.reshape(-1), x, rtol=1e-9, atol=1e-12)
    assert hist[-1][3] == pytest.approx(np.sqrt(gam), rel=1e-9)
    # run to convergence: the normal equations are solved (lstsq)
    st, xc, hist = O.cgls("sq", b, ang, 3 * A.shape[1], 1e-13, dtype=np.float64)
    xl = np.linalg.lstsq(A, bb, rcond=None)[0]
    r = np.linalg.norm(A.T @ (bb - A @ xc.reshape(-1))) / np.linalg.norm(A.T @ bb)
    assert r < 1e-9
    np.testing.assert_allclose(xc.reshape(-1), xl, rtol=1e-3, atol=1e-4)


def test_weighted_cgls_is_cgls_on_the_weighted_system():
    """cyl_cgls*_stat (areccgls_stat.cpp:49-275) keeps s = w.(b - A x)
    (:148-154), divides by sum w q^2 (:197-204) and updates s -= alpha w.q
    (:212).  With W = diag(w) that is textbook CGLS on (sqrt(W) A,
    sqrt(W) b): the same g = A^T W (b - A x), the same alpha and beta.  So
    the first iterates equal that recurrence on the dense FP64 matrix, and
    the converged volume is the weighted least-squares solution.  A weight
    dropped from the residual, the norm or the update fails both checks."""
    A, b, ang, nr = _small_problem(nx=11, ny=1, na=30, seed=8)
    b = np.abs(b) / np.abs(b).max() * 1.5  # path integrals of 0..1.5: w = exp(-b) spans 0.22..1
    w = O.init_weights(b.astype(np.float32)).astype(np.float64).reshape(-1)
    assert w.min() == 0.0 and w.max() > 0.5 and np.ptp(w[w > 0]) > 0.3  # Tukey zeros, exp(-b) spread
    sw = np.sqrt(w)
    Aw, bw = A * sw[:, None], b.reshape(-1) * sw
    x = np.zeros(A.shape[1])
    s = bw.copy()
    g = Aw.T @ s
    p = g.copy()
    gam = g @ g
    gnorms = []
    for k in range(4):
        q = Aw @ p
        alpha = gam / (q @ q)
        x = x + alpha * p
        s = s - alpha * q
        g = Aw.T @ s
        gam_new = g @ g
        p = g + (gam_new / gam) * p
        gam = gam_new
        gnorms.append(np.sqrt(gam))
    assert all(a > c for a, c in zip(gnorms, gnorms[1:]))  # no best-iterate restore at 4 iterations
    st, xo, hist = O.cgls("sq", b, ang, 4, 1e-13, dtype=np.float64, weighted=True)
    assert st == 0
    np.testing.assert_allclose(xo.reshape(-1), x, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose([h[3] for h in hist[1:]], gnorms, rtol=1e-9)
    # the FP32 solver follows the same recurrence to FP32 accuracy
    st, xf, hf = O.cgls("sq", b.astype(np.float32), ang, 4, 1e-13, weighted=True)
    assert st == 0 and np.linalg.norm(xf.reshape(-1) - x) <= 1e-4 * np.linalg.norm(x)
    # run to convergence: the weighted normal equations are solved
    st, xc, hist = O.cgls("sq", b, ang, 3 * A.shape[1], 1e-13, dtype=np.float64, weighted=True)
    xl = np.linalg.lstsq(Aw, bw, rcond=None)[0]
    r = np.linalg.norm(Aw.T @ (bw - Aw @ xc.reshape(-1))) / np.linalg.norm(Aw.T @ bw)
    assert r < 1e-9
    np.testing.assert_allclose(xc.reshape(-1), xl, rtol=1e-3, atol=1e-4)
    # and the unweighted solution is a different point (the weights matter)
    xu = np.linalg.lstsq(A, b.reshape(-1), rcond=None)[0]
    assert np.linalg.norm(xu - xl) > 1e-3 * np.linalg.norm(xl)


def test_sirt_restart_chain():
    A, b, ang, nr = _small_problem()
    st, x, hist, lam = O.sirt("sq", b.astype(np.float32), ang, 1e3, 10, 1e-9)
    restarts = sum(1 for h in hist if h[1] == 1)
    assert restarts >= 1 and lam == pytest.approx(1e3 / 2 ** restarts)


def test_init_weights_tukey_closed_form():
    from scipy.signal.windows import tukey

    nx, ny, na = 21, 2, 3
    b = np.zeros((na, ny, nx), np.float32)
    w = O.init_weights(b)
    np.testing.assert_allclose(w[0, 0], tukey(nx, 0.5), atol=1e-6)
    b[:] = 1.0
    np.testing.assert_allclose(O.init_weights(b)[1, 1], np.exp(-1) * tukey(nx, 0.5), rtol=1e-6)


# ---- alignment: closed forms / brute force ---------------------------------

def test_ccorr2d_brute_force():
    rng = np.random.default_rng(11)
    for nx, ny in [(8, 4), (6, 5), (16, 8)]:
        x = rng.random((ny, nx)).astype(np.float32)
        y = rng.random((ny, nx)).astype(np.float32)
        c = O.ccorr2d(x, y)
        bf = np.zeros((ny, nx))
        for dy in range(ny):
            for dx in range(nx):
                bf[dy, dx] = nx * ny * np.sum(x.astype(np.float64) * np.roll(np.roll(y, -dy, 0), -dx, 1))
        np.testing.assert_allclose(c, bf, rtol=1e-10)


@pytest.mark.parametrize("x0,y0", [(0.3, -0.2), (-0.7, 0.6), (0.0, 0.0), (0.95, -0.95)])
def test_quad2dfit_paraboloid(x0, y0):
    # Z(i,j) = Z[(j-1)*3+(i-1)] sampled at i = 1..3 (x), j = 1..3 (y) around a
    # peak at (2+x0, 2+y0): XSH recovers the x offset, YSH the y offset.
    z = np.zeros(9, np.float32)
    for j in range(3):
        for i in range(3):
            dx, dy = i - 1 - x0, j - 1 - y0
            z[j * 3 + i] = 10 - 2.0 * dx * dx - 1.0 * dy * dy + 0.3 * dx * dy
    xs, ys, pv = O.quad2dfit(z)
    assert xs == pytest.approx(x0, abs=1e-5)
    assert ys == pytest.approx(y0, abs=1e-5)


def test_quad2dfit_is_the_least_squares_paraboloid():
    # against numpy's least squares on the 6-term quadratic, random samples
    rng = np.random.default_rng(11)
    U = np.array([k % 3 - 1 for k in range(9)], float)
    V = np.array([k // 3 - 1 for k in range(9)], float)
    M = np.stack([np.ones(9), U, V, U * U, U * V, V * V], 1)
    for _ in range(200):
        z = rng.standard_normal(9).astype(np.float32)
        a, b, c, d, e, f = np.linalg.lstsq(M, z.astype(np.float64), rcond=None)[0]
        det = 4 * d * f - e * e
        u, v = (e * c - 2 * f * b) / det, (e * b - 2 * d * c) / det
        xs, ys, pv = O.quad2dfit(z)
        assert xs == pytest.approx(np.clip(u, -1, 1), abs=1e-5)
        assert ys == pytest.approx(np.clip(v, -1, 1), abs=1e-5)
        q = a + b * u + c * v + d * u * u + e * u * v + f * v * v
        assert pv == pytest.approx(q, rel=1e-5, abs=1e-5)


def test_speak_integer_and_subpixel():
    ny, nx = 16, 32
    yy, xx = np.mgrid[0:ny, 0:nx]
    for (px, py) in [(5, 3), (30, 15), (0, 0), (17, 9)]:
        d = 50 - ((xx - px + nx / 2) % nx - nx / 2) ** 2 - 2 * ((yy - py + ny / 2) % ny - ny / 2) ** 2
        d = np.maximum(d, 0.0)
        xt, yt = O.speak(d.astype(np.float32))
        ex = px if px <= nx // 2 else px - nx
        ey = py if py <= ny // 2 else py - ny
        assert (xt, yt) == (pytest.approx(ex, abs=1e-5), pytest.approx(ey, abs=1e-5))


def test_speak_scan_order_tie_break():
    # Equal peaks: the scan runs over x in the outer loop and takes a strict
    # ">" (align2dstack.cpp:1072-1078), so the smaller x wins even though the
    # other peak has the smaller y.
    d = np.zeros((8, 16), np.float32)
    d[2, 5] = 1.0
    d[1, 7] = 1.0
    assert O.speak(d) == (pytest.approx(5.0), pytest.approx(2.0))


def test_speak_signed_pmax_quirk():
    # pmax keeps the signed value: after a negative element takes over, the
    # next element scanned (here the zero below it) takes over unconditionally.
    d = np.zeros((8, 8), np.float32)
    d[3, 2] = -4.0  # x=2, y=3; the next scanned element is x=2, y=4 (value 0)
    xt, yt = O.speak(d)
    # the 3x3 fit around (2, 4) sees the -4 at offset (0, -1) and moves the
    # peak half a pixel away from it: 4.5 > ny/2 wraps to 4.5 - 8 = -3.5
    # (a scan that kept the -4 would centre the fit on y = 3 instead).
    assert xt == pytest.approx(2.0, abs=1e-6) and yt == pytest.approx(-3.5, abs=1e-6)


def test_fshift_integer_equals_roll():
    rng = np.random.default_rng(5)
    img = rng.random((8, 16)).astype(np.float32)
    out = O.fshift2d(img, 3.0, -2.0)
    np.testing.assert_allclose(out, np.roll(np.roll(img, 3, 1), -2, 0), atol=1e-5)


def test_fshift_fractional_direct_sum():
    rng = np.random.default_rng(6)
    ny, nx = 6, 8
    img = rng.random((ny, nx)).astype(np.float32)
    sx, sy = 0.37, -1.21
    kx = np.array([i / nx if i < np.ceil(nx / 2) else (i - nx) / nx for i in range(nx)])
    ky = np.array([j / ny if j < np.ceil(ny / 2) else (j - ny) / ny for j in range(ny)])
    F = np.fft.fft2(img.astype(np.float64))
    ph = np.exp(-2j * np.pi * (sx * kx[None, :] + sy * ky[:, None]))
    ref = np.real(np.fft.ifft2(F * ph))
    np.testing.assert_allclose(O.fshift2d(img, sx, sy), ref, atol=2e-6)


def test_cc_images_recovers_shift():
    rng = np.random.default_rng(8)
    ny, nx = 32, 32
    yy, xx = np.mgrid[0:ny, 0:nx]
    base = np.exp(-((xx - 15.2) ** 2 + (yy - 14.7) ** 2) / 18.0).astype(np.float32)
    meas = O.fshift2d(base, 2.0, -3.0)[None]
    sx, sy = O.cc_images(base[None], meas)
    assert sx[0] == pytest.approx(2.0, abs=0.1) and sy[0] == pytest.approx(-3.0, abs=0.1)
    back = O.shift_images(meas, sx, sy)
    np.testing.assert_allclose(back[0], base, atol=0.05)


def test_soft_noneg():
    x = np.array([-2.0, 0.0, 3.0, -0.5], np.float32)
    np.testing.assert_array_equal(O.soft_noneg(x), [-1.0, 0.0, 3.0, -0.25])


def _inner_object(method, nx, ny, ang, seed):
    """Views of a random non-negative object confined to 0.6 of the radius
    (zero margins on the detector), in FP64."""
    rng = np.random.default_rng(seed)
    r = 0.5 * (nx - 1)
    cord = O.cyl_geometry(r)
    inner = ((cord[:, 0] - r) ** 2 + (cord[:, 1] - r) ** 2) < (0.6 * r) ** 2
    x = rng.random((len(cord), ny)) * inner[:, None]
    return O.project(method, x, ang, nx, ny, dtype=np.float64)


@pytest.mark.parametrize("method", ["sq", "kb"])
def test_centroid_unshifted_views_give_zero(method):
    """A parallel-beam view's first moment is the object's projected
    centroid, a cos(t) + b sin(t) about the axis, so unshifted views give
    zero shifts -- up to the footprints' binning: a footprint whose width is
    not a whole number of pixels, sampled on the pixel grid, keeps its mass
    but not exactly its first moment (~1e-3 px here).  Pins the axis
    position (half a pixel off leaves ~0.3 px) and the removed gauge."""
    nx, ny, na = 41, 3, 23
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    views = _inner_object(method, nx, ny, ang, 1)
    s = O.centroid_x_shifts(method, views, ang, nx, ny)
    assert np.abs(s).max() < 5e-3


def test_centroid_recovers_integer_shifts():
    """Integer shifts (np.roll over zero margins moves the centroid exactly)
    come back up to the unobservable a cos(t) + b sin(t) -- and with the
    right sign and axis (x, not y)."""
    nx, ny, na = 41, 4, 31
    ang = (np.arange(na) * np.pi / na).astype(np.float32)
    views = _inner_object("sq", nx, ny, ang, 2)
    d = np.random.default_rng(3).integers(-4, 5, na)
    shifted = np.stack([np.roll(views[a], d[a], axis=1) for a in range(na)])
    s0 = O.centroid_x_shifts("sq", views, ang, nx, ny)
    s = O.centroid_x_shifts("sq", shifted, ang, nx, ny)
    t = ang.astype(np.float64)
    M = np.stack([np.cos(t), np.sin(t)], 1)
    want = d - M @ np.linalg.lstsq(M, d.astype(np.float64), rcond=None)[0]
    np.testing.assert_allclose(s - s0, want, atol=1e-9)
    # a y shift does not move the x centroid
    rolled_y = np.stack([np.roll(views[a], 1, axis=0) for a in range(na)])
    np.testing.assert_allclose(O.centroid_x_shifts("sq", rolled_y, ang, nx, ny), s0, atol=1e-9)


def test_take_log_closed_form():
    x = np.array([1.0, np.exp(-2.0), np.exp(3.0), 0.0, -1.5, 0.5], np.float32)
    out = O.take_log(x)
    np.testing.assert_allclose(out[:3], [0.0, 2.0, -3.0], atol=2e-7)
    assert out[3] == 0.0 and out[4] == 0.0
    assert out[5] == np.float32(-np.log(np.float32(0.5)))


def _boundary_image(u):
    """Moisan's boundary image of u, FP64, written independently."""
    u = u.astype(np.float64)
    v = np.zeros_like(u)
    v[0, :] += u[-1, :] - u[0, :]
    v[-1, :] += u[0, :] - u[-1, :]
    v[:, 0] += u[:, -1] - u[:, 0]
    v[:, -1] += u[:, 0] - u[:, -1]
    return v


@pytest.mark.parametrize("ny,nx", [(10, 12), (7, 9), (16, 16)])
def test_periodic_smooth_solves_the_periodic_poisson_equation(ny, nx):
    """The smooth part s solves Lap_per s = v (the boundary image), the
    equation Moisan's DFT quotient inverts: a wrong sign, a dropped term or
    a transposed frequency in the denominator breaks it.  s has zero mean."""
    rng = np.random.default_rng(nx * ny)
    u = rng.random((ny, nx)).astype(np.float32)
    s = O.periodic_smooth(u).astype(np.float64)
    lap = (np.roll(s, 1, 0) + np.roll(s, -1, 0) + np.roll(s, 1, 1) + np.roll(s, -1, 1) - 4 * s)
    v = _boundary_image(u)
    np.testing.assert_allclose(lap, v, atol=2e-6 * np.abs(v).max())
    assert abs(s.mean()) < 1e-6


def test_periodic_decomposition_window_and_periodic_input():
    """An image with no boundary jumps (a constant) is unchanged; otherwise
    u - out = W(d) s with W = exp(-d^2/50), d the distance to the border."""
    const = np.full((1, 8, 11), 2.5, np.float32)
    np.testing.assert_array_equal(O.periodic_decomposition(const), const)
    rng = np.random.default_rng(4)
    u = rng.random((2, 9, 13)).astype(np.float32)
    out = O.periodic_decomposition(u)
    iy, ix = np.mgrid[0:9, 0:13]
    d = np.minimum(np.minimum(ix, 12 - ix), np.minimum(iy, 8 - iy)).astype(np.float64)
    w = np.exp(-d * d / 50.0)
    for z in range(2):
        s = O.periodic_smooth(u[z]).astype(np.float64)
        np.testing.assert_allclose(u[z].astype(np.float64) - out[z], w * s, atol=3e-7)


def _fmaf(a, b, c):
    """fmaf(a, b, c) of FP32 inputs, exactly: a*b + c as a rational, rounded
    once to the nearest FP32 (ties to even)."""
    from fractions import Fraction

    a, b, c = np.float32(a), np.float32(b), np.float32(c)
    exact = Fraction(float(a)) * Fraction(float(b)) + Fraction(float(c))
    f = np.float32(float(exact))  # within one FP32 step of the exact value
    lo, hi = (np.nextafter(f, np.float32(-np.inf)), f) if Fraction(float(f)) > exact else \
        (f, np.nextafter(f, np.float32(np.inf)))
    dlo, dhi = exact - Fraction(float(lo)), Fraction(float(hi)) - exact
    if dlo != dhi:
        return lo if dlo < dhi else hi
    return lo if (lo.view(np.uint32) & 1) == 0 else hi


@pytest.mark.parametrize("method", ["sq", "kb", "linear"])
def test_device_arithmetic_oracle_is_fp32_weights_with_fma(method):
    """The device-arithmetic instantiation (an accuracy probe, not a
    restatement of the reference): an impulse's projection is the FP64
    projection rounded to FP32 (the weights are the FP64 weights rounded
    once), and a whole projection is, pixel by pixel, the chain of exact
    fmaf(w, x, y) over the columns in the reference's order (emulated here
    with rational arithmetic)."""
    nx, ny, na = 9, 1, 3
    ang = np.array([0.3, 1.1, 2.5], np.float32)
    nr = len(O.cyl_geometry(0.5 * (nx - 1)))
    W = np.zeros((na, nx, nr), np.float32)  # FP32 weight of column j at (angle, pixel)
    for j in range(nr):
        e = np.zeros((nr, ny), np.float32)
        e[j] = 1.0
        pg = O.project(method, e, ang, nx, ny, arith="device")
        pd = O.project(method, e.astype(np.float64), ang, nx, ny, dtype=np.float64)
        assert np.array_equal(pg, pd.astype(np.float32)), j
        W[:, :, j] = pg[:, 0, :]
    rng = np.random.default_rng(17)
    x = rng.random((nr, ny)).astype(np.float32) + np.float32(0.5)
    got = O.project(method, x, ang, nx, ny, arith="device")
    for a in range(na):
        for u in range(nx):
            y = np.float32(0.0)
            for j in range(nr):  # columns in order; a column adds to a pixel once
                if W[a, u, j] != 0:
                    y = _fmaf(W[a, u, j], x[j, 0], y)
            assert got[a, 0, u] == y, (a, u)


def test_device_arithmetic_solvers_close_to_fp64():
    """SIRT and CGLS with the device arithmetic share the reference's code
    and differ from it only in per-tap rounding.  SIRT stays within FP32
    rounding of the FP64 solution; CGLS in FP32 drifts from FP64 on this
    ill-conditioned problem with either arithmetic (5% after 8 iterations
    with the reference's), and the device arithmetic's drift is of the same
    size."""
    A, b, ang, nr = _small_problem(nx=11, ny=2, na=30)

    def rel(x, xd):
        return np.linalg.norm(x - xd) / np.linalg.norm(xd)

    lam = 1.0 / np.linalg.norm(A, 2) ** 2
    _, xg, hg, _ = O.sirt("sq", b, ang, lam, 20, 0.0, arith="device")
    _, xd, hd, _ = O.sirt("sq", b, ang, lam, 20, 0.0, dtype=np.float64)
    assert len(hg) == len(hd) and rel(xg, xd) <= 1e-5
    _, xg, hg = O.cgls("sq", b, ang, 8, 0.0, arith="device")
    _, xf, hf = O.cgls("sq", b, ang, 8, 0.0)
    _, xd, hd = O.cgls("sq", b, ang, 8, 0.0, dtype=np.float64)
    assert len(hg) == len(hd) == len(hf)
    assert rel(xg, xd) <= max(1e-5, 4 * rel(xf, xd))
    with pytest.raises(ValueError):
        O.cgls("sq", b, ang, 3, 0.0, weighted=True, arith="device")
    with pytest.raises(ValueError):
        O.project("sq", np.zeros((nr, 2)), ang, 11, 2, dtype=np.float64, arith="device")
