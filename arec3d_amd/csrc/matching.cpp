// Projection matching (the alignment loop of AREC3D.cpp:185-325), device
// resident: the views, their aligned copy, the reprojection and the volume
// stay in HBM; only the per-view shifts (8 bytes a view) cross to the host.
//
// One outer loop (it = 0, 1, ...):
//   prj = A x                                  (:188-191)
//   aligned = R(views0); s = 0                 (:201-204; R below)
//   cc_passes times: t = CC(prj, aligned)      (:215-223, arecCCImages)
//                    aligned = shift(aligned, -t); s += t
//   aligned = shift(R(views0), -s)             (:227-230, fresh views)
//   with the rotational step (align = 2), every third loop:
//     a = RotCC(prj, aligned); aligned = RotSafe(aligned, prj, a);
//     rot += a                                 (:235-243)
//   x = SIRT from zero, or CGLS from x         (:294-315), or SIRT from x
//   x = soft_noneg(x)                          (:318)
// where R rotates by the summed angles rot from the second loop on
// (arecRotateImages_skew; identity without the rotational step), preceded
// by the initial reconstruction from views0 (:140-165).  The
// reference moves views and slabs between these steps through MRC files
// (:197-202,249,293); here they never leave the device.
//
// Option (no counterpart in the reference): x shifts from the views'
// centroids.  For parallel-beam views of an object inside the field of
// view, a view's centroid along the detector is the object's centroid
// projected at the view's angle -- a cos(t) + b sin(t) about the rotation
// axis -- plus the view's own x shift, so the shifts follow from one pass
// over the views, without the reconstruction's part-absorption of in-plane
// misalignment that slows the reprojection loop in x.  The axis position is
// the centroid of the projected uniform cylinder (symmetric about the axis
// under every method's detector convention); a and b are not observable and
// are removed by least squares, as any alignment leaves them free.  y shifts
// still come from the cross-correlation passes (the y centroid is not used:
// objects that fill the detector rows break it).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "arec_internal.hpp"

namespace arec {

namespace {

struct Dev {
    float *p = nullptr;
    ~Dev() {
        if (p) hipFree(p);
    }
};

int dalloc(Dev &d, size_t n) {
    AREC_HIP(hipMalloc((void **)&d.p, sizeof(float) * std::max<size_t>(1, n)));
    return 0;
}

int reconstruct(const Geometry &g, hipStream_t st, int method, const float *views, float *x, float lam,
                int maxit, float tol, bool first, arec_hist *last) {
    std::vector<arec_hist> h(maxit + 64);
    int nh = 0, rc;
    if (method == 1 || method == 3)
        rc = sirt_run(g, st, views, x, lam, maxit, tol, method == 3 && !first, 0.0, 1, h.data(), (int)h.size(),
                      &nh, nullptr, use_graphs());
    else
        rc = cgls_run(g, st, views, nullptr, x, first ? 0 : 1, maxit, tol, h.data(), (int)h.size(), &nh,
                      use_graphs());
    if (rc) return rc;
    if (last) *last = nh > 0 ? h[std::min(nh, (int)h.size()) - 1] : arec_hist{0, 0, 0.0, 0.0};
    return 0;
}

// Progress lines on stderr when AREC_PROGRESS is set (long runs at full
// slab size take minutes per outer loop).
struct Progress {
    bool on;
    std::chrono::steady_clock::time_point t0;
    Progress() : on(getenv("AREC_PROGRESS") != nullptr), t0(std::chrono::steady_clock::now()) {}
    void operator()(const char *what, int loop) const {
        if (!on) return;
        double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        fprintf(stderr, "[match] %s (loop %d) at %.1f s\n", what, loop, s);
        fflush(stderr);
    }
};

}  // namespace

int centroid_x_shifts(const Geometry &g, hipStream_t st, const float *views, float *sx_host) {
    const int na = g.na;
    Dev ones, prj;
    double *mom = nullptr;
    int rc;
    if ((rc = dalloc(ones, g.nnz)) || (rc = dalloc(prj, g.nproj))) return rc;
    AREC_HIP(hipMalloc((void **)&mom, sizeof(double) * 6 * (size_t)na));
    std::vector<double> m(6 * (size_t)na);
    rc = fill(ones.p, 1.0f, g.nnz, st, nullptr);
    if (!rc) rc = launch_forward(g, ones.p, prj.p, nullptr, st, nullptr);
    if (!rc) rc = view_moments(prj.p, g.nx, g.ny, na, mom, st);
    if (!rc) rc = view_moments(views, g.nx, g.ny, na, mom + 3 * (size_t)na, st);
    if (!rc && hipMemcpyAsync(m.data(), mom, sizeof(double) * m.size(), hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = hip_fail(hipGetLastError(), "centroid_x_shifts");
    if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = hip_fail(hipGetLastError(), "centroid_x_shifts");
    hipFree(mom);
    if (rc) return rc;
    // d_a = view centroid - axis; fit d ~ a cos + b sin (2x2 normal equations)
    std::vector<double> d(na);
    double scc = 0, sss = 0, scs = 0, scd = 0, ssd = 0;
    for (int a = 0; a < na; a++) {
        const double *ax = &m[3 * (size_t)a], *vw = &m[3 * ((size_t)na + a)];
        if (!(ax[0] > 0.0) || !(vw[0] > 0.0)) {
            set_error("centroid x shifts: a view (or the projected cylinder) has no positive mass");
            return AREC_EINVAL;
        }
        d[a] = vw[1] / vw[0] - ax[1] / ax[0];
        const double c = g.ang[a].c, s = g.ang[a].s;
        scc += c * c;
        sss += s * s;
        scs += c * s;
        scd += c * d[a];
        ssd += s * d[a];
    }
    const double det = scc * sss - scs * scs;
    const double ca = det > 0.0 ? (scd * sss - ssd * scs) / det : 0.0;
    const double cb = det > 0.0 ? (ssd * scc - scd * scs) / det : 0.0;
    for (int a = 0; a < na; a++) sx_host[a] = (float)(d[a] - ca * g.ang[a].c - cb * g.ang[a].s);
    return 0;
}

int match_run(const Geometry &g, hipStream_t st, const float *views0, float *x, float lam, int outer,
              int maxit, float tol, int cc_passes, int method, int flags, float *sx_host, float *sy_host,
              float *rot_host, arec_hist *loop_hist) {
    const int64_t nproj = g.nproj, nnz = g.nnz;
    const int nx = g.nx, ny = g.ny, na = g.na;
    const bool xc = (flags & kMatchXCentroid) != 0, rotating = (flags & kMatchRotate) != 0;
    std::vector<float> hs(2 * (size_t)na), ht(2 * (size_t)na), cx(na, 0.0f), rsum(na, 0.0f), rcur(na);
    // the driver's profile radius: min(2 radius + 1, height - 10) / 2 with
    // radius (nx-1)/2 and height ny (AREC3D.cpp:237-238)
    const int r2 = std::min(nx, ny - 10) / 2;
    int rc;
    if (rotating && (r2 < 1 || (rc = rot_check(nx, ny, na, "arec_match")))) {
        if (r2 < 1) set_error("arec_match: the rotational step needs ny >= 12 and nx >= 2");
        return AREC_EINVAL;
    }
    // centroid x shifts first: their scratch (a volume and a view stack) is
    // freed before the loop's buffers are allocated
    if (xc && (rc = centroid_x_shifts(g, st, views0, cx.data()))) return rc;
    Dev aligned, prj, s, t;
    if ((rc = dalloc(aligned, nproj)) || (rc = dalloc(prj, nproj)) || (rc = dalloc(s, 2 * (size_t)na)) ||
        (rc = dalloc(t, 2 * (size_t)na)))
        return rc;
    const Progress progress;
    // views0, rotated by the summed angles when `rotate`, then shifted by the
    // host shifts h (x from h[0..na), y from h[na..2na))
    auto shifted_views = [&](const std::vector<float> &h, bool rotate) -> int {
        int r;
        AREC_HIP(hipMemcpyAsync(s.p, h.data(), sizeof(float) * 2 * na, hipMemcpyHostToDevice, st));
        if ((r = copy(aligned.p, views0, nproj, st))) return r;
        if (rotate && (r = rotate_views(aligned.p, nx, ny, na, rsum.data(), AREC_ROT_SKEW, st))) return r;
        return shift_views(aligned.p, nx, ny, na, s.p, s.p + na, -1.0f, st);
    };
    // initial reconstruction from the unshifted views (or, with centroid x
    // shifts, from the views shifted in x)
    if (xc) {
        std::copy(cx.begin(), cx.end(), hs.begin());
        if ((rc = shifted_views(hs, false))) return rc;
        progress("centroid x shifts done", -1);
    }
    if ((rc = reconstruct(g, st, method, xc ? aligned.p : views0, x, lam, maxit, tol, true, nullptr))) return rc;
    progress("initial reconstruction done", -1);
    for (int it = 0; it < outer; it++) {
        const bool rotate = rotating && it > 0;
        if ((rc = launch_forward(g, x, prj.p, nullptr, st, nullptr))) return rc;
        std::fill(hs.begin(), hs.end(), 0.0f);
        if (xc) {
            std::copy(cx.begin(), cx.end(), hs.begin());
            if ((rc = shifted_views(hs, rotate))) return rc;
        } else if ((rc = copy(aligned.p, views0, nproj, st))) {
            return rc;
        } else if (rotate && (rc = rotate_views(aligned.p, nx, ny, na, rsum.data(), AREC_ROT_SKEW, st))) {
            return rc;
        }
        for (int pass = 0; pass < cc_passes; pass++) {
            if ((rc = cc_views(prj.p, aligned.p, nx, ny, na, nullptr, t.p, t.p + na, st))) return rc;
            AREC_HIP(hipMemcpyAsync(ht.data(), t.p, sizeof(float) * 2 * na, hipMemcpyDeviceToHost, st));
            AREC_HIP(hipStreamSynchronize(st));
            if (xc) {  // x stays at the centroid shifts
                std::fill(ht.begin(), ht.begin() + na, 0.0f);
                AREC_HIP(hipMemcpyAsync(t.p, ht.data(), sizeof(float) * na, hipMemcpyHostToDevice, st));
            }
            if ((rc = shift_views(aligned.p, nx, ny, na, t.p, t.p + na, -1.0f, st))) return rc;
            for (size_t i = 0; i < hs.size(); i++) hs[i] += ht[i];  // FP32 sums (:220-223)
        }
        // the fresh views shifted by the loop's total
        if ((rc = shifted_views(hs, rotate))) return rc;
        if (rotating && (it + 1) % 3 == 0) {  // the rotational step (AREC3D.cpp:235-243)
            if ((rc = rot_cc_views(prj.p, aligned.p, nx, ny, na, r2, rcur.data(), st))) return rc;
            if ((rc = rotate_views_safe(aligned.p, prj.p, nx, ny, na, rcur.data(), nullptr, nullptr, st))) return rc;
            for (int a = 0; a < na; a++) rsum[a] = rsum[a] + rcur[a];
            progress("rotational step done", it);
        }
        progress("reprojection and alignment done", it);
        arec_hist last;
        if ((rc = reconstruct(g, st, method, aligned.p, x, lam, maxit, tol, false, &last))) return rc;
        if ((rc = soft_noneg(x, nnz, st))) return rc;
        if (loop_hist) loop_hist[it] = last;
        progress("reconstruction done", it);
    }
    AREC_HIP(hipStreamSynchronize(st));
    if (sx_host) std::copy(hs.begin(), hs.begin() + na, sx_host);
    if (sy_host) std::copy(hs.begin() + na, hs.end(), sy_host);
    if (rot_host) std::copy(rsum.begin(), rsum.end(), rot_host);
    return 0;
}

}  // namespace arec

using namespace arec;

extern "C" int arec_match_ex(arec_ctx *ctx, const arec_geom *geom, const float *views_dev, float *x_dev,
                             float lam, int32_t outer, int32_t maxit, float tol, int32_t cc_passes,
                             int32_t method, int32_t flags, float *sx, float *sy, float *rot,
                             arec_hist *loop_hist) {
    if (!ctx || !geom || !views_dev || !x_dev || outer < 0 || maxit < 1 || cc_passes < 0 || method < 1 ||
        method > 3 || (flags & ~(kMatchXCentroid | kMatchRotate))) {
        set_error("arec_match: bad argument (method 1 = SIRT, 2 = CGLS, 3 = SIRT continued; flags bits 0-1)");
        return AREC_EINVAL;
    }
    return match_run(geom->g, ctx->stream, views_dev, x_dev, lam, outer, maxit, tol, cc_passes, method, flags, sx,
                     sy, rot, loop_hist);
}

extern "C" int arec_match(arec_ctx *ctx, const arec_geom *geom, const float *views_dev, float *x_dev, float lam,
                          int32_t outer, int32_t maxit, float tol, int32_t cc_passes, int32_t method,
                          float *sx, float *sy, arec_hist *loop_hist) {
    return arec_match_ex(ctx, geom, views_dev, x_dev, lam, outer, maxit, tol, cc_passes, method, 0, sx, sy,
                         nullptr, loop_hist);
}

extern "C" int arec_align_centroid(arec_ctx *ctx, const arec_geom *geom, const float *views_dev, float *sx) {
    if (!ctx || !geom || !views_dev || !sx) {
        set_error("arec_align_centroid: NULL argument");
        return AREC_EINVAL;
    }
    return centroid_x_shifts(geom->g, ctx->stream, views_dev, sx);
}
