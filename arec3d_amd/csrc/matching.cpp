// Projection matching (the alignment loop of AREC3D.cpp:185-325), device
// resident: the views, their aligned copy, the reprojection and the volume
// stay in HBM; only the per-view shifts (8 bytes a view) cross to the host.
//
// One outer loop (align = 1, no rotational step):
//   prj = A x                                  (:188-191)
//   aligned = views0; s = 0
//   cc_passes times: t = CC(prj, aligned)      (:215-223, arecCCImages)
//                    aligned = shift(aligned, -t); s += t
//   aligned = shift(views0, -s)                (:227-229, fresh views)
//   x = SIRT from zero, or CGLS from x         (:294-315), or SIRT from x
//   x = soft_noneg(x)                          (:318)
// preceded by the initial reconstruction from views0 (:140-165).  The
// reference moves views and slabs between these steps through MRC files
// (:197-202,249,293); here they never leave the device.
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

int match_run(const Geometry &g, hipStream_t st, const float *views0, float *x, float lam, int outer,
              int maxit, float tol, int cc_passes, int method, float *sx_host, float *sy_host,
              arec_hist *loop_hist) {
    const int64_t nproj = g.nproj, nnz = g.nnz;
    const int nx = g.nx, ny = g.ny, na = g.na;
    Dev aligned, prj, s, t;
    int rc;
    if ((rc = dalloc(aligned, nproj)) || (rc = dalloc(prj, nproj)) || (rc = dalloc(s, 2 * (size_t)na)) ||
        (rc = dalloc(t, 2 * (size_t)na)))
        return rc;
    std::vector<float> hs(2 * (size_t)na), ht(2 * (size_t)na);
    const Progress progress;
    // initial reconstruction from the unshifted views
    if ((rc = reconstruct(g, st, method, views0, x, lam, maxit, tol, true, nullptr))) return rc;
    progress("initial reconstruction done", -1);
    for (int it = 0; it < outer; it++) {
        if ((rc = launch_forward(g, x, prj.p, nullptr, st, nullptr))) return rc;
        if ((rc = copy(aligned.p, views0, nproj, st))) return rc;
        std::fill(hs.begin(), hs.end(), 0.0f);
        for (int pass = 0; pass < cc_passes; pass++) {
            if ((rc = cc_views(prj.p, aligned.p, nx, ny, na, nullptr, t.p, t.p + na, st))) return rc;
            if ((rc = shift_views(aligned.p, nx, ny, na, t.p, t.p + na, -1.0f, st))) return rc;
            AREC_HIP(hipMemcpyAsync(ht.data(), t.p, sizeof(float) * 2 * na, hipMemcpyDeviceToHost, st));
            AREC_HIP(hipStreamSynchronize(st));
            for (size_t i = 0; i < hs.size(); i++) hs[i] += ht[i];  // FP32 sums (:220-223)
        }
        // the fresh views shifted by the loop's total
        AREC_HIP(hipMemcpyAsync(s.p, hs.data(), sizeof(float) * 2 * na, hipMemcpyHostToDevice, st));
        if ((rc = copy(aligned.p, views0, nproj, st))) return rc;
        if ((rc = shift_views(aligned.p, nx, ny, na, s.p, s.p + na, -1.0f, st))) return rc;
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
    return 0;
}

}  // namespace arec

using namespace arec;

extern "C" int arec_match(arec_ctx *ctx, const arec_geom *geom, const float *views_dev, float *x_dev, float lam,
                          int32_t outer, int32_t maxit, float tol, int32_t cc_passes, int32_t method,
                          float *sx, float *sy, arec_hist *loop_hist) {
    if (!ctx || !geom || !views_dev || !x_dev || outer < 0 || maxit < 1 || cc_passes < 0 || method < 1 ||
        method > 3) {
        set_error("arec_match: bad argument (method 1 = SIRT, 2 = CGLS, 3 = SIRT continued)");
        return AREC_EINVAL;
    }
    return match_run(geom->g, ctx->stream, views_dev, x_dev, lam, outer, maxit, tol, cc_passes, method, sx, sy,
                     loop_hist);
}
