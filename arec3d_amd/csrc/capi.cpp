// extern "C" entry points of libarec3d_amd.so (declared in
// include/arec3d_amd.h).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "arec_internal.hpp"

namespace arec {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int hip_fail(hipError_t e, const char *what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return AREC_EHIP;
}

bool use_graphs() {
    const char *e = getenv("AREC_NO_GRAPH");
    return !(e && *e && e[0] != '0');
}

}  // namespace arec

using namespace arec;

#define CHECK_ARG(cond, msg)      \
    do {                          \
        if (!(cond)) {            \
            set_error(msg);       \
            return AREC_EINVAL;   \
        }                         \
    } while (0)

extern "C" {

const char *arec_last_error(void) { return g_last_error.c_str(); }
const char *arec_version(void) { return "arec3d_amd 0.1 (gfx950)"; }

int arec_create(arec_ctx **ctx, int device) {
    CHECK_ARG(ctx, "ctx is NULL");
    int n = 0;
    AREC_HIP(hipGetDeviceCount(&n));
    CHECK_ARG(device >= 0 && device < n, "no such HIP device");
    AREC_HIP(hipSetDevice(device));
    // the no-allocation entry point's scratch address on this device; a
    // failure here is reported by that entry point, which tries again
    noalloc_setup(device);
    arec_ctx *c = new arec_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    *ctx = c;
    return 0;
}

int arec_destroy(arec_ctx *ctx) {
    if (!ctx) return 0;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

int arec_synchronize(arec_ctx *ctx) {
    CHECK_ARG(ctx, "ctx is NULL");
    AREC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int arec_dev_alloc(arec_ctx *ctx, void **ptr, size_t bytes) {
    CHECK_ARG(ctx && ptr, "NULL argument");
    AREC_HIP(hipSetDevice(ctx->device));
    AREC_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}

int arec_dev_free(arec_ctx *ctx, void *ptr) {
    CHECK_ARG(ctx, "ctx is NULL");
    if (ptr) AREC_HIP(hipFree(ptr));
    return 0;
}

int arec_memcpy_h2d(arec_ctx *ctx, void *dst, const void *src, size_t bytes) {
    CHECK_ARG(ctx, "ctx is NULL");
    AREC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    AREC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int arec_memcpy_d2h(arec_ctx *ctx, void *dst, const void *src, size_t bytes) {
    CHECK_ARG(ctx, "ctx is NULL");
    AREC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    AREC_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int arec_memset(arec_ctx *ctx, void *dst, int value, size_t bytes) {
    CHECK_ARG(ctx, "ctx is NULL");
    AREC_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return 0;
}

int64_t arec_cyl_nrays(double radius, int32_t *cord) {
    std::vector<int> c;
    int n = cyl_mask(radius, &c);
    if (cord) memcpy(cord, c.data(), sizeof(int) * c.size());
    return n;
}

int arec_geom_create(arec_ctx *ctx, arec_geom **geom, int method, int64_t nx, int64_t ny,
                     double radius, const float *angles, int64_t nangles) {
    CHECK_ARG(ctx && geom && angles, "NULL argument");
    CHECK_ARG(nx > 0 && ny > 0 && nangles > 0 && nx < (1 << 30) && ny < (1 << 30) && nangles < (1 << 30),
              "sizes out of range");
    CHECK_ARG((int64_t)std::lround(2 * radius + 1) == nx,
              "detector width must equal the cylinder width round(2*radius+1)");
    AREC_HIP(hipSetDevice(ctx->device));
    arec_geom *G = new arec_geom;
    std::string err;
    int rc = geometry_build_host(G->g, method, (int)nx, (int)ny, radius, angles, (int)nangles, &err);
    if (!rc) rc = geometry_upload(G->g, ctx->stream, &err);
    if (rc) {
        if (!err.empty()) set_error(err);
        geometry_free_device(G->g);
        delete G;
        return rc;
    }
    *geom = G;
    return 0;
}

int arec_geom_destroy(arec_geom *geom) {
    if (!geom) return 0;
    geometry_free_device(geom->g);
    delete geom;
    return 0;
}

int arec_geom_info(const arec_geom *geom, int64_t *nrays, int64_t *nnz, int64_t *nproj) {
    CHECK_ARG(geom, "geom is NULL");
    if (nrays) *nrays = geom->g.nrays;
    if (nnz) *nnz = geom->g.nnz;
    if (nproj) *nproj = geom->g.nproj;
    return 0;
}

int64_t arec_geom_device_bytes(const arec_geom *geom) {
    if (!geom) return 0;
    const Geometry &g = geom->g;
    const int64_t btab = (g.d_btab ? (int64_t)sizeof(uint4) * g.ntiles * g.na * kTileCols : 0) +
                         (g.d_strip ? (int64_t)sizeof(uint4) * g.strip_unit * g.strip_base.back() : 0) +
                         (g.d_sg_tab ? (int64_t)sizeof(uint4) * (g.sg_group_ent.back() + kStripBatch) +
                                           g.sg_chunk_base.back() * (int64_t)(sizeof(int) * kSgChunk +
                                                                              sizeof(int64_t) * kSgAngles)
                                     : 0);
    return (g.d_fwd ? (int64_t)sizeof(uint4) * g.na * g.nvalid : 0) + btab + (int64_t)sizeof(AngleConst) * g.na +
           (int64_t)sizeof(int) * (g.coord.size() + g.tile_cols.size() + g.tile_org.size() + g.seg.size() +
                                   g.pixstart.size());
}

int arec_project(arec_ctx *ctx, const arec_geom *geom, const float *x_dev, float *p_dev) {
    CHECK_ARG(ctx && geom && x_dev && p_dev, "NULL argument");
    return launch_forward(geom->g, x_dev, p_dev, nullptr, ctx->stream, nullptr);
}

int arec_backproject(arec_ctx *ctx, const arec_geom *geom, const float *p_dev, float *x_dev) {
    CHECK_ARG(ctx && geom && x_dev && p_dev, "NULL argument");
    return launch_backward(geom->g, p_dev, x_dev, kBpStore, nullptr, nullptr, ctx->stream, nullptr);
}

int arec_sirt(arec_ctx *ctx, const arec_geom *geom, const float *b_dev, float *x_dev, float lam,
              int32_t maxit, float tol, arec_hist *hist, int32_t hist_cap, int32_t *nhist,
              float *lam_out) {
    CHECK_ARG(ctx && geom && b_dev && x_dev, "NULL argument");
    CHECK_ARG(hist_cap >= 0 && (hist || hist_cap == 0), "bad history buffer");
    return sirt_run(geom->g, ctx->stream, b_dev, x_dev, lam, maxit, tol, 0, 0.0, 1, hist, hist_cap, nhist,
                    lam_out, use_graphs());
}

int arec_sirt_warm(arec_ctx *ctx, const arec_geom *geom, const float *b_dev, float *x_dev, float lam,
                   int32_t maxit, float tol, double reject_above, int32_t check_pass, arec_hist *hist,
                   int32_t hist_cap, int32_t *nhist, float *lam_out) {
    CHECK_ARG(ctx && geom && b_dev && x_dev, "NULL argument");
    CHECK_ARG(hist_cap >= 0 && (hist || hist_cap == 0), "bad history buffer");
    CHECK_ARG(check_pass >= 1, "check_pass must be >= 1");
    return sirt_run(geom->g, ctx->stream, b_dev, x_dev, lam, maxit, tol, 1, reject_above, check_pass, hist,
                    hist_cap, nhist, lam_out, use_graphs());
}

int arec_sirt_batch(arec_ctx *ctx, int32_t nprob, const arec_geom *const *geoms, const float *const *b_dev,
                    float *const *x_dev, const float *lam, int32_t maxit, float tol, arec_hist *hist,
                    int32_t hist_cap, int32_t *nhist, float *lam_out) {
    CHECK_ARG(ctx && nprob >= 0 && (nprob == 0 || (geoms && b_dev && x_dev && lam)), "NULL argument");
    CHECK_ARG(hist_cap >= 0 && (hist || hist_cap == 0), "bad history buffer");
    std::vector<SirtProblem> P;
    for (int k = 0; k < nprob; k++) {
        CHECK_ARG(geoms[k] && b_dev[k] && x_dev[k], "NULL problem");
        P.push_back(SirtProblem{&geoms[k]->g, b_dev[k], x_dev[k], lam[k]});
    }
    return sirt_batch_run(P, ctx->stream, maxit, tol, 0, 0.0, 1, hist, hist_cap, nhist, lam_out, use_graphs());
}

int arec_geom_plan(arec_ctx *ctx, arec_geom *geom, int32_t rows_per_lane, int32_t min_groups) {
    CHECK_ARG(ctx && geom, "NULL argument");
    CHECK_ARG(min_groups >= 0 && rows_per_lane >= 0, "bad plan");
    return geometry_set_forward(geom->g, 1, rows_per_lane, min_groups, ctx->stream);
}

int arec_geom_set_forward(arec_ctx *ctx, arec_geom *geom, int32_t design, int32_t rows_per_lane,
                          int32_t min_groups) {
    CHECK_ARG(ctx && geom, "NULL argument");
    CHECK_ARG(min_groups >= 0 && rows_per_lane >= 0, "bad plan");
    return geometry_set_forward(geom->g, design, rows_per_lane, min_groups, ctx->stream);
}

int32_t arec_geom_forward_design(const arec_geom *geom) { return geom ? geom->g.fwd_design : 0; }

int arec_geom_set_back(arec_ctx *ctx, arec_geom *geom, int32_t rows, int32_t angles) {
    CHECK_ARG(ctx && geom, "NULL argument");
    for (int s = 0; s < kBpShapes; s++)
        if (kBpShapeRows[s] == rows && kBpShapeAngles[s] == angles) return geometry_set_back(geom->g, s);
    set_error("back projector: tile shape must be 32 x 2, 16 x 4 or 64 x 2 (rows x angles)");
    return AREC_EINVAL;
}

int arec_geom_get_back(const arec_geom *geom, int32_t *rows, int32_t *angles) {
    CHECK_ARG(geom, "NULL argument");
    if (rows) *rows = kBpShapeRows[geom->g.bp_shape];
    if (angles) *angles = kBpShapeAngles[geom->g.bp_shape];
    return 0;
}

int arec_geom_get_plan(const arec_geom *geom, int32_t *rows_per_lane, int32_t *min_groups, int32_t *nseg) {
    CHECK_ARG(geom, "NULL argument");
    if (rows_per_lane) *rows_per_lane = geom->g.fwd_design >= 2 ? geom->g.strip_rpl : geom->g.fwd_rpl;
    if (min_groups) *min_groups = geom->g.fwd_min_groups;
    if (nseg) *nseg = geom->g.nseg;
    return 0;
}

int arec_cgls(arec_ctx *ctx, const arec_geom *geom, const float *b_dev, const float *w_dev,
              float *x_dev, int32_t warm, int32_t maxit, float tol, arec_hist *hist,
              int32_t hist_cap, int32_t *nhist) {
    CHECK_ARG(ctx && geom && b_dev && x_dev, "NULL argument");
    CHECK_ARG(hist_cap >= 0 && (hist || hist_cap == 0), "bad history buffer");
    return cgls_run(geom->g, ctx->stream, b_dev, w_dev, x_dev, warm, maxit, tol, hist, hist_cap,
                    nhist, use_graphs());
}

int arec_init_weights(arec_ctx *ctx, const float *b_dev, float *w_dev, int64_t nx, int64_t ny,
                      int64_t nz) {
    CHECK_ARG(ctx && b_dev && w_dev && nx > 1 && ny > 0 && nz > 0, "bad argument");
    return init_weights(b_dev, w_dev, nx, ny, nz, ctx->stream);
}

int arec_soft_noneg(arec_ctx *ctx, float *x_dev, int64_t n) {
    CHECK_ARG(ctx && x_dev && n >= 0, "bad argument");
    return soft_noneg(x_dev, n, ctx->stream);
}

int arec_dot(arec_ctx *ctx, const float *a_dev, const float *b_dev, int64_t n, double *out) {
    CHECK_ARG(ctx && a_dev && b_dev && out && n >= 0, "bad argument");
    return dot(a_dev, b_dev, n, out, ctx->stream);
}

// Average duration of `reps` back-to-back launches of one operator, measured
// with HIP events on the context's stream (the stream the kernels run on).
int arec_time_operator(arec_ctx *ctx, const arec_geom *geom, int which, const float *in_dev,
                       float *out_dev, int32_t reps, float *ms_per_launch) {
    CHECK_ARG(ctx && geom && in_dev && out_dev && ms_per_launch && reps > 0, "bad argument");
    CHECK_ARG(which == 0 || which == 1, "which: 0 forward, 1 back");
    hipEvent_t e0, e1;
    AREC_HIP(hipEventCreate(&e0));
    AREC_HIP(hipEventCreate(&e1));
    AREC_HIP(hipEventRecord(e0, ctx->stream));
    int rc = 0;
    for (int r = 0; r < reps && !rc; r++)
        rc = which == 0 ? launch_forward(geom->g, in_dev, out_dev, nullptr, ctx->stream, nullptr)
                        : launch_backward(geom->g, in_dev, out_dev, kBpStore, nullptr, nullptr, ctx->stream, nullptr);
    AREC_HIP(hipEventRecord(e1, ctx->stream));
    AREC_HIP(hipEventSynchronize(e1));
    float ms = 0;
    AREC_HIP(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms_per_launch = ms / reps;
    return rc;
}

int arec_copy(arec_ctx *ctx, float *dst_dev, const float *src_dev, int64_t n) {
    CHECK_ARG(ctx && dst_dev && src_dev && n >= 0, "bad argument");
    return copy(dst_dev, src_dev, n, ctx->stream);
}

int arec_scale(arec_ctx *ctx, float *x_dev, int64_t n, float a) {
    CHECK_ARG(ctx && x_dev && n >= 0, "bad argument");
    return scale(x_dev, a, n, ctx->stream);
}

int arec_max(arec_ctx *ctx, const float *x_dev, int64_t n, float *out) {
    CHECK_ARG(ctx && x_dev && out && n > 0, "bad argument");
    return reduce_max(x_dev, n, out, ctx->stream);
}

// ---- host-pointer layer ----------------------------------------------------

namespace {

// One process-wide context for the host-pointer entry points (the reference's
// functions take no context argument).
arec_ctx *host_ctx(int *rc) {
    static arec_ctx *ctx = nullptr;
    if (!ctx) {
        *rc = arec_create(&ctx, 0);
        if (*rc) ctx = nullptr;
    }
    return ctx;
}

struct HostScratch {
    arec_ctx *ctx;
    std::vector<void *> ptrs;
    explicit HostScratch(arec_ctx *c) : ctx(c) {}
    ~HostScratch() {
        for (void *p : ptrs) hipFree(p);
    }
    float *get(size_t n, int *rc) {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(float)) != hipSuccess) {
            set_error("hipMalloc failed");
            *rc = AREC_EHIP;
            return nullptr;
        }
        ptrs.push_back(p);
        return (float *)p;
    }
};

}  // namespace

// arecProject2D* status rules (arecproject.cpp:39-68, 551-580): not a
// cylinder -> 1; projstack->nx != cylinder width -> 2.
int arec_h_project(int method, double radius, int32_t nx, int32_t ny, const float *x,
                   const float *angles, int32_t nangles, float *p) {
    CHECK_ARG(x && p && angles && nangles > 0 && ny > 0, "bad argument");
    if ((int)std::lround(2 * radius + 1) != nx) {
        set_error("arecProject2D: inconsistent dimension!");
        return 2;
    }
    int rc = 0;
    arec_ctx *ctx = host_ctx(&rc);
    if (!ctx) return rc;
    arec_geom *G = nullptr;
    if ((rc = arec_geom_create(ctx, &G, method, nx, ny, radius, angles, nangles))) return rc;
    HostScratch s(ctx);
    const Geometry &g = G->g;
    float *dx = s.get(g.nnz, &rc), *dp = rc ? nullptr : s.get(g.nproj, &rc);
    if (!rc) rc = arec_memcpy_h2d(ctx, dx, x, sizeof(float) * g.nnz);
    if (!rc) rc = arec_project(ctx, G, dx, dp);
    if (!rc) rc = arec_memcpy_d2h(ctx, p, dp, sizeof(float) * g.nproj);
    arec_geom_destroy(G);
    return rc;
}

// arecBackProject2D* status rules (arecproject.cpp:117-140, 654-676):
// nz != nangles -> 1; bvol->nx != images.nx -> -1.
int arec_h_backproject(int method, double radius, int32_t nx, int32_t ny, int32_t nz,
                       const float *p, const float *angles, int32_t nangles, float *x) {
    CHECK_ARG(x && p && angles && nangles > 0 && ny > 0, "bad argument");
    if (nz != nangles) return 1;
    if ((int)std::lround(2 * radius + 1) != nx) {
        set_error("arecBackProject2D: image and volume sizes don't match");
        return -1;
    }
    int rc = 0;
    arec_ctx *ctx = host_ctx(&rc);
    if (!ctx) return rc;
    arec_geom *G = nullptr;
    if ((rc = arec_geom_create(ctx, &G, method, nx, ny, radius, angles, nangles))) return rc;
    HostScratch s(ctx);
    const Geometry &g = G->g;
    float *dx = s.get(g.nnz, &rc), *dp = rc ? nullptr : s.get(g.nproj, &rc);
    if (!rc) rc = arec_memcpy_h2d(ctx, dp, p, sizeof(float) * g.nproj);
    if (!rc) rc = arec_backproject(ctx, G, dp, dx);
    if (!rc) rc = arec_memcpy_d2h(ctx, x, dx, sizeof(float) * g.nnz);
    arec_geom_destroy(G);
    return rc;
}

// cyl_sirt* / cyl_cgls* / cyl_cgls*_stat.  radius 0.5(nx-1) (arecsirt.cpp:37,
// areccgls.cpp:40); the KB CGLS and the weighted solvers use the integer
// nx/2 (areccgls_KB.cpp:39, areccgls_stat.cpp:78), which for even nx makes
// the cylinder one wider than the detector and the first back projection
// fail with -1.
int arec_h_solve(int solver, int method, const float *b, int32_t nx, int32_t ny, int32_t nz,
                 const float *angles, float lam, int32_t maxit, float tol, int32_t warm, float *x,
                 arec_hist *hist, int32_t hist_cap, int32_t *nhist, float *lam_out) {
    CHECK_ARG(b && x && angles && nx > 0 && ny > 0 && nz > 0, "bad argument");
    CHECK_ARG(solver >= 1 && solver <= 3, "solver must be 1 (SIRT), 2 (CGLS) or 3 (weighted CGLS)");
    double radius = 0.5 * (nx - 1);
    if ((solver == 2 && method == kKB) || solver == 3) radius = (double)(nx / 2);
    if ((int)std::lround(2 * radius + 1) != nx) return -1;
    int rc = 0;
    arec_ctx *ctx = host_ctx(&rc);
    if (!ctx) return rc;
    arec_geom *G = nullptr;
    if ((rc = arec_geom_create(ctx, &G, method, nx, ny, radius, angles, nz))) return rc;
    HostScratch s(ctx);
    const Geometry &g = G->g;
    float *db = s.get(g.nproj, &rc);
    float *dx = rc ? nullptr : s.get(g.nnz, &rc);
    float *dw = (rc || solver != 3) ? nullptr : s.get(g.nproj, &rc);
    if (!rc) rc = arec_memcpy_h2d(ctx, db, b, sizeof(float) * g.nproj);
    if (!rc && warm && solver != 1) rc = arec_memcpy_h2d(ctx, dx, x, sizeof(float) * g.nnz);
    if (!rc && dw) rc = arec_init_weights(ctx, db, dw, nx, ny, nz);
    if (!rc) {
        if (solver == 1)
            rc = arec_sirt(ctx, G, db, dx, lam, maxit, tol, hist, hist_cap, nhist, lam_out);
        else
            rc = arec_cgls(ctx, G, db, dw, dx, warm, maxit, tol, hist, hist_cap, nhist);
    }
    if (!rc) rc = arec_memcpy_d2h(ctx, x, dx, sizeof(float) * g.nnz);
    arec_geom_destroy(G);
    return rc;
}

}  // extern "C"
