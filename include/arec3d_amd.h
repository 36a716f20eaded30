/*
 * arec3d_amd -- MI355X-native reconstruction hot path of AREC3D.
 *
 * C ABI of libarec3d_amd.so: plain pointers, 64-bit sizes, no C++ or torch
 * types.  Two layers:
 *
 *  1. Host-pointer entry points (arec_h_*).  Same arguments, meaning and
 *     status codes as the reference's projector / solver / alignment
 *     functions; data is uploaded, computed on the GPU and downloaded.  Each
 *     cites the reference function it replaces.  A C++ caller that wants the
 *     reference's exact symbols (arecImage by value, MPI_Comm) links
 *     libarec3d_dropin.so, whose declarations are in arec3d_dropin.h.
 *
 *  2. Device-resident entry points (arec_*).  Pointers are device pointers;
 *     geometry tables live on the GPU; work is queued on the context's
 *     stream.  These are what the solvers and benchmarks use.
 *
 * Data layouts are the reference's (AREC3D/AREC3D/arecImage.h:5-22):
 *   volume      x[nrays][ny]      (compact cylinder, y fastest)
 *   projections p[na][ny][nx]     (x fastest)
 *   cord        (ix, iz) per column, iz-major
 * Projector methods follow the reference's "pmethod" key: 1 = linear,
 * 2 = Kaiser-Bessel, 3 = SQ (square-voxel footprint).
 *
 * Every function returns 0 on success; a negative AREC_E* code for errors
 * this library adds; or the reference's own status code where the reference
 * defines one.  arec_last_error() describes the last failure of the calling
 * thread.
 */
#ifndef AREC3D_AMD_H
#define AREC3D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AREC_OK 0
#define AREC_EINVAL (-101)  /* bad argument */
#define AREC_EHIP (-102)    /* HIP runtime failure (no GPU, out of memory, ...) */
#define AREC_ESIZE (-103)   /* size beyond what the tables encode */
#define AREC_EGEOM (-104)   /* geometry self-check failed on the device */

#define AREC_LINEAR 1
#define AREC_KB 2
#define AREC_SQ 3

typedef struct arec_ctx arec_ctx;
typedef struct arec_geom arec_geom;

/* One line of a solver's residual history, as the reference prints it
 * (arecsirt.cpp:131-132, areccgls.cpp:197-198).  event: 0 iteration,
 * 1 SIRT restart with lam halved, 2 termination, 3 warm start rejected
 * (arec_sirt_warm with reject_above; no counterpart in the reference). */
typedef struct {
    int32_t iter;
    int32_t event;
    double relnrm;
    double rnorm;
} arec_hist;

const char *arec_last_error(void);
const char *arec_version(void);

/* ---- context, memory ------------------------------------------------ */
int arec_create(arec_ctx **ctx, int device);
int arec_destroy(arec_ctx *ctx);
int arec_synchronize(arec_ctx *ctx);
int arec_dev_alloc(arec_ctx *ctx, void **ptr, size_t bytes);
int arec_dev_free(arec_ctx *ctx, void *ptr);
int arec_memcpy_h2d(arec_ctx *ctx, void *dst, const void *src, size_t bytes);
int arec_memcpy_d2h(arec_ctx *ctx, void *dst, const void *src, size_t bytes);
int arec_memset(arec_ctx *ctx, void *dst, int value, size_t bytes);

/* ---- geometry ------------------------------------------------------- */
/* Cylinder mask of arecAllocateCylImage (arecImage.cpp:29-86).  Returns
 * nrays; fills 2*nrays ints when cord is not NULL.  Host only. */
int64_t arec_cyl_nrays(double radius, int32_t *cord);

/* Geometry of one projector: detector nx x ny, cylinder radius, FP32 angles
 * in radians.  Builds the per-angle constants, back-projection tiles and the
 * forward table on the device.  nx must equal round(2 radius + 1). */
int arec_geom_create(arec_ctx *ctx, arec_geom **geom, int method, int64_t nx, int64_t ny,
                     double radius, const float *angles, int64_t nangles);
int arec_geom_destroy(arec_geom *geom);
/* nrays, nnz = nrays*ny, nproj = nx*ny*na; any pointer may be NULL. */
int arec_geom_info(const arec_geom *geom, int64_t *nrays, int64_t *nnz, int64_t *nproj);
int64_t arec_geom_device_bytes(const arec_geom *geom);
/* Launch plan of the forward projector (no counterpart in the reference):
 * detector rows per lane (1, 2 or 4; 0 = default: 4 if ny >= 256 and
 * ny % 4 == 0, else 2 if ny >= 128 and even, else 1) and the minimum number
 * of workgroups (detector segments are added until the grid reaches it;
 * 0 = 1024).  Results do not depend on the plan.  The autotuner of the
 * Python package picks it by timing. */
int arec_geom_plan(arec_ctx *ctx, arec_geom *geom, int32_t rows_per_lane, int32_t min_groups);
int arec_geom_get_plan(const arec_geom *geom, int32_t *rows_per_lane, int32_t *min_groups, int32_t *nseg);
/* Forward projector design (no counterpart in the reference): 1 = detector
 * sweep (the default; arec_geom_plan sets its plan), 2 = depth strips (16-pixel
 * detector strips walked in depth order by 4 neighbouring angles in
 * lockstep; its table, ~1.125 x the sweep table, is built when design 2 is
 * selected, refused above 32 GiB, and released when design 1 is selected),
 * 3 = depth strips walked by angle pairs (each wave carries two angles, one
 * table entry per column of either, rows per lane 1 or 2; table ~1.13 x the
 * sweep table, same limits), 4 = staged strips (groups of 8 angles walk
 * chunks of 32 columns staged in LDS; rows per lane 1, 2 or 4; table about
 * 1.2 x the sweep table, refused where it would not leave room in free
 * device memory for a SIRT's work vectors).  Designs 2-4 replace the sweep
 * table while selected: it is released and rebuilt when design 1 is
 * selected again.  Results agree to FP32 summation order.
 * arec_geom_get_plan reports the selected design's rows per lane. */
int arec_geom_set_forward(arec_ctx *ctx, arec_geom *geom, int32_t design, int32_t rows_per_lane,
                          int32_t min_groups);
int32_t arec_geom_forward_design(const arec_geom *geom);
/* Back projector tile shape (no counterpart in the reference): detector rows
 * per workgroup x angles staged per LDS pass, one of 32 x 2 (the default),
 * 16 x 4 and 64 x 2; other pairs return AREC_EINVAL.  Every shape sums each
 * voxel's angles in the same order, so results are the same bits. */
int arec_geom_set_back(arec_ctx *ctx, arec_geom *geom, int32_t rows, int32_t angles);
int arec_geom_get_back(const arec_geom *geom, int32_t *rows, int32_t *angles);

/* ---- device-resident operators --------------------------------------- */
/* p = A x   (arecProject2D / _KB / _SQ, arecproject.cpp:19-99,351-437,540-648) */
int arec_project(arec_ctx *ctx, const arec_geom *geom, const float *x_dev, float *p_dev);
/* x = A^T p (arecBackProject2D / _KB / _SQ, arecproject.cpp:109-183,447-535,650-765) */
int arec_backproject(arec_ctx *ctx, const arec_geom *geom, const float *p_dev, float *x_dev);

/* SIRT (Landweber) of cyl_sirt / _KB / _SQ (arecsirt.cpp:11-191,193-373,
 * arecsirt_KB.cpp:11-191): x is zeroed, the restart chain halves lam.  The
 * whole iteration runs on the device; hist (host memory) receives at most
 * hist_cap lines. */
int arec_sirt(arec_ctx *ctx, const arec_geom *geom, const float *b_dev, float *x_dev, float lam,
              int32_t maxit, float tol, arec_hist *hist, int32_t hist_cap, int32_t *nhist,
              float *lam_out);

/* SIRT continued from the volume in x_dev (no counterpart in the
 * reference, whose SIRT always starts from zero, arecsirt.cpp:53-56).  The
 * relative residual keeps the reference's fixed denominator ||A^T b||; the
 * first printed line is the residual of the given volume; a restart still
 * zeroes x.  reject_above > 0 makes SIRT's residual decide on the given
 * volume: if the relative residual printed at pass check_pass (1 = the
 * given volume's own) is >= reject_above, or a restart would come before
 * that pass, the solve ends with a last history line of event 3 (x_dev then
 * holds the volume after check_pass - 1 updates; with check_pass = 1 it is
 * unchanged).  Used for the learned warm start. */
int arec_sirt_warm(arec_ctx *ctx, const arec_geom *geom, const float *b_dev, float *x_dev, float lam,
                   int32_t maxit, float tol, double reject_above, int32_t check_pass, arec_hist *hist,
                   int32_t hist_cap, int32_t *nhist, float *lam_out);

/* Batched SIRT over nprob independent problems (no counterpart in the
 * reference; each problem is one cyl_sirt* call, arecsirt.cpp:11-191), all
 * with the same projector method and the same maxit/tol, each with its own
 * geometry, data b_dev[k], volume x_dev[k] (zeroed) and lam[k].  One pass
 * runs every live problem's forward and back projection as persistent
 * work-queue launches (one forward launch per rows-per-lane class), then
 * the per-problem control; the pass is one hipGraph.  Every problem's
 * history equals a single arec_sirt call's bit for bit.  hist: nprob blocks
 * of hist_cap lines; nhist and lam_out: nprob entries (either may be NULL). */
int arec_sirt_batch(arec_ctx *ctx, int32_t nprob, const arec_geom *const *geoms, const float *const *b_dev,
                    float *const *x_dev, const float *lam, int32_t maxit, float tol, arec_hist *hist,
                    int32_t hist_cap, int32_t *nhist, float *lam_out);

/* CGLS of cyl_cgls / _KB / _SQ (areccgls.cpp:12-243,245-478,
 * areccgls_KB.cpp:12-240).  warm != 0 starts from x_dev (reference:
 * xcvol->data != NULL); else x is zeroed.  w_dev != NULL gives the weighted
 * solver of cyl_cgls*_stat (areccgls_stat.cpp:49-731) with weights w_dev
 * [na][ny][nx] (see arec_init_weights). */
int arec_cgls(arec_ctx *ctx, const arec_geom *geom, const float *b_dev, const float *w_dev,
              float *x_dev, int32_t warm, int32_t maxit, float tol, arec_hist *hist,
              int32_t hist_cap, int32_t *nhist);

/* init_weights (areccgls_stat.cpp:12-47): w = exp(-b) times the x-Tukey
 * window (alpha 0.5) of each view. */
int arec_init_weights(arec_ctx *ctx, const float *b_dev, float *w_dev, int64_t nx, int64_t ny,
                      int64_t nz);

/* soft_noneg (arecImage.cpp:812-826): negative values halved. */
int arec_soft_noneg(arec_ctx *ctx, float *x_dev, int64_t n);

/* Dot products for checks: FP64 sum of FP32 products (both compensated). */
int arec_dot(arec_ctx *ctx, const float *a_dev, const float *b_dev, int64_t n, double *out);

/* Average duration in ms of `reps` launches of one operator (which: 0
 * forward A, 1 back A^T), timed with HIP events on the context's stream. */
int arec_time_operator(arec_ctx *ctx, const arec_geom *geom, int which, const float *in_dev,
                       float *out_dev, int32_t reps, float *ms_per_launch);

/* dst = src, n floats, by a streaming copy kernel (also the reference
 * kernel for calibrating HBM byte counters). */
int arec_copy(arec_ctx *ctx, float *dst_dev, const float *src_dev, int64_t n);
/* x *= a (FP32) */
int arec_scale(arec_ctx *ctx, float *x_dev, int64_t n, float a);

/* Largest value of x (reduction), e.g. to scale noise to 1% of max(b). */
int arec_max(arec_ctx *ctx, const float *x_dev, int64_t n, float *out);

/* ---- alignment ------------------------------------------------------ */
/* Correlation maps c = IDFT_unnorm(conj(DFT x) DFT y) of nviews view pairs
 * [ny][nx] (ccorr2d, align2dstack.cpp:480-521). */
int arec_align_ccorr(arec_ctx *ctx, const float *x_dev, const float *y_dev, int32_t nx, int32_t ny,
                     int32_t nviews, float *c_dev);
/* Per view: ccorr2d(projection, measurement) then speak's sub-pixel peak
 * (arecCCImages, arecImage.cpp:384-434; speak/quad2dfit,
 * align2dstack.cpp:1059-1166).  sx_dev, sy_dev: nviews floats each. */
int arec_align_cc(arec_ctx *ctx, const float *proj_dev, const float *meas_dev, int32_t nx, int32_t ny,
                  int32_t nviews, float *sx_dev, float *sy_dev);
/* arec_align_cc without allocation (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results as arec_align_cc, bit for bit.  Its work buffers and
 * FFT twiddles live in a fixed 128 MiB device array that each GPU the
 * library runs on holds until the process ends; views run in batches that
 * fit it.  A view needing more than the array (nx*ny above about 4.7 M
 * pixels, or a side above 32768) is AREC_ESIZE.  It calls no hipMalloc or
 * hipFree; every call on one GPU shares the array, so calls that overlap on
 * one GPU race. */
int arec_align_cc_noalloc(arec_ctx *ctx, const float *proj_dev, const float *meas_dev, int32_t nx, int32_t ny,
                          int32_t nviews, float *sx_dev, float *sy_dev);
/* speak (align2dstack.cpp:1059-1105) on nviews maps [ny][nx]. */
int arec_align_peak(arec_ctx *ctx, const float *c_dev, int32_t nx, int32_t ny, int32_t nviews,
                    float *sx_dev, float *sy_dev);
/* Fourier shift of each view by (sign*sx, sign*sy) (fshift2d,
 * align2dstack.cpp:566-630); arecShiftImages (arecImage.cpp:653-671) is
 * sign = -1. */
int arec_align_shift(arec_ctx *ctx, float *data_dev, int32_t nx, int32_t ny, int32_t nviews,
                     const float *sx_dev, const float *sy_dev, float sign);

/* The reference's per-view helpers on host arrays (align2dstack.h:9,13,
 * 24-25), the drop-in's versions: ccorr2d (align2dstack.cpp:480-521), the
 * correlation map c of one view pair [ny][nx]; fshift2d (:566-630), one
 * image shifted by (+sx, +sy) into imageout (which may be imagein);
 * speak (:1059-1105), the sub-pixel peak of one map; quad2dfit
 * (:1125-1166), the 3x3 paraboloid fit (z[3*row + column]) with its peak
 * value, on the host (no GPU; with a singular fit the shifts are 0 and
 * *peakv is left as it was, as in the reference). */
int arec_h_ccorr2d(int32_t nx, int32_t ny, const float *x, const float *y, float *c);
int arec_h_fshift2d(int32_t nx, int32_t ny, const float *imagein, float *imageout, float sx, float sy);
int arec_h_speak(int32_t nx, int32_t ny, const float *data, float *xt, float *yt);
int arec_quad2dfit(const float *z, float *xsh, float *ysh, float *peakv);

/* ---- preprocessing of the measured views (AREC3D.cpp:81,105) ---------- */
/* -log of every value > 0, 0 elsewhere, in place (arecImageTakeLog,
 * arecImage.cpp:358-381; n values, e.g. nx*ny*nz of a view stack). */
int arec_take_log(arec_ctx *ctx, float *data_dev, int64_t n);
/* Moisan's periodic plus smooth decomposition of every view [nviews][ny][nx]
 * in place (periodicDecomposition, imagetools.cpp:17-114): the smooth part s
 * solves the periodic discrete Poisson equation with the boundary jumps as
 * source (2-D FFT), and each view loses W(d) s, W = exp(-d^2/50), d the
 * distance to the nearest border. */
int arec_periodic_decomposition(arec_ctx *ctx, float *views_dev, int32_t nx, int32_t ny, int32_t nviews);

/* Projection matching: the reference driver's alignment loop (AREC3D.cpp:
 * 140-165 initial reconstruction, :185-325 outer loops, translational
 * alignment only) with every buffer on the device.  views_dev [na][ny][nx]
 * are the measured views (never modified); x_dev receives the volume.  Per
 * outer loop: reproject, cc_passes rounds of arecCCImages + arecShiftImages
 * on a copy of the views, shift of the fresh views by the loop's summed
 * shifts, reconstruction (method 1: SIRT from zero, lam/maxit/tol as
 * cyl_sirt*; 2: CGLS continued from x, as cyl_cgls*; 3: SIRT continued from
 * x, no counterpart in the reference), soft_noneg.  sx, sy
 * (host, na floats each, may be NULL) receive the last loop's summed
 * shifts; loop_hist (host, outer lines, may be NULL) each reconstruction's
 * last history line. */
int arec_match(arec_ctx *ctx, const arec_geom *geom, const float *views_dev, float *x_dev, float lam,
               int32_t outer, int32_t maxit, float tol, int32_t cc_passes, int32_t method, float *sx,
               float *sy, arec_hist *loop_hist);
/* arec_match with flags.  Bit 0 (AREC_MATCH_X_CENTROID, no counterpart in
 * the reference) takes the x shifts from the views' centroids
 * (arec_align_centroid, once, before the initial reconstruction) and only
 * the y shifts from the cross-correlation passes.  Bit 1 (AREC_MATCH_ROTATE)
 * adds the driver's rotational step (its align = 2, AREC3D.cpp:203-204,
 * 229-243): from the second loop on, the fresh views are first rotated by
 * the summed angles (arecRotateImages_skew), and every third loop the
 * shifted views are matched in rotation against the reprojection
 * (arecRotCCImages with r2 = min(nx, ny - 10) / 2, then
 * arecRotateImages_skew_safe) before the reconstruction; rot (host, na
 * floats, may be NULL) receives the summed angles in degrees.  flags 0 is
 * arec_match. */
#define AREC_MATCH_X_CENTROID 1
#define AREC_MATCH_ROTATE 2
int arec_match_ex(arec_ctx *ctx, const arec_geom *geom, const float *views_dev, float *x_dev, float lam,
                  int32_t outer, int32_t maxit, float tol, int32_t cc_passes, int32_t method, int32_t flags,
                  float *sx, float *sy, float *rot, arec_hist *loop_hist);
/* x shift of every view from its centroid (no counterpart in the reference;
 * parallel beam, object inside the field of view): the view's centroid
 * along the detector minus the rotation axis (the centroid of the projected
 * uniform cylinder), with the unobservable a cos(t) + b sin(t) (a
 * translation of the object in the plane) removed by least squares.  sx:
 * host, na floats.  AREC_EINVAL if a view has no positive mass. */
int arec_align_centroid(arec_ctx *ctx, const arec_geom *geom, const float *views_dev, float *sx);

/* Rotational alignment of view stacks [nviews][ny][nx] (views of at least
 * 2x2 and below 2^31 pixels, at most 65535 a call).  Angles are host arrays
 * of nviews floats, in degrees; a positive angle rotates x towards -y.
 * arec_rot_cc (arecRotCCImages, arecImage.cpp:555-599): the rotation of each
 * view of images2 against images1 from the 1-D correlation of their radial
 * profiles (L = (int)(2 pi r2) directions, radii below r2, 1 <= r2 <=
 * min(nx, ny)/2), wrapped to (-180, 180], 0 beyond 45 degrees.
 * arec_rotate_views: rule AREC_ROT_SKEW is arecRotateImages_skew
 * (arecImage.cpp:693-741: angles wrapped by 360 and updated, 0 and no
 * rotation beyond 45, bilinear rotate2d beyond 10, else three circular
 * shears); AREC_ROT_BILINEAR is arecRotateImages (:673-691, rotate2d for
 * every view).  arec_rotate_views_safe (arecRotateImages_skew_safe,
 * :743-794): the three shears (angles within +-90), kept where the
 * Tukey-windowed correlation coefficient with ref rises by more than 1e-5,
 * the angle then halved; else the view restored and its angle 0; kept and
 * mean_cc (may be NULL) get the count kept and the mean of the chosen
 * coefficients. */
#define AREC_ROT_SKEW 0
#define AREC_ROT_BILINEAR 1
int arec_rot_cc(arec_ctx *ctx, const float *images1_dev, const float *images2_dev, int32_t nx, int32_t ny,
                int32_t nviews, int32_t r2, float *angles);
int arec_rotate_views(arec_ctx *ctx, float *views_dev, int32_t nx, int32_t ny, int32_t nviews, float *angles,
                      int32_t rule);
int arec_rotate_views_safe(arec_ctx *ctx, float *views_dev, const float *ref_dev, int32_t nx, int32_t ny,
                           int32_t nviews, float *angles, int32_t *kept, double *mean_cc);

/* ---- synthetic inputs (no counterpart in the reference) --------------- */
/* Sum of ellipsoids at the voxel centres of the compact cylinder of the
 * given radius and height ny, zero outside `support` (radius about the
 * axis).  params: 16 doubles per ellipsoid (cx, cy, cz, 3x3 world->body
 * rotation, semi-axes ax, ay, az, density); drawn by arec3d_amd.phantom. */
int arec_phantom(arec_ctx *ctx, double radius, int64_t ny, double support, const double *params,
                 int32_t ne, float *x_dev);
/* mode 0: b += sigma N(0,1).  mode 1: Beer-Lambert counts at i0 photons per
 * ray with attenuation scale, b = -ln(counts/i0).  Counter-based RNG: the
 * result depends only on (seed, index). */
int arec_add_noise(arec_ctx *ctx, float *b_dev, int64_t n, uint64_t seed, int32_t mode,
                   double sigma, double i0, double scale);

/* ---- learned warm start (no counterpart in the reference) ------------- */
/* 3-level 3-D U-Net (16/32/64 channels, 3x3x3 convolutions, ReLU, BF16
 * weights/activations, FP32 accumulation) on volumes D x H x W (multiples of
 * 4; training also needs W <= 128 and (D/4)(H/4) divisible by 128).  y = x +
 * net(x*scale)/scale on FP32 cubes [D][H][W]. */
typedef struct arec_unet arec_unet;
int arec_unet_create(arec_ctx *ctx, arec_unet **net, int32_t D, int32_t H, int32_t W,
                     int32_t training, uint64_t seed);
int arec_unet_destroy(arec_unet *net);
int64_t arec_unet_nparams(const arec_unet *net);
int arec_unet_get_params(arec_unet *net, float *host);
int arec_unet_set_params(arec_unet *net, const float *host);
int arec_unet_set_scale(arec_unet *net, float scale);
int arec_unet_forward(arec_unet *net, const float *x_cube, float *y_cube);
/* One Adam step on the mean-squared error to t_cube; *loss before the step. */
int arec_unet_train_step(arec_unet *net, const float *x_cube, const float *t_cube, float lr,
                         double *loss);
/* Loss and its gradient with respect to every parameter (host array). */
int arec_unet_gradient(arec_unet *net, const float *x_cube, const float *t_cube, double *loss,
                       float *grad_host);
/* One MFMA convolution (FP32 in/out, BF16 inside), for checks: input
 * channels [0,C1) from in1 [D][H][W][C1], [C1,C1+C2) from in2 at half
 * resolution (nearest upsampling); w [27][Cout][C1+C2]. */
int arec_conv3_check(arec_ctx *ctx, const float *in1, int32_t C1, const float *in2, int32_t C2,
                     const float *w, const float *bias, float *out, int32_t Cout, int32_t D,
                     int32_t H, int32_t W, int32_t relu);
int arec_conv3_wgrad_check(arec_ctx *ctx, const float *in1, int32_t C1, const float *in2,
                           int32_t C2, const float *dz, int32_t Cout, int32_t D, int32_t H,
                           int32_t W, float *dw);
/* compact cylinder [nrays][ny] <-> bounding cube [ny][n][n], n = round(2r+1) */
int arec_cyl_to_cube(arec_ctx *ctx, const float *x_dev, double radius, int64_t ny, float *cube_dev);
int arec_cube_to_cyl(arec_ctx *ctx, const float *cube_dev, double radius, int64_t ny, float *x_dev);

/* ---- host-pointer drop-in layer -------------------------------------- */
/* arecProject2D* (arecproject.h:55-62) with the cylinder given by radius and
 * nrays from arecAllocateCylImage.  Status codes as the reference: 1 not a
 * cylinder, 2 inconsistent dimension. */
int arec_h_project(int method, double radius, int32_t nx, int32_t ny, const float *x,
                   const float *angles, int32_t nangles, float *p);
/* arecBackProject2D*: status 1 when nz != nangles, -1 size mismatch. */
int arec_h_backproject(int method, double radius, int32_t nx, int32_t ny, int32_t nz,
                       const float *p, const float *angles, int32_t nangles, float *x);
/* cyl_sirt* / cyl_cgls* on host arrays; x receives nrays*ny values (the
 * caller sizes it with arec_cyl_nrays(0.5*(nx-1))).  solver: 1 SIRT, 2 CGLS,
 * 3 weighted CGLS.  For CGLS warm != 0 uses x as the start. */
int arec_h_solve(int solver, int method, const float *b, int32_t nx, int32_t ny, int32_t nz,
                 const float *angles, float lam, int32_t maxit, float tol, int32_t warm,
                 float *x, arec_hist *hist, int32_t hist_cap, int32_t *nhist, float *lam_out);

/* arecCCImages (arecImage.cpp:384-434) on host stacks: -1 when the two
 * stacks differ in size. */
int arec_h_cc_images(int32_t nx1, int32_t ny1, int32_t nz1, const float *images1, int32_t nx2,
                     int32_t ny2, int32_t nz2, const float *images2, float *sx, float *sy);
/* arecShiftImages (arecImage.cpp:653-671): each view shifted by (-sx, -sy). */
int arec_h_shift_images(int32_t nx, int32_t ny, int32_t nz, float *data, const float *sx,
                        const float *sy);
/* arecImageTakeLog / periodicDecomposition on host data (upload, compute,
 * download; the drop-in's versions). */
int arec_h_take_log(float *data, int64_t n);
int arec_h_periodic_decomposition(int32_t nx, int32_t ny, int32_t nz, float *data);
/* arecRotCCImages / arecRotateImages(_skew) / arecRotateImages_skew_safe on
 * host stacks (the drop-in's versions): -1 when the two stacks differ in
 * size, as the reference. */
int arec_h_rot_cc_images(int32_t nx1, int32_t ny1, int32_t nz1, const float *images1, int32_t nx2, int32_t ny2,
                         int32_t nz2, const float *images2, int32_t r2, float *angles);
int arec_h_rotate_images(int32_t nx, int32_t ny, int32_t nz, float *data, float *angles, int32_t rule);
int arec_h_rotate_images_safe(int32_t nx, int32_t ny, int32_t nz, float *data, const float *ref, float *angles,
                              int32_t *kept, double *mean_cc);

/* MRC files (arecImageIO.cpp:10-173, header arecImageIO.h:5-63), host only:
 * a 1024-byte header, then nx*ny*nz values, x fastest, host byte order.
 * arec_mrc_read_header (arecReadImageHeader) reads the sizes and the mode;
 * any pointer may be NULL.  arec_mrc_read (arecReadImage) reads mode 2
 * (float32) and modes 1 and 6 (16-bit, converted as signed) into data, which
 * holds capacity floats (AREC_ESIZE when the image is larger); -1 for a file
 * that fails to open or read or has another mode, as the reference.
 * arec_mrc_write (arecWriteImage) writes mode 2 with the reference's header:
 * cell 1x1x1 at 90 degrees, axes 1/2/3, amin/amax/amean of the data, "MAP",
 * rms 1, no labels; the fields the reference leaves unset are zero except
 * machinestamp, which is the little-endian stamp 0x4444. */
int arec_mrc_read_header(const char *path, int32_t *nx, int32_t *ny, int32_t *nz, int32_t *mode);
int arec_mrc_read(const char *path, float *data, int64_t capacity);
int arec_mrc_write(const char *path, int32_t nx, int32_t ny, int32_t nz, const float *data);

#ifdef __cplusplus
}
#endif
#endif
