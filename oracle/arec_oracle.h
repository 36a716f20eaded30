/*
 * CPU oracle for the AREC3D reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * is compared against; it is linked or loaded only by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py.  The product never calls it.
 *
 * It restates the reference's CPU algorithms loop for loop, including their
 * floating-point semantics, so that on identical inputs it produces the
 * reference's numbers:
 *   - the FP32 entry points (suffix _f) store and accumulate like the
 *     reference (FP32 storage, FP64 geometry/weights, every "+=" rounded to
 *     FP32);
 *   - the FP64 entry points (suffix _d) run the same algorithm with FP64
 *     storage and accumulation and are the accuracy oracle.
 * Every function cites the reference file:line it follows.  Build flags must
 * keep -ffp-contract=off: the reference's own products and sums are separate
 * roundings.
 *
 * Projector method numbers follow the reference's "pmethod" input key
 * (AREC3D/AREC3D/AREC3D.cpp:145-169): 1 = linear, 2 = Kaiser-Bessel, 3 = SQ.
 */
#ifndef AREC_ORACLE_H
#define AREC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Cylinder mask: AREC3D/AREC3D/arecImage.cpp:29-86.  Returns nrays; when
 * cord is not NULL it receives 2*nrays ints (ix, iz) pairs, iz-major. */
int orc_cyl_geometry(double radius, int height, int *cord);

/* Projectors.  x: [nrays][ny] (y fastest); p: [na][ny][nx] (x fastest).
 * nx is the detector width = cylinder width round(2r+1).  Return the
 * reference's status codes. */
int orc_project_f(int method, const int *cord, int nrays, double radius, int nx, int ny,
                  const float *x, const float *angles, int na, float *p);
int orc_backproject_f(int method, const int *cord, int nrays, double radius, int nx, int ny,
                      const float *p, const float *angles, int na, float *x);
int orc_project_d(int method, const int *cord, int nrays, double radius, int nx, int ny,
                  const double *x, const float *angles, int na, double *p);
int orc_backproject_d(int method, const int *cord, int nrays, double radius, int nx, int ny,
                      const double *p, const float *angles, int na, double *x);

/* One line of the solvers' printed residual history. */
typedef struct {
    int iter;      /* iteration number as printed */
    int event;     /* 0 normal, 1 restart (SIRT lam halved), 2 terminate */
    double relnrm; /* rnorm / bnorm as printed */
    double rnorm;
} orc_hist;

/* SIRT (Landweber), AREC3D/AREC3D/arecsirt.cpp:11-191 (and its _SQ/_KB
 * copies).  b: [na][ny][nx].  x: [nrays][ny], zeroed on entry like the
 * reference.  hist receives at most hist_cap lines; *nhist the count.
 * *lam_out receives the final lam (after restarts). */
int orc_sirt_f(int method, const float *b, int nx, int ny, int na, const float *angles,
               float lam, int maxit, float tol, float *x, orc_hist *hist, int hist_cap,
               int *nhist, float *lam_out);
int orc_sirt_d(int method, const double *b, int nx, int ny, int na, const float *angles,
               double lam, int maxit, double tol, double *x, orc_hist *hist, int hist_cap,
               int *nhist, double *lam_out);

/* CGLS, AREC3D/AREC3D/areccgls.cpp:12-243 (and _SQ/_KB).  warm != 0 uses x
 * as the starting volume (reference: xcvol->data != NULL); otherwise x is
 * zeroed.  Method 2 (KB) uses cyl_cgls_KB's integer radius nx/2
 * (AREC3D/AREC3D/areccgls_KB.cpp:39): even nx fails with status -1. */
int orc_cgls_f(int method, const float *b, int nx, int ny, int na, const float *angles,
               int maxit, float tol, int warm, float *x, orc_hist *hist, int hist_cap,
               int *nhist);
int orc_cgls_d(int method, const double *b, int nx, int ny, int na, const float *angles,
               int maxit, double tol, int warm, double *x, orc_hist *hist, int hist_cap,
               int *nhist);

/* The same projectors and solvers with the GPU library's per-tap arithmetic
 * (weights rounded to FP32, y = fmaf(w, v, y)) in the reference's loop order:
 * an accuracy probe, not a restatement of the reference. */
int orc_project_g(int method, const int *cord, int nrays, double radius, int nx, int ny,
                  const float *x, const float *angles, int na, float *p);
int orc_backproject_g(int method, const int *cord, int nrays, double radius, int nx, int ny,
                      const float *p, const float *angles, int na, float *x);
int orc_sirt_g(int method, const float *b, int nx, int ny, int na, const float *angles,
               float lam, int maxit, float tol, float *x, orc_hist *hist, int hist_cap,
               int *nhist, float *lam_out);
int orc_cgls_g(int method, const float *b, int nx, int ny, int na, const float *angles,
               int maxit, float tol, int warm, float *x, orc_hist *hist, int hist_cap,
               int *nhist);

/* Weighted CGLS, AREC3D/AREC3D/areccgls_stat.cpp:12-731, weights from
 * init_weights (:12-47) with tukey_filter_x_inplace (imagetools.cpp:369-391).
 * Integer radius nx/2 as in the reference: even nx fails with its status. */
int orc_init_weights_f(const float *b, int nx, int ny, int na, float *w);
int orc_cgls_stat_f(int method, const float *b, int nx, int ny, int na, const float *angles,
                    int maxit, float tol, int warm, float *x, orc_hist *hist, int hist_cap,
                    int *nhist);
int orc_cgls_stat_d(int method, const double *b, int nx, int ny, int na, const float *angles,
                    int maxit, double tol, int warm, double *x, orc_hist *hist, int hist_cap,
                    int *nhist);

/* Alignment (FP64 arithmetic, FP32 interfaces like the reference). */
/* c = IDFT_unnormalised(conj(DFT x) * DFT y): AREC3D/AREC3D/align2dstack.cpp:480-521 */
void orc_ccorr2d(int nx, int ny, const float *x, const float *y, double *c);
/* AREC3D/AREC3D/align2dstack.cpp:1125-1166 */
void orc_quad2dfit(const float *Z, float *xsh, float *ysh, float *peakv);
/* AREC3D/AREC3D/align2dstack.cpp:1059-1105 on a float map */
void orc_speak(int nx, int ny, const float *data, float *xt, float *yt);
/* AREC3D/AREC3D/align2dstack.cpp:566-630 */
void orc_fshift2d(int nx, int ny, const float *in, float *out, float sx, float sy);
/* AREC3D/AREC3D/arecImage.cpp:384-434 (one view per z) */
void orc_cc_images(int nx, int ny, int nz, const float *proj, const float *meas, float *sx,
                   float *sy);
/* AREC3D/AREC3D/arecImage.cpp:653-671 */
void orc_shift_images(int nx, int ny, int nz, float *data, const float *sx, const float *sy);
/* AREC3D/AREC3D/arecImage.cpp:812-826 */
void orc_soft_noneg(float *x, long n);
void orc_take_log(float *x, long n);
void orc_periodic_smooth(int nx, int ny, const float *u, float *s);
void orc_periodic_decomposition(int nx, int ny, int nz, float *data);

/* Rotational alignment: AREC3D/AREC3D/arecImage.cpp:489-599 (radial profiles,
 * 1-D correlation peak, angles), :693-794 (rotations), imagetools.cpp:
 * 244-291 (shears), :332-367, 415-447 (windowed correlation coefficient),
 * align2dstack.cpp:437-475 (ccorr1d), :774-822 (rotate2d). */
void orc_radial_profile(int nx, int ny, const float *img, int length, int R, float *out);
void orc_ccorr1d(int n, const float *x, const float *y, double *c);
float orc_findpeak1d(int n, const float *c);
void orc_rot_cc_images(int nx, int ny, int nz, const float *img1, const float *img2, int r2, float *angles);
void orc_shear_x(float *img, float shear, int nx, int ny);
void orc_shear_y(float *img, float shear, int nx, int ny);
void orc_rotate2d(int nx, int ny, const float *in, float *out, float deg);
void orc_rotate_skew(int nx, int ny, int nz, float *data, float *angles);
double orc_tukey_ncc(const float *a, const float *b, int nx, int ny);
int orc_rotate_skew_safe(int nx, int ny, int nz, float *data, const float *ref, float *angles, double *mean_cc);

#ifdef __cplusplus
}
#endif
#endif
