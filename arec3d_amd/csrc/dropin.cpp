// libarec3d_dropin.so: the reference's C++ symbols (include/arec3d_dropin.h)
// on top of the C ABI of libarec3d_amd.so.  Host code only; every projector
// and solver call runs on the GPU.
#include "arec3d_dropin.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "arec3d_amd.h"

namespace {

int one_rank(MPI_Comm comm) {
    int init = 0, size = 1;
    MPI_Initialized(&init);
    if (init) MPI_Comm_size(comm, &size);
    if (size != 1) {
        fprintf(stderr, "arec3d_dropin: one GPU per process, communicator of size %d not supported\n", size);
        return 0;
    }
    return 1;
}

int is_rank0(MPI_Comm comm) {
    int init = 0, rank = 0;
    MPI_Initialized(&init);
    if (init) MPI_Comm_rank(comm, &rank);
    return rank == 0;
}

int project(int method, arecImage cylvol, float *angles, int nangles, arecImage *projstack) {
    if (!cylvol.is_cyl) {
        fprintf(stderr, "invalid input volume format. Must be cylindrical!\n");
        return 1;
    }
    if (projstack->nx != cylvol.nx || projstack->ny != cylvol.ny) {
        fprintf(stderr, "arecProject2D: inconsistent dimension!\n");
        return 2;
    }
    int rc = arec_h_project(method, cylvol.radius, cylvol.nx, cylvol.ny, cylvol.data, angles, nangles,
                            projstack->data);
    if (rc < 0) fprintf(stderr, "arec3d_amd: %s\n", arec_last_error());
    return rc;
}

int backproject(int method, arecImage images, float *angles, int nangles, arecImage *bvol) {
    if (images.nz != nangles) return 1;
    if (bvol->nx != images.nx || bvol->ny != images.ny) {
        fprintf(stderr, "arecBackProject2D: image and volume sizes don't match\n");
        return -1;
    }
    if (!bvol->is_cyl) {
        fprintf(stderr, "arecBackProject2D: cylindrical volume not properly set!\n");
        return -1;
    }
    int rc = arec_h_backproject(method, bvol->radius, images.nx, images.ny, images.nz, images.data,
                                angles, nangles, bvol->data);
    if (rc < 0) fprintf(stderr, "arec3d_amd: %s\n", arec_last_error());
    return rc;
}

// cyl_sirt*/cyl_cgls*: allocate xvol with calloc when xvol->data == NULL
// (arecsirt.cpp:46-52, areccgls.cpp:87-101), run on the GPU, then print the
// residual history the way the reference does.
int solve(int solver, int method, MPI_Comm comm, arecImage images, float *angles, arecImage *xvol,
          float lam, int maxit, float tol) {
    if (!one_rank(comm)) return -2;
    int nx = images.nx, ny = images.ny, nz = images.nz;
    double radius = 0.5 * (nx - 1);
    if ((solver == 2 && method == AREC_KB) || solver == 3) radius = (double)(nx / 2);
    int warm = 0;
    if (xvol->data == NULL) {
        if (arecAllocateCylImage(xvol, radius, ny) != 0) {
            fprintf(stderr, "failed to allocate xcvol\n");
            return -1;
        }
    } else {
        warm = solver != 1;
    }
    if (xvol->nx != nx) {
        fprintf(stderr, "arecBackProject2D: image and volume sizes don't match\n");
        return -1;
    }
    const int cap = maxit * 22 + 64;
    std::vector<arec_hist> hist(cap);
    int nh = 0;
    float lam_out = lam;
    int rc = arec_h_solve(solver, method, images.data, nx, ny, nz, angles, lam, maxit, tol, warm,
                          xvol->data, hist.data(), cap, &nh, &lam_out);
    if (rc < 0 && rc <= AREC_EINVAL) fprintf(stderr, "arec3d_amd: %s\n", arec_last_error());
    if (rc == 0 && is_rank0(comm)) {
        printf("nrays = %d, nnz = %d\n", xvol->nrays, xvol->nnz);
        float l = lam;
        for (int i = 0; i < nh && i < cap; i++) {
            const arec_hist &h = hist[i];
            if (solver == 1) {
                printf("iter = %3d, rnorm / bnorm = %11.3e, rnorm = %11.3e\n", h.iter, h.relnrm, h.rnorm);
                if (h.event == 1) {
                    l /= 2.0f;
                    printf("reducing lam to %11.3e, restarting\n", l);
                }
            } else if (h.iter == 0) {
                printf("CG iter = 0, resnrm = %11.3e\n", h.rnorm);
            } else {
                printf("CG iter = %3d, rnorm / bnorm = %11.3e, rnorm = %11.3e\n", h.iter, h.relnrm, h.rnorm);
            }
            if (h.event == 2) printf("Terminating with rnorm/bnorm = %11.3e, tol = %11.3e\n", h.relnrm, tol);
        }
    }
    return rc;
}

}  // namespace

// ---- image helpers (arecImage.cpp) -------------------------------------

int arecAllocateCBImage(arecImage *cbimage, int nx, int ny, int nz) {
    cbimage->nx = nx;
    cbimage->ny = ny;
    cbimage->nz = nz;
    cbimage->data = (float *)calloc((size_t)nx * ny * nz, sizeof(float));
    cbimage->is_cyl = 0;
    cbimage->cord = NULL;
    return cbimage->data ? 0 : -1;
}

int arecAllocateCylImage(arecImage *cylimage, double radius, int height) {
    int n = (int)std::lround(2 * radius + 1);
    int64_t nrays = arec_cyl_nrays(radius, NULL);
    cylimage->nx = n;
    cylimage->ny = height;
    cylimage->nz = n;
    cylimage->radius = radius;
    cylimage->height = height;
    cylimage->nrays = (int)nrays;
    cylimage->nnz = (int)(nrays * height);
    cylimage->data = (float *)calloc((size_t)nrays * height, sizeof(float));
    cylimage->cord = (int *)calloc((size_t)nrays * 2, sizeof(int));
    if (cylimage->cord) arec_cyl_nrays(radius, cylimage->cord);
    cylimage->is_cyl = 1;
    return cylimage->data ? 0 : -1;
}

// cube[iz][iy][ix] -> cyl[j][iy] (arecImage.cpp:88-161).  As the reference:
// the cube's own nx, ny, nz are walked, iz outer, ix inner; a column is in
// the cylinder when (ix - c)^2 + (iz - c)^2 <= (int)(r^2) with the centre
// c = (int)r (for a half-integer radius a different set than
// arecAllocateCylImage's truncated mask, which the count check then
// reports); every kept column takes the cube's ny values and rewrites cord.
// Status 1 for an unset cylinder, -1 unallocated, -2 when the counts differ
// from the cylinder's nnz / nrays.  Where the reference would write past the
// cylinder's buffers (more columns in the cube than the cylinder holds) the
// extra values are counted but not stored.
int ImageCB2Cyl(arecImage cbimage, arecImage cylimage, int, int, double, int) {
    if (!cylimage.data) {
        fprintf(stderr, "ImageCB2Cyl: cylimage not properly allocated\n");
        return -1;
    }
    const double radius = cylimage.radius;
    const int height = cylimage.height, nnz0 = cylimage.nnz, nrays0 = cylimage.nrays;
    if (nnz0 <= 0 || nrays0 <= 0 || radius <= 0 || height <= 0) {
        fprintf(stderr, "ImageCB2Cy: cylimage not properly created!\n");
        return 1;
    }
    if (!cylimage.cord) {
        fprintf(stderr, "ImageCB2Cyl: cylimage not properly allocated\n");
        return -1;
    }
    const int nx = cbimage.nx, ny = cbimage.ny, nz = cbimage.nz;
    const int cent = (int)radius, r2 = (int)(radius * radius);
    int64_t nnz = 0, nrays = 0;
    for (int iz = 0; iz < nz; iz++) {
        const int zs = iz - cent;
        for (int ix = 0; ix < nx; ix++) {
            const int xs = ix - cent;
            if (xs * xs + zs * zs > r2) continue;
            for (int iy = 0; iy < ny; iy++, nnz++)
                if (nnz < nnz0) cylimage.data[nnz] = cbimage.data[((size_t)iz * ny + iy) * nx + ix];
            if (nrays < nrays0) {
                cylimage.cord[2 * nrays] = ix;
                cylimage.cord[2 * nrays + 1] = iz;
            }
            nrays++;
        }
    }
    if (nnz != nnz0 || nrays != nrays0) {
        fprintf(stderr, "ImageCB2Cyl: inconsistency!\n");
        fprintf(stderr, "nnz, nnz0, nrays, nrays0 = %lld, %d, %lld, %d\n", (long long)nnz, nnz0, (long long)nrays,
                nrays0);
        return -2;
    }
    return 0;
}

// cyl[j][iy] -> cube[iz][iy][ix] (arecImage.cpp:163-232): status -1 for
// unallocated data, 1 for sizes that disagree with the cube or the
// cylinder, -2 for a column outside the cube or more values than nnz (the
// reference's check lets one value past nnz through; here none is read).
int ImageCyl2CB(arecImage cylimage, arecImage cbimage, int nx, int ny, int nz) {
    if (!cylimage.data) {
        fprintf(stderr, "ImageCyl2CB: cylimage not properly allocated!\n");
        return -1;
    }
    if (nx != cbimage.nx || ny != cbimage.ny || nz != cbimage.nz) {
        fprintf(stderr, "ImageCyl2CB: inconsistent dimension\n");
        return 1;
    }
    if (nx != 2 * cylimage.radius + 1 || nz != 2 * cylimage.radius + 1 || ny != cylimage.height) {
        fprintf(stderr, "ImageCyl2CB: inconsistent dimension\n");
        return 1;
    }
    if (!cbimage.data) {
        fprintf(stderr, "ImageCyl2CB: cbimage not properly allocated!\n");
        return -1;
    }
    int64_t nnz = 0;
    for (int j = 0; j < cylimage.nrays; j++) {
        const int ix = cylimage.cord[2 * j], iz = cylimage.cord[2 * j + 1];
        if (!(iz < nz && ix < nx)) {
            printf("ImageCyl2CB: ix, iz, out of bound\n");
            printf("ImageCyl2CB: ix = %d, iz = %d \n", ix, iz);
            return -2;
        }
        for (int iy = 0; iy < ny; iy++, nnz++) {
            if (nnz >= cylimage.nnz) {
                printf("ImageCyl2CB: iz out of bound, nnz = %lld, nnz0 = %d\n", (long long)nnz, cylimage.nnz);
                return -2;
            }
            cbimage.data[((size_t)iz * ny + iy) * nx + ix] = cylimage.data[nnz];
        }
    }
    return 0;
}

void arecImageFree(arecImage *image) {
    if (image->is_cyl) {
        free(image->cord);
        image->cord = NULL;
    }
    free(image->data);
    image->data = NULL;
}

void reset(arecImage *image) {
    size_t n = image->is_cyl ? (size_t)image->nnz : (size_t)image->nx * image->ny * image->nz;
    memset(image->data, 0, n * sizeof(float));
}

void soft_noneg(arecImage *image) {
    size_t n = image->is_cyl ? (size_t)image->nnz : (size_t)image->nx * image->ny * image->nz;
    for (size_t i = 0; i < n; i++)
        if (image->data[i] < 0.0f) image->data[i] *= 0.5f;
}

// ---- projectors ----------------------------------------------------------

int arecProject2D(arecImage c, float *a, int n, arecImage *p) { return project(AREC_LINEAR, c, a, n, p); }
int arecProject2D_KB(arecImage c, float *a, int n, arecImage *p) { return project(AREC_KB, c, a, n, p); }
int arecProject2D_SQ(arecImage c, float *a, int n, arecImage *p) { return project(AREC_SQ, c, a, n, p); }
int arecBackProject2D(arecImage i, float *a, int n, arecImage *b) { return backproject(AREC_LINEAR, i, a, n, b); }
int arecBackProject2D_KB(arecImage i, float *a, int n, arecImage *b) { return backproject(AREC_KB, i, a, n, b); }
int arecBackProject2D_SQ(arecImage i, float *a, int n, arecImage *b) { return backproject(AREC_SQ, i, a, n, b); }

// ---- solvers ---------------------------------------------------------------

int cyl_sirt(MPI_Comm c, arecImage i, float *a, arecImage *x, float l, int m, float t) { return solve(1, AREC_LINEAR, c, i, a, x, l, m, t); }
int cyl_sirt_KB(MPI_Comm c, arecImage i, float *a, arecImage *x, float l, int m, float t) { return solve(1, AREC_KB, c, i, a, x, l, m, t); }
int cyl_sirt_SQ(MPI_Comm c, arecImage i, float *a, arecImage *x, float l, int m, float t) { return solve(1, AREC_SQ, c, i, a, x, l, m, t); }
int cyl_cgls(MPI_Comm c, arecImage i, float *a, arecImage *x, float l, int m, float t) { return solve(2, AREC_LINEAR, c, i, a, x, l, m, t); }
int cyl_cgls_KB(MPI_Comm c, arecImage i, float *a, arecImage *x, float l, int m, float t) { return solve(2, AREC_KB, c, i, a, x, l, m, t); }
int cyl_cgls_SQ(MPI_Comm c, arecImage i, float *a, arecImage *x, float l, int m, float t) { return solve(2, AREC_SQ, c, i, a, x, l, m, t); }
int cyl_cgls_stat(MPI_Comm c, arecImage i, float *a, arecImage *x, float l, int m, float t) { return solve(3, AREC_LINEAR, c, i, a, x, l, m, t); }
int cyl_cgls_KB_stat(MPI_Comm c, arecImage i, float *a, arecImage *x, float l, int m, float t) { return solve(3, AREC_KB, c, i, a, x, l, m, t); }
int cyl_cgls_SQ_stat(MPI_Comm c, arecImage i, float *a, arecImage *x, float l, int m, float t) { return solve(3, AREC_SQ, c, i, a, x, l, m, t); }

// init_weights (areccgls_stat.cpp:12-47), on the GPU.
int init_weights(MPI_Comm, arecImage projections, arecImage *weights) {
    int nx = projections.nx, ny = projections.ny, nz = projections.nz;
    if (arecAllocateCBImage(weights, nx, ny, nz) != 0) return -1;
    arec_ctx *ctx = NULL;
    if (arec_create(&ctx, 0) != 0) return -1;
    size_t n = (size_t)nx * ny * nz;
    void *b = NULL, *w = NULL;
    int rc = arec_dev_alloc(ctx, &b, n * 4);
    if (!rc) rc = arec_dev_alloc(ctx, &w, n * 4);
    if (!rc) rc = arec_memcpy_h2d(ctx, b, projections.data, n * 4);
    if (!rc) rc = arec_init_weights(ctx, (const float *)b, (float *)w, nx, ny, nz);
    if (!rc) rc = arec_memcpy_d2h(ctx, weights->data, w, n * 4);
    arec_dev_free(ctx, b);
    arec_dev_free(ctx, w);
    arec_destroy(ctx);
    return rc ? -1 : 0;
}

// ---- alignment ---------------------------------------------------------------

int arecCCImages(MPI_Comm comm, arecImage images1, arecImage images2, float *sx, float *sy) {
    (void)comm;
    int rc = arec_h_cc_images(images1.nx, images1.ny, images1.nz, images1.data, images2.nx,
                              images2.ny, images2.nz, images2.data, sx, sy);
    if (rc == -1) {
        fprintf(stderr, "arecCCImages: mismatch in dimension:\n");
        fprintf(stderr, "nx1,nx2,ny1,ny2,nz1,nz2=%d, %d, %d, %d, %d, %d\n", images1.nx, images2.nx,
                images1.ny, images2.ny, images1.nz, images2.nz);
    }
    return rc;
}

void arecShiftImages(arecImage images, float *sx, float *sy) {
    arec_h_shift_images(images.nx, images.ny, images.nz, images.data, sx, sy);
}

// align2dstack.h:9,13,24-25: one view at a time, as the reference's callers
// use them.
void ccorr2d(int nx, int ny, float *x, float *y, float *c) {
    if (arec_h_ccorr2d(nx, ny, x, y, c)) fprintf(stderr, "ccorr2d: %s\n", arec_last_error());
}

void fshift2d(int nx, int ny, float *imagein, float *imageout, float sx, float sy) {
    if (arec_h_fshift2d(nx, ny, imagein, imageout, sx, sy)) fprintf(stderr, "fshift2d: %s\n", arec_last_error());
}

void quad2dfit(float *Z, float *XSH, float *YSH, float *PEAKV) { arec_quad2dfit(Z, XSH, YSH, PEAKV); }

void speak(int nx, int ny, float *data, float *xt, float *yt) {
    if (arec_h_speak(nx, ny, data, xt, yt)) fprintf(stderr, "speak: %s\n", arec_last_error());
}

// ---- rotational alignment -------------------------------------------------------

// arecImage.cpp:555-599: per-view angles (degrees) of images2 against images1.
int arecRotCCImages(MPI_Comm comm, arecImage images1, arecImage images2, int r2, float *angles) {
    (void)comm;
    int rc = arec_h_rot_cc_images(images1.nx, images1.ny, images1.nz, images1.data, images2.nx, images2.ny,
                                  images2.nz, images2.data, r2, angles);
    if (rc == -1) {
        fprintf(stderr, "arecRotCCImages2: mismatch in dimension:\n");
        fprintf(stderr, "(%d,%d,%d) (%d,%d,%d)\n", images1.nx, images1.ny, images1.nz, images2.nx, images2.ny,
                images2.nz);
    } else if (rc) {
        fprintf(stderr, "%s\n", arec_last_error());
    }
    return rc;
}

// arecImage.cpp:673-691: rotate2d for every view.
void arecRotateImages(arecImage *images, float *angles) {
    if (arec_h_rotate_images(images->nx, images->ny, images->nz, images->data, angles, AREC_ROT_BILINEAR))
        fprintf(stderr, "%s\n", arec_last_error());
}

// arecImage.cpp:693-741, with the reference's messages for the angles it
// wraps, drops or hands to the bilinear rotation.
void arecRotateImages_skew(arecImage *images, float *angles) {
    for (int j = 0; j < images->nz; j++) {
        float a = angles[j];
        if (a < -180 || a > 180) {
            const float b = a < -180 ? a + 360 : a - 360;
            printf("Angle changed from %4.2f to %4.2f \n", a, b);
            a = b;
        }
        if (std::fabs(a) > 45)
            printf("Angle %4.2f too large, probably bad CC, do nothing \n", a);
        else if (std::fabs(a) > 10)
            printf("Angle %4.2f too large, using bilinear instead\n", a);
    }
    if (arec_h_rotate_images(images->nx, images->ny, images->nz, images->data, angles, AREC_ROT_SKEW))
        fprintf(stderr, "%s\n", arec_last_error());
}

// arecImage.cpp:743-794
void arecRotateImages_skew_safe(arecImage *images, arecImage *images_ref, float *angles) {
    int32_t kept = 0;
    double mean_cc = 0.0;
    if (arec_h_rotate_images_safe(images->nx, images->ny, images->nz, images->data, images_ref->data, angles, &kept,
                                  &mean_cc)) {
        fprintf(stderr, "%s\n", arec_last_error());
        return;
    }
    printf("arecRotateImages_skew_safe: Mean CC=%4.4f n_rot=%d/%d\n", mean_cc, kept, images->nz);
}

// ---- preprocessing -------------------------------------------------------------

// arecImage.cpp:358-381: a cylinder image holds nnz values, a stack nx*ny*nz.
void arecImageTakeLog(arecImage image) {
    const int64_t n = image.is_cyl == 1 ? (int64_t)image.nnz : (int64_t)image.nx * image.ny * image.nz;
    arec_h_take_log(image.data, n);
}

// imagetools.cpp:86-114: every view of the stack, in place.
void periodicDecomposition(arecImage img) {
    arec_h_periodic_decomposition(img.nx, img.ny, img.nz, img.data);
}

// arecImageIO.cpp:10-31
int arecReadImageHeader(char *fname, int *nx, int *ny, int *nz) {
    const int rc = arec_mrc_read_header(fname, nx, ny, nz, nullptr);
    if (rc) fprintf(stderr, "%s\n", arec_last_error());
    return rc;
}

// arecImageIO.cpp:33-82: the buffer is malloc'ed here, as the reference does.
int arecReadImage(char *fname, arecImage *image) {
    int nx = 0, ny = 0, nz = 0;
    int rc = arec_mrc_read_header(fname, &nx, &ny, &nz, nullptr);
    if (rc == 0) {
        image->nx = nx;
        image->ny = ny;
        image->nz = nz;
        const int64_t n = (int64_t)nx * ny * nz;
        image->data = (float *)malloc((size_t)(n > 0 ? n : 1) * sizeof(float));
        if (!image->data) {
            fprintf(stderr, "arecReadImage: out of memory for %s\n", fname);
            return -1;
        }
        rc = arec_mrc_read(fname, image->data, n);
    }
    if (rc) fprintf(stderr, "%s\n", arec_last_error());
    return rc;
}

// arecImageIO.cpp:84-173
int arecWriteImage(char *fname, arecImage image) {
    const int rc = arec_mrc_write(fname, image.nx, image.ny, image.nz, image.data);
    if (rc) fprintf(stderr, "%s\n", arec_last_error());
    return rc;
}
