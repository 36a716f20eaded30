/*
 * Drop-in C++ interface of libarec3d_dropin.so: the reference's projector,
 * solver, alignment and image-helper symbols with their exact signatures,
 * backed by the HIP implementation in libarec3d_amd.so.  A program written
 * against the reference's headers links this library instead of the
 * reference's arecproject.o / arecsirt*.o / areccgls*.o / arecImage.o.
 *
 * The struct layout is the reference's (AREC3D/AREC3D/arecImage.h:5-18,
 * 64 bytes on LP64) because callers pass it by value.  MPI_Comm comes from
 * the caller's MPI (MPICH: an int), as in the reference's headers.  One GPU
 * per process: a communicator with more than one rank is rejected (status
 * -2); several ranks would each hold one Y-slab on their own GPU, which
 * this library does not do.
 */
#ifndef AREC3D_DROPIN_H
#define AREC3D_DROPIN_H

#include <mpi.h>

typedef struct Image {
    int nx;
    int ny;
    int nz;
    double radius;
    int height;
    int nnz;
    int nrays;
    int *cord;
    int is_cyl;
    float *data;
} arecImage;

/* arecImage.h:27-32,56-57 (arecImage.cpp:15-241,796-826) */
int arecAllocateCBImage(arecImage *cbimage, int nx, int ny, int nz);
int arecAllocateCylImage(arecImage *cylimage, double radius, int height);
int ImageCB2Cyl(arecImage cbimage, arecImage cylimage, int xcent, int ycent, double radius,
                int height);
int ImageCyl2CB(arecImage cylimage, arecImage cbimage, int nx, int ny, int nz);
void arecImageFree(arecImage *image);
void soft_noneg(arecImage *image);
void reset(arecImage *image);

/* arecproject.h:55-62 */
int arecBackProject2D(arecImage images, float *angles, int nangles, arecImage *bvol);
int arecProject2D(arecImage cylvol, float *angles, int nangles, arecImage *projstack);
int arecBackProject2D_KB(arecImage images, float *angles, int nangles, arecImage *bvol);
int arecProject2D_KB(arecImage cylvol, float *angles, int nangles, arecImage *projstack);
int arecBackProject2D_SQ(arecImage images, float *angles, int nangles, arecImage *bvol);
int arecProject2D_SQ(arecImage cylvol, float *angles, int nangles, arecImage *projstack);

/* arecsirt.h:6-10, arecsirt_KB.h:6-7, areccgls.h:6-10, areccgls_KB.h:6-7,
 * areccgls_stat.h:6-13 */
int cyl_sirt(MPI_Comm comm, arecImage images, float *angles, arecImage *xvol, float lam,
             int maxit, float tol);
int cyl_sirt_SQ(MPI_Comm comm, arecImage images, float *angles, arecImage *xvol, float lam,
                int maxit, float tol);
int cyl_sirt_KB(MPI_Comm comm, arecImage images, float *angles, arecImage *xvol, float lam,
                int maxit, float tol);
int cyl_cgls(MPI_Comm comm, arecImage images, float *angles, arecImage *xvol, float lam,
             int maxit, float tol);
int cyl_cgls_SQ(MPI_Comm comm, arecImage images, float *angles, arecImage *xvol, float lam,
                int maxit, float tol);
int cyl_cgls_KB(MPI_Comm comm, arecImage images, float *angles, arecImage *xvol, float lam,
                int maxit, float tol);
int init_weights(MPI_Comm comm, arecImage projections, arecImage *weights);
int cyl_cgls_stat(MPI_Comm comm, arecImage images, float *angles, arecImage *xvol, float lam,
                  int maxit, float tol);
int cyl_cgls_KB_stat(MPI_Comm comm, arecImage images, float *angles, arecImage *xvol, float lam,
                     int maxit, float tol);
int cyl_cgls_SQ_stat(MPI_Comm comm, arecImage images, float *angles, arecImage *xvol, float lam,
                     int maxit, float tol);

/* arecImage.h:36,45 */
int arecCCImages(MPI_Comm comm, arecImage images1, arecImage images2, float *sx, float *sy);
void arecShiftImages(arecImage images, float *sx, float *sy);

/* align2dstack.h:9,13,24-25 (align2dstack.cpp:480-521,566-630,1059-1166):
 * one view's correlation map, Fourier shift, sub-pixel peak and 3x3 fit */
void ccorr2d(int nx, int ny, float *x, float *y, float *c);
void fshift2d(int nx, int ny, float *imagein, float *imageout, float sx, float sy);
void quad2dfit(float *Z, float *XSH, float *YSH, float *PEAKV);
void speak(int nx, int ny, float *data, float *xt, float *yt);

/* arecImage.h:37,46-48 (rotational alignment) */
int arecRotCCImages(MPI_Comm comm, arecImage images1, arecImage images2, int r2, float *angle);
void arecRotateImages(arecImage *images, float *angles);
void arecRotateImages_skew(arecImage *images, float *angles);
void arecRotateImages_skew_safe(arecImage *images, arecImage *images_ref, float *angles);

/* arecImage.h:34, imagetools.h:5 (preprocessing of the measured views) */
void arecImageTakeLog(arecImage image);
void periodicDecomposition(arecImage img);

/* arecImageIO.h:64-66 (arecImageIO.cpp:10-173): MRC files.  arecReadImage
 * mallocs image->data (free with arecImageFree) and sets nx, ny, nz. */
int arecReadImageHeader(char *fname, int *nx, int *ny, int *nz);
int arecReadImage(char *fname, arecImage *image);
int arecWriteImage(char *fname, arecImage image);

#endif
