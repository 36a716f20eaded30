// Link-time companion of the reference build in oracle/_ref (test
// infrastructure only; never part of the product).
//
// 1. The reference's projector and solver translation units need four
//    helpers from arecImage.cpp, which cannot compile here (it includes
//    fftw3.h, absent from this image).  They are restated below from
//    AREC3D/AREC3D/arecImage.cpp:15-27, 29-86, 234-241 and 796-810; the
//    cylinder mask is the same as orc_cyl_geometry in arec_oracle.c.
// 2. extern "C" wrappers give ctypes a flat interface to the reference's
//    arecProject2D*/arecBackProject2D*, cyl_sirt*, cyl_cgls* and the MRC
//    reader and writer (arecReadImage / arecWriteImage).
#include "arecImage.h"
#include "arecImageIO.h"
#include "arecproject.h"
#include "areccgls.h"
#include "areccgls_KB.h"
#include "arecsirt.h"
#include "arecsirt_KB.h"

#include <math.h>
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

int arecAllocateCBImage(arecImage *cbimage, int nx, int ny, int nz) {
    cbimage->nx = nx;
    cbimage->ny = ny;
    cbimage->nz = nz;
    cbimage->data = (float *)calloc((size_t)nx * ny * nz, sizeof(float));
    cbimage->is_cyl = 0;
    cbimage->cord = NULL;
    return cbimage->data ? 0 : -1;
}

int arecAllocateCylImage(arecImage *cyl, double radius, int height) {
    int n = static_cast<int>(round(2 * radius + 1));
    cyl->nx = n;
    cyl->ny = height;
    cyl->nz = n;
    cyl->radius = radius;
    cyl->height = height;
    int r2 = radius * radius;
    int nrays = 0;
    for (int pass = 0; pass < 2; pass++) {
        int k = 0;
        for (int iz = 0; iz < n; iz++) {
            int zs = iz - radius;
            for (int ix = 0; ix < n; ix++) {
                int xs = ix - radius;
                if (zs * zs + xs * xs <= r2) {
                    if (pass) {
                        cyl->cord[2 * k] = ix;
                        cyl->cord[2 * k + 1] = iz;
                    }
                    k++;
                }
            }
        }
        if (!pass) {
            nrays = k;
            cyl->nrays = nrays;
            cyl->nnz = nrays * height;
            cyl->data = (float *)calloc((size_t)nrays * height, sizeof(float));
            // two spare ints: the linear projector reads up to 3 past the end
            // (its 1-based loop over the 0-based cord macro); they hold -1e6
            // so that those two columns fall off the detector.
            cyl->cord = (int *)calloc((size_t)nrays * 2 + 4, sizeof(int));
            for (int e = 0; e < 4; e++) cyl->cord[2 * nrays + e] = -1000000;
        }
    }
    cyl->is_cyl = 1;
    return cyl->data ? 0 : -1;
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

static void ensure_mpi() {
    int ok = 0;
    MPI_Initialized(&ok);
    if (!ok) MPI_Init(NULL, NULL);
}

// The reference prints progress to stdout; keep the test logs quiet.
struct Quiet {
    int saved;
    Quiet() {
        fflush(stdout);
        saved = dup(1);
        FILE *f = fopen("/dev/null", "w");
        dup2(fileno(f), 1);
        fclose(f);
    }
    ~Quiet() {
        fflush(stdout);
        dup2(saved, 1);
        close(saved);
    }
};

extern "C" {

int ref_cyl_geometry(double radius, int height, int *nrays, int *cord) {
    arecImage c;
    int st = arecAllocateCylImage(&c, radius, height);
    *nrays = c.nrays;
    if (cord) memcpy(cord, c.cord, sizeof(int) * 2 * c.nrays);
    arecImageFree(&c);
    return st;
}

// x: [nrays][ny] -> p: [na][ny][nx]
int ref_project(int method, double radius, int nx, int ny, const float *x, const float *angles,
                int na, float *p) {
    arecImage cyl, prj;
    arecAllocateCylImage(&cyl, radius, ny);
    arecAllocateCBImage(&prj, nx, ny, na);
    memcpy(cyl.data, x, sizeof(float) * cyl.nnz);
    float *ang = (float *)malloc(sizeof(float) * na);
    memcpy(ang, angles, sizeof(float) * na);
    int st;
    {
        Quiet q;
        st = method == 1 ? arecProject2D(cyl, ang, na, &prj)
             : method == 2 ? arecProject2D_KB(cyl, ang, na, &prj)
                           : arecProject2D_SQ(cyl, ang, na, &prj);
    }
    memcpy(p, prj.data, sizeof(float) * (size_t)nx * ny * na);
    arecImageFree(&cyl);
    arecImageFree(&prj);
    free(ang);
    return st;
}

int ref_backproject(int method, double radius, int nx, int ny, const float *p,
                    const float *angles, int na, float *x) {
    arecImage cyl, prj;
    arecAllocateCylImage(&cyl, radius, ny);
    arecAllocateCBImage(&prj, nx, ny, na);
    memcpy(prj.data, p, sizeof(float) * (size_t)nx * ny * na);
    float *ang = (float *)malloc(sizeof(float) * na);
    memcpy(ang, angles, sizeof(float) * na);
    int st;
    {
        Quiet q;
        st = method == 1 ? arecBackProject2D(prj, ang, na, &cyl)
             : method == 2 ? arecBackProject2D_KB(prj, ang, na, &cyl)
                           : arecBackProject2D_SQ(prj, ang, na, &cyl);
    }
    memcpy(x, cyl.data, sizeof(float) * cyl.nnz);
    arecImageFree(&cyl);
    arecImageFree(&prj);
    free(ang);
    return st;
}

// solver: 1 = SIRT, 2 = CGLS.  x must hold nnz floats; warm != 0 passes x
// as the starting volume (CGLS only, as in the reference).
int ref_solve(int solver, int method, const float *b, int nx, int ny, int na,
              const float *angles, float lam, int maxit, float tol, int warm, float *x,
              int *nnz_out) {
    ensure_mpi();
    arecImage img, xv;
    arecAllocateCBImage(&img, nx, ny, na);
    memcpy(img.data, b, sizeof(float) * (size_t)nx * ny * na);
    float *ang = (float *)malloc(sizeof(float) * na);
    memcpy(ang, angles, sizeof(float) * na);
    memset(&xv, 0, sizeof xv);
    if (warm) {
        arecAllocateCylImage(&xv, 0.5 * (nx - 1), ny);
        memcpy(xv.data, x, sizeof(float) * xv.nnz);
    }
    int st;
    {
        Quiet q;
        if (solver == 1)
            st = method == 1 ? cyl_sirt(MPI_COMM_WORLD, img, ang, &xv, lam, maxit, tol)
                 : method == 2 ? cyl_sirt_KB(MPI_COMM_WORLD, img, ang, &xv, lam, maxit, tol)
                               : cyl_sirt_SQ(MPI_COMM_WORLD, img, ang, &xv, lam, maxit, tol);
        else
            st = method == 1 ? cyl_cgls(MPI_COMM_WORLD, img, ang, &xv, lam, maxit, tol)
                 : method == 2 ? cyl_cgls_KB(MPI_COMM_WORLD, img, ang, &xv, lam, maxit, tol)
                               : cyl_cgls_SQ(MPI_COMM_WORLD, img, ang, &xv, lam, maxit, tol);
    }
    if (nnz_out) *nnz_out = xv.nnz;
    if (xv.data) memcpy(x, xv.data, sizeof(float) * xv.nnz);
    if (xv.data) arecImageFree(&xv);
    arecImageFree(&img);
    free(ang);
    return st;
}

// arecWriteImage of a [nz][ny][nx] float array.
int ref_mrc_write(const char *path, int nx, int ny, int nz, const float *data) {
    arecImage img;
    memset(&img, 0, sizeof(img));
    img.nx = nx;
    img.ny = ny;
    img.nz = nz;
    img.data = (float *)data;
    return arecWriteImage((char *)path, img);
}

// arecReadImage into data (capacity floats); sizes into nx, ny, nz.
int ref_mrc_read(const char *path, int *nx, int *ny, int *nz, float *data, long capacity) {
    arecImage img;
    memset(&img, 0, sizeof(img));
    int st = arecReadImage((char *)path, &img);
    *nx = img.nx;
    *ny = img.ny;
    *nz = img.nz;
    const long n = (long)img.nx * img.ny * img.nz;
    if (st == 0 && n <= capacity && img.data) memcpy(data, img.data, sizeof(float) * n);
    free(img.data);
    return st;
}

}  // extern "C"
