// Caller written against the reference's interface (arec3d_dropin.h mirrors
// AREC3D/AREC3D/arecImage.h, arecproject.h, arecsirt.h, areccgls.h), linked
// with libarec3d_dropin.so.  Re-checks the reference's unit tests
// (AREC3D/AREC3D_test/AREC3D_test.cpp:8-189, 216-243) and runs a solver.
// Exit status 0 = all checks passed.
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "arec3d_dropin.h"

static int fails = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            fprintf(stderr, "FAILED %s:%d %s\n", __FILE__, __LINE__, #c); \
            fails++;                                               \
        }                                                          \
    } while (0)

static bool approx(double a, double b, double eps = 1e-5) { return std::fabs(a - b) <= eps * (1 + std::fabs(b)); }

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    {  // Cylvol, odd
        arecImage c;
        arecAllocateCylImage(&c, 1, 2);
        CHECK(c.nrays == 5 && c.nnz == 10);
        int want[10] = {1, 0, 0, 1, 1, 1, 2, 1, 1, 2};
        for (int i = 0; i < 10; i++) CHECK(c.cord[i] == want[i]);
        arecImageFree(&c);
    }
    {  // projection / backprojection, odd (SQ, theta = 0)
        arecImage cyl, prj;
        arecAllocateCylImage(&cyl, 1, 2);
        arecAllocateCBImage(&prj, 3, 2, 1);
        cyl.data[4] = 1.0f;
        cyl.data[5] = 2.0f;
        float angle = 0.0f;
        CHECK(arecProject2D_SQ(cyl, &angle, 1, &prj) == 0);
        double s = 0;
        for (int i = 0; i < 6; i++) s += prj.data[i];
        CHECK(approx(s, 3.0) && approx(prj.data[1], 1.0) && approx(prj.data[4], 2.0));
        reset(&prj);
        prj.data[1] = 1.0f;
        prj.data[4] = 2.0f;
        CHECK(arecBackProject2D_SQ(prj, &angle, 1, &cyl) == 0);
        s = 0;
        for (int i = 0; i < cyl.nnz; i++) s += cyl.data[i];
        CHECK(approx(s, 9.0));
        arecImageFree(&cyl);
        arecImageFree(&prj);
    }
    {  // A vs AT, odd: impulse responses equal exactly
        float angles[3] = {1.475f, 21.123f, 2.123f};
        arecImage cyl, prj;
        arecAllocateCylImage(&cyl, 1.0, 2);
        arecAllocateCBImage(&prj, cyl.nx, 2, 3);
        int np = prj.nx * prj.ny * prj.nz;
        for (int a = 0; a < cyl.nnz; a++) {
            reset(&cyl);
            cyl.data[a] = 1.0f;
            arecProject2D_SQ(cyl, angles, 3, &prj);
            float *col = (float *)malloc(sizeof(float) * np);
            for (int b = 0; b < np; b++) col[b] = prj.data[b];
            for (int b = 0; b < np; b++) {
                reset(&prj);
                prj.data[b] = 1.0f;
                arecBackProject2D_SQ(prj, angles, 3, &cyl);
                CHECK(col[b] == cyl.data[a]);
            }
            free(col);
        }
        arecImageFree(&cyl);
        arecImageFree(&prj);
    }
    {  // solver: SIRT from a NULL volume, then CGLS warm-started from it
        int nx = 33, ny = 4, na = 24;
        float *angles = (float *)malloc(sizeof(float) * na);
        for (int i = 0; i < na; i++) angles[i] = (float)(i * M_PI / na);
        arecImage truth, img, x;
        arecAllocateCylImage(&truth, 0.5 * (nx - 1), ny);
        for (int i = 0; i < truth.nnz; i++) truth.data[i] = (float)((i * 37) % 11) / 11.0f;
        arecAllocateCBImage(&img, nx, ny, na);
        CHECK(arecProject2D_SQ(truth, angles, na, &img) == 0);
        x.data = NULL;
        CHECK(cyl_sirt_SQ(MPI_COMM_WORLD, img, angles, &x, 1e-3f, 10, 1e-6f) == 0);
        CHECK(x.nrays == truth.nrays && x.nnz == truth.nnz);
        CHECK(cyl_cgls_SQ(MPI_COMM_WORLD, img, angles, &x, 0.0f, 10, 1e-6f) == 0);
        double err = 0, nrm = 0;
        for (int i = 0; i < x.nnz; i++) {
            err += (x.data[i] - truth.data[i]) * (x.data[i] - truth.data[i]);
            nrm += truth.data[i] * truth.data[i];
        }
        printf("relative error after SIRT+CGLS: %.3e\n", std::sqrt(err / nrm));
        CHECK(std::sqrt(err / nrm) < 0.5);
        arecImage even;
        arecAllocateCBImage(&even, 32, ny, na);
        arecImage xe;
        xe.data = NULL;
        CHECK(cyl_cgls_KB(MPI_COMM_WORLD, even, angles, &xe, 0.0f, 3, 1e-6f) == -1);  // F8
        arecImageFree(&truth);
        arecImageFree(&img);
        arecImageFree(&x);
        arecImageFree(&even);
        arecImageFree(&xe);
        free(angles);
    }
    {  // rotational alignment: a stack turned by +5 degrees (rotate2d) measures about -5
        const int nx = 64, ny = 64, nz = 2;
        arecImage a, b;
        arecAllocateCBImage(&a, nx, ny, nz);
        arecAllocateCBImage(&b, nx, ny, nz);
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    const double dx = x - 44.0 - 3 * z, dy = y - 30.0, ex = x - 20.0, ey = y - 40.0 + 2 * z;
                    a.data[(z * ny + y) * nx + x] = (float)(std::exp(-(dx * dx + dy * dy) / 18.0) +
                                                            0.6 * std::exp(-(ex * ex + ey * ey) / 24.0));
                }
        for (int i = 0; i < nx * ny * nz; i++) b.data[i] = a.data[i];
        float turn[2] = {5.0f, 5.0f}, ang[2] = {0.0f, 0.0f};
        arecRotateImages(&b, turn);
        CHECK(arecRotCCImages(MPI_COMM_WORLD, a, b, 27, ang) == 0);
        CHECK(std::fabs(ang[0] + 5.0f) < 0.3f && std::fabs(ang[1] + 5.0f) < 0.3f);
        arecRotateImages_skew_safe(&b, &a, ang);
        CHECK(std::fabs(ang[0] + 2.5f) < 0.2f);  // kept and halved
        arecImage c;
        arecAllocateCBImage(&c, nx, ny, 1);
        CHECK(arecRotCCImages(MPI_COMM_WORLD, a, c, 27, ang) == -1);  // size mismatch
        arecImageFree(&a);
        arecImageFree(&b);
        arecImageFree(&c);
    }
    MPI_Finalize();
    printf(fails ? "dropin driver: %d failures\n" : "dropin driver: all checks passed%.0d\n", fails);
    return fails ? 1 : 0;
}
