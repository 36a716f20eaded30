// Caller written against the reference's interface (arec3d_dropin.h mirrors
// AREC3D/AREC3D/arecImage.h, arecproject.h, arecsirt.h, areccgls.h,
// align2dstack.h), linked with libarec3d_dropin.so.  Re-checks the
// reference's unit tests (AREC3D/AREC3D_test/AREC3D_test.cpp:8-189, 216-243,
// 326-363), runs a solver and the per-view alignment helpers.  With --host
// only the checks that need no GPU run (cylinder, cube <-> cylinder,
// quad2dfit).  Exit status 0 = all checks passed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

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

// The cube <-> cylinder checks and the 3x3 fit: host code only.
static void host_checks() {
    {  // ImageCyl2CB then ImageCB2Cyl (AREC3D_test.cpp:326-363)
        arecImage cylvol;
        arecAllocateCylImage(&cylvol, 1, 2);
        for (int i = 0; i < cylvol.nnz; i++) cylvol.data[i] = 1.0f * i;
        double sum_cyl = 0.0;
        for (int i = 0; i < cylvol.nnz; i++) sum_cyl += cylvol.data[i];
        arecImage xcb;
        arecAllocateCBImage(&xcb, cylvol.nx, cylvol.ny, cylvol.nz);
        CHECK(ImageCyl2CB(cylvol, xcb, cylvol.nx, cylvol.ny, cylvol.nz) == 0);
        double sum_cub = 0.0;
        for (int i = 0; i < xcb.nx * xcb.ny * xcb.nz; i++) sum_cub += xcb.data[i];
        CHECK(sum_cub == sum_cyl);
        arecImage cylvol2;
        arecAllocateCylImage(&cylvol2, 1, 2);
        CHECK(ImageCB2Cyl(xcb, cylvol2, 2, 2, 2, 2) == 0);
        for (int i = 0; i < cylvol.nnz; i++) CHECK(cylvol.data[i] == cylvol2.data[i]);
        for (int i = 0; i < 2 * cylvol.nrays; i++) CHECK(cylvol.cord[i] == cylvol2.cord[i]);
        // a cube of another height: the count check reports it (-2)
        arecImage tall;
        arecAllocateCBImage(&tall, 3, 3, 3);
        CHECK(ImageCB2Cyl(tall, cylvol2, 0, 0, 0, 0) == -2);
        // sizes that disagree with the cylinder: status 1
        CHECK(ImageCyl2CB(cylvol, tall, 3, 3, 3) == 1);
        arecImageFree(&cylvol);
        arecImageFree(&cylvol2);
        arecImageFree(&xcb);
        arecImageFree(&tall);
    }
    {  // half-integer radius: the allocator's truncated mask keeps all 16
       // columns of a 4 x 4 cross-section, ImageCB2Cyl's (int)r-centred
       // mask 9 of them, so the reference reports -2
        arecImage cyl;
        arecAllocateCylImage(&cyl, 1.5, 2);
        CHECK(cyl.nrays == 16 && cyl.nx == 4);
        for (int i = 0; i < cyl.nnz; i++) cyl.data[i] = 1.0f + i;
        arecImage cb;
        arecAllocateCBImage(&cb, 4, 2, 4);
        CHECK(ImageCyl2CB(cyl, cb, 4, 2, 4) == 0);
        arecImage back;
        arecAllocateCylImage(&back, 1.5, 2);
        CHECK(ImageCB2Cyl(cb, back, 0, 0, 0, 0) == -2);
        // its 9 columns, iz outer: (0..2, 0..2) with (ix-1)^2 + (iz-1)^2 <= 2
        for (int j = 0; j < 9; j++) {
            const int ix = j % 3, iz = j / 3;
            CHECK(back.cord[2 * j] == ix && back.cord[2 * j + 1] == iz);
            CHECK(back.data[2 * j] == cb.data[(iz * 2 + 0) * 4 + ix]);
            CHECK(back.data[2 * j + 1] == cb.data[(iz * 2 + 1) * 4 + ix]);
        }
        arecImageFree(&cyl);
        arecImageFree(&cb);
        arecImageFree(&back);
    }
    {  // quad2dfit on a sampled paraboloid: centre (0.3, -0.4), peak 2
        float z[9], xs = 0, ys = 0, pv = 0;
        for (int k = 0; k < 9; k++) {
            const double u = k % 3 - 1, v = k / 3 - 1;
            z[k] = (float)(2.0 - 0.5 * (u - 0.3) * (u - 0.3) - 0.8 * (v + 0.4) * (v + 0.4) + 0.1 * (u - 0.3) * (v + 0.4));
        }
        quad2dfit(z, &xs, &ys, &pv);
        CHECK(approx(xs, 0.3, 1e-5) && approx(ys, -0.4, 1e-5) && approx(pv, 2.0, 1e-5));
        float flat[9] = {1, 1, 1, 1, 1, 1, 1, 1, 1};
        pv = 7.0f;
        quad2dfit(flat, &xs, &ys, &pv);
        CHECK(xs == 0.0f && ys == 0.0f && pv == 7.0f);  // singular: no shift, peak value untouched
    }
}

// ccorr2d, speak and fshift2d on one 16 x 8 view (GPU).
static void view_checks() {
    const int nx = 16, ny = 8, n = nx * ny;
    float x[n], y[n], c[n], out[n];
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const double a = i - 6.3, b = j - 3.1;
            x[j * nx + i] = (float)(std::exp(-(a * a) / 6.0 - (b * b) / 3.0) + 0.01 * ((i * 7 + j * 3) % 5));
        }
    for (int j = 0; j < ny; j++)  // y = x shifted by (+3, +2), circularly
        for (int i = 0; i < nx; i++) y[j * nx + i] = x[((j - 2 + ny) % ny) * nx + (i - 3 + nx) % nx];
    ccorr2d(nx, ny, x, y, c);
    double worst = 0, scale = 0;  // c[n] = N sum_m x[m] y[m + n]
    for (int dj = 0; dj < ny; dj++)
        for (int di = 0; di < nx; di++) {
            double s = 0;
            for (int j = 0; j < ny; j++)
                for (int i = 0; i < nx; i++) s += (double)x[j * nx + i] * y[((j + dj) % ny) * nx + (i + di) % nx];
            s *= n;
            worst = std::fmax(worst, std::fabs(s - c[dj * nx + di]));
            scale = std::fmax(scale, std::fabs(s));
        }
    CHECK(worst <= 1e-5 * scale);
    float sx = 0, sy = 0;
    speak(nx, ny, c, &sx, &sy);
    CHECK(std::fabs(sx - 3.0f) < 0.5f && std::fabs(sy - 2.0f) < 0.5f);
    fshift2d(nx, ny, x, out, 3.0f, 2.0f);  // an integer shift is the circular shift
    double e = 0;
    for (int i = 0; i < n; i++) e = std::fmax(e, std::fabs(out[i] - y[i]));
    CHECK(e < 1e-5);
    fshift2d(nx, ny, y, y, -3.0f, -2.0f);  // in place, back
    e = 0;
    for (int i = 0; i < n; i++) e = std::fmax(e, std::fabs(y[i] - x[i]));
    CHECK(e < 1e-5);
}

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    host_checks();
    if (argc > 1 && !std::strcmp(argv[1], "--host")) {
        MPI_Finalize();
        printf(fails ? "dropin driver (host): %d failures\n" : "dropin driver (host): all checks passed%.0d\n", fails);
        return fails ? 1 : 0;
    }
    view_checks();
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
