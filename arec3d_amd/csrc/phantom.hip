// Synthetic inputs: ellipsoid phantoms on the compact cylinder and seeded
// noise, evaluated on the device.  Parameters are drawn on the host
// (arec3d_amd/phantom.py, SplitMix64) so that CPU tests and GPU runs see the
// same phantom.
#include <hip/hip_runtime.h>

#include <cmath>

#include "arec_internal.hpp"

namespace arec {

// One ellipsoid: centre (cx, cy, cz) in voxel units about the cylinder axis
// (cx along ix, cz along iz, cy along the rows), rotation matrix rows
// r0..r8 (world -> body), semi-axes (ax, ay, az), density.
struct Ellipsoid {
    double cx, cy, cz;
    double r[9];
    double ax, ay, az;
    double dens;
};

__global__ void phantom_kernel(const int *__restrict__ cord, int nrays, int ny, double radius,
                               double support, const Ellipsoid *__restrict__ E, int ne,
                               float *__restrict__ x) {
    int64_t n = (int64_t)nrays * ny;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int j = (int)(i / ny), y = (int)(i % ny);
        double px = cord[2 * j] - radius, pz = cord[2 * j + 1] - radius;
        double py = y - 0.5 * (ny - 1);
        double v = 0.0;
        if (px * px + pz * pz <= support * support) {
            for (int e = 0; e < ne; e++) {
                double dx = px - E[e].cx, dy = py - E[e].cy, dz = pz - E[e].cz;
                double u = E[e].r[0] * dx + E[e].r[1] * dy + E[e].r[2] * dz;
                double w = E[e].r[3] * dx + E[e].r[4] * dy + E[e].r[5] * dz;
                double t = E[e].r[6] * dx + E[e].r[7] * dy + E[e].r[8] * dz;
                double q = (u / E[e].ax) * (u / E[e].ax) + (w / E[e].ay) * (w / E[e].ay) +
                           (t / E[e].az) * (t / E[e].az);
                if (q <= 1.0) v += E[e].dens;
            }
        }
        x[i] = (float)v;
    }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t z) { return (double)(z >> 11) * (1.0 / 9007199254740992.0); }

// mode 0: b += sigma * N(0,1).  mode 1: Beer-Lambert counts with I0 photons,
// attenuation scale s: lam = I0 exp(-s b), counts ~ N(lam, lam) rounded and
// clipped at 0 (the normal approximation of Poisson), b = -ln(counts/I0),
// zero counts giving 0 like arecImageTakeLog (arecImage.cpp:358-381).
__global__ void noise_kernel(float *__restrict__ b, int64_t n, uint64_t seed, int mode, double sigma,
                             double i0, double scale) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t h = splitmix64(seed ^ splitmix64((uint64_t)i));
        double u1 = u01(splitmix64(h)), u2 = u01(splitmix64(h + 1));
        if (u1 < 1e-300) u1 = 1e-300;
        double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        if (mode == 0) {
            b[i] = (float)((double)b[i] + sigma * g);
        } else {
            double lam = i0 * exp(-scale * (double)b[i]);
            double c = rint(lam + sqrt(lam) * g);
            b[i] = c > 0 ? (float)(-log(c / i0)) : 0.0f;
        }
    }
}

}  // namespace arec

using namespace arec;

extern "C" {

// Evaluate a phantom of ne ellipsoids (packed as 16 doubles each: cx, cy, cz,
// r[9], ax, ay, az, density) on the compact cylinder of the given radius.
int arec_phantom(arec_ctx *ctx, double radius, int64_t ny, double support, const double *params,
                 int32_t ne, float *x_dev) {
    if (!ctx || !x_dev || ny <= 0 || ne < 0) {
        set_error("arec_phantom: bad argument");
        return AREC_EINVAL;
    }
    std::vector<int> cord;
    int nrays = cyl_mask(radius, &cord);
    int *d_cord = nullptr;
    Ellipsoid *d_e = nullptr;
    AREC_HIP(hipMalloc((void **)&d_cord, sizeof(int) * cord.size()));
    AREC_HIP(hipMalloc((void **)&d_e, sizeof(Ellipsoid) * (ne > 0 ? ne : 1)));
    AREC_HIP(hipMemcpyAsync(d_cord, cord.data(), sizeof(int) * cord.size(), hipMemcpyHostToDevice, ctx->stream));
    if (ne > 0)
        AREC_HIP(hipMemcpyAsync(d_e, params, sizeof(Ellipsoid) * ne, hipMemcpyHostToDevice, ctx->stream));
    phantom_kernel<<<2048, 256, 0, ctx->stream>>>(d_cord, nrays, (int)ny, radius, support, d_e, ne, x_dev);
    AREC_HIP(hipGetLastError());
    AREC_HIP(hipStreamSynchronize(ctx->stream));
    hipFree(d_cord);
    hipFree(d_e);
    return 0;
}

int arec_add_noise(arec_ctx *ctx, float *b_dev, int64_t n, uint64_t seed, int32_t mode, double sigma,
                   double i0, double scale) {
    if (!ctx || !b_dev || n < 0 || mode < 0 || mode > 1) {
        set_error("arec_add_noise: bad argument");
        return AREC_EINVAL;
    }
    noise_kernel<<<2048, 256, 0, ctx->stream>>>(b_dev, n, seed, mode, sigma, i0, scale);
    AREC_HIP(hipGetLastError());
    return 0;
}

}  // extern "C"
