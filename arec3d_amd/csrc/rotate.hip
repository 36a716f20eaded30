// Rotational alignment of measured views against reprojections, gfx950
// (the driver's align = 2 step, AREC3D.cpp:232-243).
//
//  arecRotCCImages (arecImage.cpp:555-599): per view, the radial profiles of
//      both stacks (getRadialProfile, :516-553: for each of L = (int)(2 pi r2)
//      directions, sum over r < r2 of r times the bilinear sample at radius
//      r), their circular 1-D correlation (ccorr1d, align2dstack.cpp:437-475)
//      and its peak (findPeak1d + PRB1D, arecImage.cpp:489-514,
//      align2dstack.cpp:1028-1054) as an angle in degrees.
//  arecRotateImages_skew (arecImage.cpp:693-741): angles wrapped by 360,
//      beyond 45 degrees dropped, beyond 10 the bilinear rotate2d
//      (align2dstack.cpp:774-822), else three circular shears
//      (shearX_circ / shearY_circ, imagetools.cpp:244-291).
//  arecRotateImages_skew_safe (arecImage.cpp:743-794): the three shears,
//      kept only where the Tukey-windowed correlation coefficient with the
//      reference view (imagetools.cpp:332-367, 415-447) rises by > 1e-5.
//
// Transcendental values (profile directions, shear factors, the bilinear
// rotation's cosine and sine, the Tukey windows) are evaluated on the host
// in the reference's precision and passed in, so the device arithmetic --
// float and double operations in the reference's order -- reproduces the
// reference's rounding.  Each kernel takes whole views: a profile or
// correlation lag per thread, a pixel per thread for the rotations, one
// workgroup per view for the correlation coefficients.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "arec_internal.hpp"

namespace arec {

namespace {

constexpr double kPi = 3.14159265358979323846;  // 4 atan(1) in double
constexpr int kRotThreads = 256;

// Per-view rotation: kind 0 copies, 1 runs the three shears (a, b, a),
// 2 the bilinear rotation with cod, sid.
struct RotView {
    int kind;
    float a, b, pad;
    double cod, sid;
};

__global__ void __launch_bounds__(kRotThreads) radial_profile_kernel(const float *__restrict__ img, int nx, int ny,
                                                                     int L, int R, const double2 *__restrict__ cs,
                                                                     float *__restrict__ out) {
    const int i = blockIdx.x * kRotThreads + threadIdx.x;
    if (i >= L) return;
    const int64_t v = blockIdx.y;
    const float *im = img + v * nx * (int64_t)ny;
    const double xc = 0.5 * (nx - 1), yc = 0.5 * (ny - 1), ca = cs[i].x, sa = cs[i].y;
    float acc = 0.0f;
    for (int r = 0; r < R; r++) {
        const float x = (float)(xc + ca * r), y = (float)(yc + sa * r);
        const int ix = (int)x, iy = (int)y;
        double fx = (double)(x - (float)ix), fy = (double)(y - (float)iy);
        fx = fx < 0.0 ? 0.0 : fx;
        fy = fy < 0.0 ? 0.0 : fy;
        const float *row = im + (int64_t)iy * nx + ix;
        const double v00 = row[0], v10 = row[1], v01 = row[nx], v11 = row[nx + 1];
        const double top = v01 + fx * (v11 - v01), bot = v00 + fx * (v10 - v00);
        const float s = (float)(bot + fy * (top - bot));
        acc += s * (float)r;
    }
    out[v * L + i] = acc;
}

// c[k] = L * sum_m p1[m] p2[(m + k) mod L], m ascending, in double.
__global__ void __launch_bounds__(kRotThreads) ccorr1d_kernel(const float *__restrict__ p1,
                                                              const float *__restrict__ p2, int L,
                                                              float *__restrict__ c) {
    const int k = blockIdx.x * kRotThreads + threadIdx.x;
    if (k >= L) return;
    const int64_t v = blockIdx.y;
    const float *x = p1 + v * L, *y = p2 + v * L;
    double acc = 0.0;
    for (int m = 0; m < L - k; m++) acc += (double)x[m] * (double)y[m + k];
    for (int m = L - k; m < L; m++) acc += (double)x[m] * (double)y[m + k - L];
    c[v * L + k] = (float)((double)L * acc);
}

// findPeak1d: the first strict maximum, then the 7-point parabola (PRB1D).
__global__ void __launch_bounds__(kRotThreads) peak1d_kernel(const float *__restrict__ c, int L,
                                                             float *__restrict__ pos) {
    __shared__ float sv[kRotThreads];
    __shared__ int si[kRotThreads];
    const int64_t v = blockIdx.x;
    const float *cv = c + v * L;
    const int t = threadIdx.x;
    float best = cv[0];
    int bi = 0;
    for (int i = t; i < L; i += kRotThreads)  // ascending: the first maximum of this thread's share
        if (cv[i] > best) {
            best = cv[i];
            bi = i;
        }
    sv[t] = best;
    si[t] = bi;
    __syncthreads();
    for (int s = kRotThreads / 2; s > 0; s >>= 1) {
        if (t < s) {
            const float o = sv[t + s];
            const int oi = si[t + s];
            if (o > sv[t] || (o == sv[t] && oi < si[t])) {
                sv[t] = o;
                si[t] = oi;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        const int m = si[0];
        double b[7];
        for (int k = -3; k <= 3; k++) b[k + 3] = (double)cv[(m + k + L) % L];
        const double c2 = 49. * b[0] + 6. * b[1] - 21. * b[2] - 32. * b[3] - 27. * b[4] - 6. * b[5] + 31. * b[6];
        const double c3 = 5. * b[0] - 3. * b[2] - 4. * b[3] - 3. * b[4] + 5. * b[6];
        const double p = c3 != 0.0 ? c2 / (2. * c3) - 4 : 0.0;
        pos[v] = (float)(p + m);
    }
}

__device__ __forceinline__ int wrap(int64_t i, int n) {
    int64_t j = i % n;
    return (int)(j < 0 ? j + n : j);
}

// One pass of the rotation (pass 0: x shear or bilinear; 1: y shear;
// 2: x shear); views whose kind has nothing to do in a pass are copied.
__global__ void __launch_bounds__(kRotThreads) rotate_pass_kernel(const float *__restrict__ src,
                                                                  float *__restrict__ dst, int nx, int ny,
                                                                  const RotView *__restrict__ prm, int pass) {
    const int n = nx * ny;  // < 2^31 per view (checked by the host)
    const int p = blockIdx.x * kRotThreads + threadIdx.x;
    if (p >= n) return;
    const int64_t v = blockIdx.y;
    const float *s = src + v * n;
    float *d = dst + v * n;
    const int x = p % nx, y = p / nx;
    const RotView q = prm[v];
    float out;
    if (q.kind == 1 && pass != 1) {  // shearX_circ: row y moves by a (y - (ny-1)/2)
        const float skew = (float)((double)q.a * ((double)y - 0.5 * (ny - 1)));
        const int k = (int)floorf(skew);
        const float f = (float)((double)skew - 1.0 * k);
        const float *row = s + (int64_t)y * nx;
        const float r = row[wrap((int64_t)x - k, nx)], l = row[wrap((int64_t)x - 1 - k, nx)];
        out = (r - r * f) + l * f;
    } else if (q.kind == 1) {  // shearY_circ: column x moves by b (x - (nx-1)/2)
        const float skew = (float)((double)q.b * ((double)x - 0.5 * (nx - 1)));
        const int k = (int)floorf(skew);
        const float f = (float)((double)skew - 1.0 * k);
        const float r = s[(int64_t)wrap((int64_t)y - k, ny) * nx + x];
        const float l = s[(int64_t)wrap((int64_t)y - 1 - k, ny) * nx + x];
        out = (r - r * f) + l * f;
    } else if (q.kind == 2 && pass == 0) {  // rotate2d, 1-based (K, I) = (x+1, y+1)
        const int kc = nx / 2 + 1, ic = ny / 2 + 1;
        const double yy = (double)(y + 1 - ic), ycod = yy * q.cod + ic, ysid = -yy * q.sid + kc;
        const double xx = (double)(x + 1 - kc), xold = xx * q.cod + ysid, yold = xx * q.sid + ycod;
        const int iy = (int)yold, ix = (int)xold;
        out = 0.0f;
        if (iy >= 1 && iy <= ny - 1 && ix >= 1 && ix <= nx - 1) {
            const double yd = yold - iy, yr = 1.0 - yd, xd = xold - ix, xr = 1.0 - xd;
            const float *a = s + (int64_t)(iy - 1) * nx + (ix - 1);
            out = (float)(yd * (a[nx] * xr + a[nx + 1] * xd) + yr * (a[0] * xr + a[1] * xd));
        }
    } else {
        out = s[p];
    }
    d[p] = out;
}

// Correlation coefficient of a*w and b*w per view, w = (float)(hx[x] hy[y]);
// means and sums in double (getNormalizedCrossCorrelationWithFilter).
constexpr int kNccThreads = 1024;

__device__ double block_sum(double v, double *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = kNccThreads / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kNccThreads) tukey_ncc_kernel(const float *__restrict__ a,
                                                                const float *__restrict__ b, int nx, int ny,
                                                                const double *__restrict__ hx,
                                                                const double *__restrict__ hy,
                                                                double *__restrict__ out) {
    __shared__ double sh[kNccThreads];
    const int n = nx * ny;  // < 2^31 per view (checked by the host)
    const int64_t v = blockIdx.x;
    const float *av = a + v * n, *bv = b + v * n;
    double sa = 0.0, sb = 0.0;
    for (int i = threadIdx.x; i < n; i += kNccThreads) {
        const float w = (float)(hx[i % nx] * hy[i / nx]);
        sa += av[i] * w;
        sb += bv[i] * w;
    }
    const double ma = block_sum(sa, sh) / (1.0 * nx * ny), mb = block_sum(sb, sh) / (1.0 * nx * ny);
    double cv = 0.0, va = 0.0, vb = 0.0;
    for (int i = threadIdx.x; i < n; i += kNccThreads) {
        const float w = (float)(hx[i % nx] * hy[i / nx]);
        const double da = av[i] * w - ma, db = bv[i] * w - mb;
        cv += da * db;
        va += da * da;
        vb += db * db;
    }
    cv = block_sum(cv, sh);
    va = block_sum(va, sh);
    vb = block_sum(vb, sh);
    if (threadIdx.x == 0) out[v] = cv / sqrt(va * vb);
}

// dst[v] = src[v] for the views with keep[v] == 0.
__global__ void __launch_bounds__(kRotThreads) restore_views_kernel(float *__restrict__ dst,
                                                                    const float *__restrict__ src, int64_t n,
                                                                    const int *__restrict__ keep) {
    const int64_t p = (int64_t)blockIdx.x * kRotThreads + threadIdx.x, v = blockIdx.y;
    if (p >= n || keep[v]) return;
    dst[v * n + p] = src[v * n + p];
}

// Device buffers of one call, freed on exit.
struct Buf {
    void *p = nullptr;
    ~Buf() {
        if (p) hipFree(p);
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

int balloc(Buf &b, size_t bytes) {
    AREC_HIP(hipMalloc(&b.p, std::max<size_t>(bytes, 16)));
    return 0;
}

dim3 view_grid(int64_t per_view, int nv) {
    return dim3((unsigned)((per_view + kRotThreads - 1) / kRotThreads), (unsigned)nv);
}

// The Tukey window of length n, alpha 0.3 (tukey_filter, imagetools.cpp:332-367).
std::vector<double> tukey_window(int n) {
    const double alpha = 0.3, A = 2.0 * kPi / alpha;
    const double ta = 0.5 * alpha * (n - 1), tb = (1.0 - 0.5 * alpha) * (n - 1);
    std::vector<double> w(n, 1.0);
    for (int i = 0; i < n; i++) {
        if (i < ta)
            w[i] = 0.5 * (1.0 + cos(A * (1.0 * i / (n - 1) - alpha / 2)));
        else if (i > tb)
            w[i] = 0.5 * (1.0 + cos(A * (1.0 * i / (n - 1) - 1.0 + alpha / 2)));
    }
    return w;
}

// The three shears of a rotation by deg degrees (arecImage.cpp:733-736):
// theta in radians as a float, tan(theta/2) in double, the float sine.
RotView shear_params(float deg) {
    const float rad = (float)(deg * kPi / 180.0);
    RotView q{};
    q.kind = 1;
    q.a = (float)(-std::tan(-0.5 * rad));
    q.b = std::sin(-rad);
    return q;
}

// rotate2d's angle handling (align2dstack.cpp:783-792): float radians
// wrapped to [-pi, pi], cosine and sine of the float angle.
RotView bilinear_params(float deg) {
    float th = (float)(deg * kPi / 180.0);
    if (th > kPi) th = (float)(-2.0 * kPi + th);
    if (th < -kPi) th = (float)(2. * kPi + th);
    RotView q{};
    q.kind = 2;
    q.cod = std::cos(th);
    q.sid = std::sin(th);
    return q;
}

int run_rotation(float *views, float *tmp, int nx, int ny, int nv, const std::vector<RotView> &q,
                 hipStream_t st) {
    const int64_t n = (int64_t)nx * ny;
    Buf prm;
    int rc;
    if ((rc = balloc(prm, sizeof(RotView) * q.size()))) return rc;
    AREC_HIP(hipMemcpyAsync(prm.p, q.data(), sizeof(RotView) * q.size(), hipMemcpyHostToDevice, st));
    const dim3 g = view_grid(n, nv);
    rotate_pass_kernel<<<g, kRotThreads, 0, st>>>(views, tmp, nx, ny, prm.as<RotView>(), 0);
    rotate_pass_kernel<<<g, kRotThreads, 0, st>>>(tmp, views, nx, ny, prm.as<RotView>(), 1);
    rotate_pass_kernel<<<g, kRotThreads, 0, st>>>(views, tmp, nx, ny, prm.as<RotView>(), 2);
    AREC_HIP(hipGetLastError());
    AREC_HIP(hipMemcpyAsync(views, tmp, sizeof(float) * n * nv, hipMemcpyDeviceToDevice, st));
    // the parameters must outlive the kernels
    AREC_HIP(hipStreamSynchronize(st));
    return 0;
}

}  // namespace

int rot_check(int nx, int ny, int nv, const char *what) {
    if (nx < 2 || ny < 2 || nv < 0 || nv > 65535 || (int64_t)nx * ny > INT32_MAX) {
        set_error(std::string(what) + ": need views of at least 2x2 pixels, below 2^31 pixels, and at most "
                                      "65535 views a call");
        return AREC_EINVAL;
    }
    return 0;
}

int rot_cc_views(const float *img1, const float *img2, int nx, int ny, int nv, int r2, float *angles,
                 hipStream_t st) {
    int rc;
    if ((rc = rot_check(nx, ny, nv, "arecRotCCImages"))) return rc;
    if (r2 < 1 || 2 * (int64_t)r2 > std::min(nx, ny)) {  // keeps every bilinear sample inside the view
        set_error("arecRotCCImages: need 1 <= r2 <= min(nx, ny)/2");
        return AREC_EINVAL;
    }
    if (nv == 0) return 0;
    const int L = (int)(r2 * 2 * kPi);
    std::vector<double2> cs(L);
    for (int i = 0; i < L; i++) {
        const double alpha = 2 * kPi * i / L;
        cs[i] = make_double2(std::cos(alpha), std::sin(alpha));
    }
    Buf dcs, prof, corr, pos;
    if ((rc = balloc(dcs, sizeof(double2) * L)) || (rc = balloc(prof, sizeof(float) * 2 * (size_t)L * nv)) ||
        (rc = balloc(corr, sizeof(float) * (size_t)L * nv)) || (rc = balloc(pos, sizeof(float) * nv)))
        return rc;
    AREC_HIP(hipMemcpyAsync(dcs.p, cs.data(), sizeof(double2) * L, hipMemcpyHostToDevice, st));
    float *p1 = prof.as<float>(), *p2 = p1 + (size_t)L * nv;
    const dim3 g = view_grid(L, nv);
    radial_profile_kernel<<<g, kRotThreads, 0, st>>>(img1, nx, ny, L, r2, dcs.as<double2>(), p1);
    radial_profile_kernel<<<g, kRotThreads, 0, st>>>(img2, nx, ny, L, r2, dcs.as<double2>(), p2);
    ccorr1d_kernel<<<g, kRotThreads, 0, st>>>(p1, p2, L, corr.as<float>());
    peak1d_kernel<<<nv, kRotThreads, 0, st>>>(corr.as<float>(), L, pos.as<float>());
    AREC_HIP(hipGetLastError());
    AREC_HIP(hipMemcpyAsync(angles, pos.p, sizeof(float) * nv, hipMemcpyDeviceToHost, st));
    AREC_HIP(hipStreamSynchronize(st));
    // peak index -> degrees (arecImage.cpp:584-592), float as the reference
    for (int v = 0; v < nv; v++) {
        float a = angles[v];
        a = (float)(a * (2 * kPi / L));
        if (a > kPi) a = (float)(a - 2.0 * kPi);
        a = (float)(a * (180.0 / kPi));
        if (std::fabs(a) > 45) a = 0.0f;
        angles[v] = a;
    }
    return 0;
}

int rotate_views(float *views, int nx, int ny, int nv, float *angles, int rule, hipStream_t st) {
    int rc;
    if ((rc = rot_check(nx, ny, nv, "arecRotateImages"))) return rc;
    if (nv == 0) return 0;
    std::vector<RotView> q(nv);
    for (int v = 0; v < nv; v++) {
        if (rule == AREC_ROT_BILINEAR) {  // arecRotateImages: rotate2d for every view
            q[v] = bilinear_params(angles[v]);
            continue;
        }
        float &a = angles[v];  // arecRotateImages_skew's rules, angles updated
        if (a < -180) a += 360;
        if (a > 180) a -= 360;
        if (std::fabs(a) > 45) {
            a = 0.0f;
            q[v] = RotView{};
        } else if (std::fabs(a) > 10) {
            q[v] = bilinear_params(a);
        } else {
            q[v] = shear_params(a);
        }
    }
    Buf tmp;
    if ((rc = balloc(tmp, sizeof(float) * (size_t)nx * ny * nv))) return rc;
    return run_rotation(views, tmp.as<float>(), nx, ny, nv, q, st);
}

int rotate_views_safe(float *views, const float *ref, int nx, int ny, int nv, float *angles, int *kept,
                      double *mean_cc, hipStream_t st) {
    if (kept) *kept = 0;
    if (mean_cc) *mean_cc = 0.0;
    int rc;
    if ((rc = rot_check(nx, ny, nv, "arecRotateImages_skew_safe"))) return rc;
    if (nv == 0) return 0;
    for (int v = 0; v < nv; v++)
        if (!(std::fabs(angles[v]) <= 90.0f)) {
            set_error("arecRotateImages_skew_safe: the three shears take angles within +-90 degrees");
            return AREC_EINVAL;
        }
    const int64_t n = (int64_t)nx * ny;
    std::vector<RotView> q(nv);
    for (int v = 0; v < nv; v++) q[v] = shear_params(angles[v]);
    const std::vector<double> hx = tukey_window(nx), hy = tukey_window(ny);
    Buf backup, tmp, dhx, dhy, cc, keep;
    if ((rc = balloc(backup, sizeof(float) * n * nv)) || (rc = balloc(tmp, sizeof(float) * n * nv)) ||
        (rc = balloc(dhx, sizeof(double) * nx)) || (rc = balloc(dhy, sizeof(double) * ny)) ||
        (rc = balloc(cc, sizeof(double) * 2 * nv)) || (rc = balloc(keep, sizeof(int) * nv)))
        return rc;
    AREC_HIP(hipMemcpyAsync(backup.p, views, sizeof(float) * n * nv, hipMemcpyDeviceToDevice, st));
    AREC_HIP(hipMemcpyAsync(dhx.p, hx.data(), sizeof(double) * nx, hipMemcpyHostToDevice, st));
    AREC_HIP(hipMemcpyAsync(dhy.p, hy.data(), sizeof(double) * ny, hipMemcpyHostToDevice, st));
    if ((rc = run_rotation(views, tmp.as<float>(), nx, ny, nv, q, st))) return rc;
    double *c = cc.as<double>();
    tukey_ncc_kernel<<<nv, kNccThreads, 0, st>>>(views, ref, nx, ny, dhx.as<double>(), dhy.as<double>(), c);
    tukey_ncc_kernel<<<nv, kNccThreads, 0, st>>>(backup.as<float>(), ref, nx, ny, dhx.as<double>(),
                                                 dhy.as<double>(), c + nv);
    AREC_HIP(hipGetLastError());
    std::vector<double> h(2 * (size_t)nv);
    AREC_HIP(hipMemcpyAsync(h.data(), c, sizeof(double) * 2 * nv, hipMemcpyDeviceToHost, st));
    AREC_HIP(hipStreamSynchronize(st));
    std::vector<int> k(nv);
    double sum = 0.0;
    int nk = 0;
    for (int v = 0; v < nv; v++) {
        const double c1 = h[v], c2 = h[nv + v];
        k[v] = c2 + 0.00001 < c1;
        if (k[v]) {
            sum += c1;
            nk++;
            angles[v] *= 0.5;  // the reference's bookkeeping (arecImage.cpp:781)
        } else {
            angles[v] = 0.0f;
            sum += c2;
        }
    }
    AREC_HIP(hipMemcpyAsync(keep.p, k.data(), sizeof(int) * nv, hipMemcpyHostToDevice, st));
    restore_views_kernel<<<view_grid(n, nv), kRotThreads, 0, st>>>(views, backup.as<float>(), n, keep.as<int>());
    AREC_HIP(hipGetLastError());
    AREC_HIP(hipStreamSynchronize(st));
    if (kept) *kept = nk;
    if (mean_cc) *mean_cc = sum / nv;
    return 0;
}

}  // namespace arec

using namespace arec;

namespace {

// Upload n floats of each host stack, run op on the device copies, download
// the first stack when `back` is set.
template <class Op>
int host_stacks(const char *what, size_t n, float *a_host, const float *b_host, bool back, Op op) {
    hipStream_t st;
    AREC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    float *a = nullptr, *b = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&a, sizeof(float) * std::max<size_t>(n, 1)) != hipSuccess ||
        (b_host && hipMalloc((void **)&b, sizeof(float) * std::max<size_t>(n, 1)) != hipSuccess)) {
        rc = hip_fail(hipErrorOutOfMemory, what);
    } else {
        hipMemcpyAsync(a, a_host, sizeof(float) * n, hipMemcpyHostToDevice, st);
        if (b_host) hipMemcpyAsync(b, b_host, sizeof(float) * n, hipMemcpyHostToDevice, st);
        rc = op(a, b, st);
        if (!rc && back) {
            hipMemcpyAsync(a_host, a, sizeof(float) * n, hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) rc = hip_fail(hipGetLastError(), what);
        }
    }
    hipFree(a);
    hipFree(b);
    hipStreamDestroy(st);
    return rc;
}

}  // namespace

extern "C" {

int arec_rot_cc(arec_ctx *ctx, const float *images1_dev, const float *images2_dev, int32_t nx, int32_t ny,
                int32_t nviews, int32_t r2, float *angles) {
    if (!ctx || !images1_dev || !images2_dev || (!angles && nviews > 0)) {
        set_error("arec_rot_cc: NULL argument");
        return AREC_EINVAL;
    }
    return rot_cc_views(images1_dev, images2_dev, nx, ny, nviews, r2, angles, ctx->stream);
}

int arec_rotate_views(arec_ctx *ctx, float *views_dev, int32_t nx, int32_t ny, int32_t nviews, float *angles,
                      int32_t rule) {
    if (!ctx || !views_dev || (!angles && nviews > 0) || (rule != AREC_ROT_SKEW && rule != AREC_ROT_BILINEAR)) {
        set_error("arec_rotate_views: bad argument (rule AREC_ROT_SKEW or AREC_ROT_BILINEAR)");
        return AREC_EINVAL;
    }
    return rotate_views(views_dev, nx, ny, nviews, angles, rule, ctx->stream);
}

int arec_rotate_views_safe(arec_ctx *ctx, float *views_dev, const float *ref_dev, int32_t nx, int32_t ny,
                           int32_t nviews, float *angles, int32_t *kept, double *mean_cc) {
    if (!ctx || !views_dev || !ref_dev || (!angles && nviews > 0)) {
        set_error("arec_rotate_views_safe: NULL argument");
        return AREC_EINVAL;
    }
    return rotate_views_safe(views_dev, ref_dev, nx, ny, nviews, angles, kept, mean_cc, ctx->stream);
}

int arec_h_rot_cc_images(int32_t nx1, int32_t ny1, int32_t nz1, const float *images1, int32_t nx2, int32_t ny2,
                         int32_t nz2, const float *images2, int32_t r2, float *angles) {
    if (nx1 != nx2 || ny1 != ny2 || nz1 != nz2) {
        set_error("arecRotCCImages: mismatch in dimension");
        return -1;
    }
    if (!images1 || !images2 || !angles || nz1 < 1) {
        set_error("arec_h_rot_cc_images: bad argument");
        return AREC_EINVAL;
    }
    int rc;
    if ((rc = rot_check(nx1, ny1, nz1, "arecRotCCImages"))) return rc;
    return host_stacks("arec_h_rot_cc_images", (size_t)nx1 * ny1 * nz1, const_cast<float *>(images1), images2,
                       false, [&](float *a, float *b, hipStream_t st) {
                           return rot_cc_views(a, b, nx1, ny1, nz1, r2, angles, st);
                       });
}

int arec_h_rotate_images(int32_t nx, int32_t ny, int32_t nz, float *data, float *angles, int32_t rule) {
    if (!data || !angles || nz < 1 || (rule != AREC_ROT_SKEW && rule != AREC_ROT_BILINEAR)) {
        set_error("arec_h_rotate_images: bad argument");
        return AREC_EINVAL;
    }
    int rc;
    if ((rc = rot_check(nx, ny, nz, "arecRotateImages"))) return rc;
    return host_stacks("arec_h_rotate_images", (size_t)nx * ny * nz, data, nullptr, true,
                       [&](float *a, float *, hipStream_t st) { return rotate_views(a, nx, ny, nz, angles, rule, st); });
}

int arec_h_rotate_images_safe(int32_t nx, int32_t ny, int32_t nz, float *data, const float *ref, float *angles,
                              int32_t *kept, double *mean_cc) {
    if (!data || !ref || !angles || nz < 1) {
        set_error("arec_h_rotate_images_safe: bad argument");
        return AREC_EINVAL;
    }
    int rc;
    if ((rc = rot_check(nx, ny, nz, "arecRotateImages_skew_safe"))) return rc;
    return host_stacks("arec_h_rotate_images_safe", (size_t)nx * ny * nz, data, ref, true,
                       [&](float *a, float *b, hipStream_t st) {
                           return rotate_views_safe(a, b, nx, ny, nz, angles, kept, mean_cc, st);
                       });
}

}  // extern "C"
