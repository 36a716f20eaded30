// Translational alignment of measured views against reprojections, gfx950:
// batched 2-D FFTs, conj-multiply, sub-pixel peak search and Fourier shift.
//
//  ccorr2d  (align2dstack.cpp:480-521):  c = IDFT_unnorm(conj(DFT x) DFT y)
//  speak    (align2dstack.cpp:1059-1105): x-major scan for |c| > pmax with
//           the signed pmax, circular 3x3 neighbourhood, quad2dfit
//           (:1125-1166), wrap to (-n/2, n/2]
//  fshift2d (align2dstack.cpp:566-630):  DFT, phase ramp with the FP32
//           frequency tables and FP32 phase, inverse DFT, Re/(nx ny)
//
// FFTs are complex-to-complex in FP32 with FP64-accurate twiddle tables.
// Power-of-two lengths run a radix-2 transform in LDS (one workgroup per
// group of transforms); other lengths use a direct DFT (one output per
// thread), so every view size the reference accepts is handled on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "arec_internal.hpp"

namespace arec {

constexpr int kFftThreads = 256;
constexpr int kFftLds = 4096;  // complex values per workgroup

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Twiddles w[k] = exp(-2 pi i k / n), k < n (FP64-accurate, stored FP32).
struct Twiddles {
    int n = 0;
    float2 *d = nullptr;
};

static std::map<int, Twiddles> &twiddle_cache() {
    static std::map<int, Twiddles> c;
    return c;
}

// w[k] = exp(-2 pi i k / n) in FP64, rounded to FP32, k < n.
static void fill_twiddles(float2 *h, int n) {
    const double pi = 3.141592653589793238462643;
    for (int k = 0; k < n; k++) {
        double a = -2.0 * pi * (double)k / (double)n;
        h[k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
}

static int get_twiddles(int n, hipStream_t st, float2 **out) {
    auto &c = twiddle_cache();
    auto it = c.find(n);
    if (it == c.end()) {
        std::vector<float2> h(n);
        fill_twiddles(h.data(), n);
        Twiddles t;
        t.n = n;
        AREC_HIP(hipMalloc((void **)&t.d, sizeof(float2) * n));
        AREC_HIP(hipMemcpyAsync(t.d, h.data(), sizeof(float2) * n, hipMemcpyHostToDevice, st));
        AREC_HIP(hipStreamSynchronize(st));
        it = c.emplace(n, t).first;
    }
    *out = it->second.d;
    return 0;
}

// Radix-2 FFT of `cnt` transforms of length n = 2^logn, in place.  Element m
// of transform t of batch b sits at data[b*bstride + t*tstride + m*estride].
// One workgroup handles `cnt` transforms (cnt*n <= kFftLds) of one batch.
// dir = -1 forward, +1 inverse (unnormalised).
__global__ __launch_bounds__(kFftThreads) void fft_pow2_kernel(float2 *__restrict__ data, int n,
                                                               int logn, int64_t estride,
                                                               int64_t tstride, int64_t bstride,
                                                               int ntrans, int cnt,
                                                               const float2 *__restrict__ tw, int dir) {
    __shared__ float2 s[kFftLds];
    const int groups = (ntrans + cnt - 1) / cnt;
    const int b = blockIdx.x / groups;
    const int t0 = (blockIdx.x % groups) * cnt;
    const int c_here = min(cnt, ntrans - t0);
    float2 *base = data + (int64_t)b * bstride + (int64_t)t0 * tstride;
    // load with bit-reversed placement; consecutive threads take consecutive
    // transforms first so that column transforms read contiguous rows
    for (int e = threadIdx.x; e < c_here * n; e += blockDim.x) {
        int t = e % c_here, m = e / c_here;
        int r = __brev(m) >> (32 - logn);
        s[t * n + r] = base[(int64_t)t * tstride + (int64_t)m * estride];
    }
    __syncthreads();
    for (int half = 1; half < n; half <<= 1) {
        const int step = n / (2 * half);
        for (int e = threadIdx.x; e < c_here * (n / 2); e += blockDim.x) {
            int t = e / (n / 2), k = e % (n / 2);
            int grp = k / half, j = k % half;
            int i0 = t * n + grp * 2 * half + j, i1 = i0 + half;
            float2 w = tw[j * step];
            if (dir > 0) w.y = -w.y;
            float2 u = s[i0], v = cmul(s[i1], w);
            s[i0] = make_float2(u.x + v.x, u.y + v.y);
            s[i1] = make_float2(u.x - v.x, u.y - v.y);
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < c_here * n; e += blockDim.x) {
        int t = e % c_here, m = e / c_here;
        base[(int64_t)t * tstride + (int64_t)m * estride] = s[t * n + m];
    }
}

// Direct DFT for lengths that are not powers of two: out-of-place, one
// output element per thread.
__global__ void dft_direct_kernel(const float2 *__restrict__ in, float2 *__restrict__ out, int n,
                                  int64_t estride, int64_t tstride, int64_t bstride, int ntrans,
                                  int64_t total, const float2 *__restrict__ tw, int dir) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int k = (int)(e % n);
        int64_t rest = e / n;
        int t = (int)(rest % ntrans);
        int64_t b = rest / ntrans;
        const float2 *src = in + b * bstride + (int64_t)t * tstride;
        double sr = 0.0, si = 0.0;
        for (int m = 0; m < n; m++) {
            float2 w = tw[(int)(((int64_t)k * m) % n)];
            double wi = dir > 0 ? -(double)w.y : (double)w.y;
            float2 v = src[(int64_t)m * estride];
            sr += (double)v.x * w.x - (double)v.y * wi;
            si += (double)v.x * wi + (double)v.y * w.x;
        }
        out[b * bstride + (int64_t)t * tstride + (int64_t)k * estride] = make_float2((float)sr, (float)si);
    }
}

__global__ void copy2_kernel(float2 *__restrict__ dst, const float2 *__restrict__ src, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// One batch of 1-D transforms along one axis.  tw: the length's twiddles,
// or NULL for the cached table.
static int fft_axis(float2 *data, float2 *tmp, int n, int64_t estride, int64_t tstride,
                    int64_t bstride, int ntrans, int nbatch, int dir, hipStream_t st,
                    const float2 *tw_given = nullptr) {
    float2 *tw = const_cast<float2 *>(tw_given);
    if (!tw) {
        int rc = get_twiddles(n, st, &tw);
        if (rc) return rc;
    }
    if (n == 1) return 0;
    if ((n & (n - 1)) == 0 && n <= kFftLds) {
        int logn = 0;
        while ((1 << logn) < n) logn++;
        int cnt = std::max(1, std::min(kFftLds / n, 16));
        cnt = std::min(cnt, ntrans);
        int groups = (ntrans + cnt - 1) / cnt;
        fft_pow2_kernel<<<groups * nbatch, kFftThreads, 0, st>>>(data, n, logn, estride, tstride, bstride,
                                                                ntrans, cnt, tw, dir);
    } else {
        int64_t total = (int64_t)n * ntrans * nbatch;
        dft_direct_kernel<<<(unsigned)std::min<int64_t>((total + 255) / 256, 65535), 256, 0, st>>>(
            data, tmp, n, estride, tstride, bstride, ntrans, total, tw, dir);
        // copy back: the layout of tmp equals data's for the touched elements
        int64_t span = (int64_t)nbatch * bstride;
        copy2_kernel<<<(unsigned)std::min<int64_t>((span + 255) / 256, 65535), 256, 0, st>>>(data, tmp, span);
    }
    AREC_HIP(hipGetLastError());
    return 0;
}

// 2-D transform of nv views of [ny][nx] complex values (twx, twy: the
// twiddles of nx and ny, or NULL for the cached tables).
static int fft2(float2 *data, float2 *tmp, int nx, int ny, int nv, int dir, hipStream_t st,
                const float2 *twx = nullptr, const float2 *twy = nullptr) {
    int rc = fft_axis(data, tmp, nx, 1, nx, (int64_t)nx * ny, ny, nv, dir, st, twx);
    if (!rc) rc = fft_axis(data, tmp, ny, nx, 1, (int64_t)nx * ny, nx, nv, dir, st, twy);
    return rc;
}

__global__ void pack_real_kernel(const float *__restrict__ x, float2 *__restrict__ c, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c[i] = make_float2(x[i], 0.0f);
}

// conj(X) * Y (align2dstack.cpp:503-506, FP32)
__global__ void conj_mul_kernel(float2 *__restrict__ X, const float2 *__restrict__ Y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float2 a = X[i], b = Y[i];
        X[i] = make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
    }
}

// Re(c) / div (fshift2d divides by (float)(nx*ny), align2dstack.cpp:619-621)
__global__ void real_part_kernel(const float2 *__restrict__ c, float *__restrict__ out, int64_t n, float div) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = c[i].x / div;
}

// Sub-pixel peak of a 3x3 neighbourhood (quad2dfit, align2dstack.cpp:1125-
// 1166), FP64 on FP32 samples z[k], k = 3*row + column.  Least-squares fit
// of q(u,v) = a + b u + c v + d u^2 + e u v + f v^2 on u, v in {-1, 0, 1}
// (u along x): the grid's moments give a = (S0 - 6 (d + f))/9,
// b = Su/6, c = Sv/6, e = Suv/4, d + f = (Suu + Svv)/2 - 2 S0/3 and
// d - f = (Suu - Svv)/2; the stationary point solves [2d e; e 2f] (u, v) =
// -(b, c) and is clamped to [-1, 1]; *peakv (when not NULL) gets q at the
// unclamped stationary point.  With a singular system the shifts are 0 and
// *peakv is left as it was, as in the reference.  Host and device share it.
__host__ __device__ inline void quad2dfit_dev(const float *z, float *xsh, float *ysh, float *peakv = nullptr) {
    double s0 = 0, su = 0, sv = 0, suu = 0, svv = 0, suv = 0;
    for (int k = 0; k < 9; k++) {
        double u = (double)(k % 3 - 1), v = (double)(k / 3 - 1), w = (double)z[k];
        s0 += w;
        su += u * w;
        sv += v * w;
        suu += u * u * w;
        svv += v * v * w;
        suv += u * v * w;
    }
    double b = su / 6.0, c = sv / 6.0, e = suv / 4.0;
    double dpf = 0.5 * (suu + svv) - 2.0 * s0 / 3.0, dmf = 0.5 * (suu - svv);
    double d = 0.5 * (dpf + dmf), f = 0.5 * (dpf - dmf);
    double det = 4.0 * d * f - e * e;
    float xs = 0.0f, ys = 0.0f;
    if (det != 0.0) {
        const double u = (e * c - 2.0 * f * b) / det, v = (e * b - 2.0 * d * c) / det;
        xs = (float)fmin(1.0, fmax(-1.0, u));
        ys = (float)fmin(1.0, fmax(-1.0, v));
        if (peakv) {
            const double a = (s0 - 6.0 * dpf) / 9.0;
            *peakv = (float)(a + b * u + c * v + d * u * u + e * u * v + f * v * v);
        }
    }
    *xsh = xs;
    *ysh = ys;
}

// speak on each view's correlation map (one workgroup per view).  The scan
// order key is x*ny + y; the reference keeps the first strictly larger |c|
// and stores the signed value.  A positive winner is the sequential scan's
// answer; after a negative winner the sequential state machine is replayed
// exactly by one thread from that point on.
__global__ __launch_bounds__(256) void speak_kernel(const float *__restrict__ cmap, int nx, int ny,
                                                    float *__restrict__ sx, float *__restrict__ sy) {
    __shared__ float s_v[256];
    __shared__ int s_k[256];
    __shared__ int s_ix, s_iy;
    const int v = blockIdx.x;
    const float *c = cmap + (int64_t)v * nx * ny;
    const int n = nx * ny;
    float best = 0.0f;
    int bestk = n;  // none
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        int x = k / ny, y = k % ny;  // scan order: x outer, y inner
        float a = fabsf(c[(int64_t)y * nx + x]);
        if (a > best) {  // NaN never wins, like the reference's comparison
            best = a;
            bestk = k;
        }
    }
    s_v[threadIdx.x] = best;
    s_k[threadIdx.x] = bestk;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            float o = s_v[threadIdx.x + off];
            int ok = s_k[threadIdx.x + off];
            if (o > s_v[threadIdx.x] || (o == s_v[threadIdx.x] && ok < s_k[threadIdx.x])) {
                s_v[threadIdx.x] = o;
                s_k[threadIdx.x] = ok;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        int k = s_k[0];
        int ix = 1, iy = 1;  // no element above 0: reference indices undefined
        if (k < n) {
            float val = c[(int64_t)(k % ny) * nx + k / ny];
            if (val > 0.0f) {
                ix = k / ny + 1;
                iy = k % ny + 1;
            } else {
                // exact replay of the sequential scan from the first element
                float pmax = 0.0f;
                for (int kk = 0; kk < n; kk++) {
                    int x = kk / ny, y = kk % ny;
                    float d = c[(int64_t)y * nx + x];
                    if (fabsf(d) > pmax) {
                        ix = x + 1;
                        iy = y + 1;
                        pmax = d;
                    }
                }
            }
        }
        s_ix = ix;
        s_iy = iy;
        float rsq[9];
        for (int j = -1; j <= 1; j++) {
            int jc = iy + j;
            if (jc < 1) jc += ny;
            if (jc > ny) jc -= ny;
            for (int i = -1; i <= 1; i++) {
                int ir = ix + i;
                if (ir < 1) ir += nx;
                if (ir > nx) ir -= nx;
                rsq[(j + 1) * 3 + i + 1] = c[(int64_t)(jc - 1) * nx + (ir - 1)];
            }
        }
        float xsh, ysh;
        quad2dfit_dev(rsq, &xsh, &ysh);
        float xt = xsh + (float)(ix - 1), yt = ysh + (float)(iy - 1);
        if (xt > nx / 2) xt -= (float)nx;
        if (yt > ny / 2) yt -= (float)ny;
        sx[v] = xt;
        sy[v] = yt;
    }
}

// Phase ramp of fshift2d with the FP32 tables and FP32 phase
// (align2dstack.cpp:593-613); shift by (sgn*sx[v], sgn*sy[v]).
__global__ void phase_ramp_kernel(float2 *__restrict__ F, int nx, int ny, int nv,
                                  const float *__restrict__ sx, const float *__restrict__ sy, float sgn) {
    const int64_t total = (int64_t)nx * ny * nv;
    const int hx = (nx + 1) / 2, hy = (ny + 1) / 2;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        int i = (int)(e % nx);
        int64_t r = e / nx;
        int j = (int)(r % ny);
        int v = (int)(r / ny);
        float kx = i < hx ? (float)i / (float)nx : (float)(i - nx) / (float)nx;
        float ky = j < hy ? (float)j / (float)ny : (float)(j - ny) / (float)ny;
        float ssx = sgn * sx[v], ssy = sgn * sy[v];
        float t1 = __fmul_rn(ssx, kx), t2 = __fmul_rn(ssy, ky);
        float phase = (float)(-2.0 * 3.141592653589793 * (double)__fadd_rn(t1, t2));
        float cph, sph;
        sincosf(phase, &sph, &cph);
        float2 f = F[e];
        F[e] = make_float2(f.x * cph - f.y * sph, f.x * sph + f.y * cph);
    }
}

static dim3 ew_grid(int64_t n) {
    return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)));
}

// Workspace for batches of views.
struct AlignWork {
    float2 *a = nullptr, *b = nullptr, *t = nullptr;
    size_t cap = 0;
    ~AlignWork() {
        hipFree(a);
        hipFree(b);
        hipFree(t);
    }
    int ensure(size_t n) {
        if (n <= cap) return 0;
        hipFree(a);
        hipFree(b);
        hipFree(t);
        a = b = t = nullptr;
        AREC_HIP(hipMalloc((void **)&a, sizeof(float2) * n));
        AREC_HIP(hipMalloc((void **)&b, sizeof(float2) * n));
        AREC_HIP(hipMalloc((void **)&t, sizeof(float2) * n));
        cap = n;
        return 0;
    }
};

static int views_per_batch(int64_t nx, int64_t ny, int64_t nv) {
    int64_t per = nx * ny;
    int64_t vb = std::max<int64_t>(1, (int64_t)(256ll << 20) / per);  // <= 2 GiB per buffer
    return (int)std::min<int64_t>(vb, nv);
}

// The correlation of views [v0, v0 + n) in batches of vb views through the
// work buffers a, b, t (vb*nx*ny complex values each) and cbuf (vb*nx*ny
// floats, used when cmap_out is NULL).  Every buffer element a batch reads
// it has written first.
static int cc_batches(const float *proj, const float *meas, int nx, int ny, int nv, float *cmap_out, float *sx,
                      float *sy, hipStream_t st, int vb, float2 *a, float2 *b, float2 *t, float *cbuf,
                      const float2 *twx, const float2 *twy) {
    const int64_t per = (int64_t)nx * ny;
    int rc = 0;
    for (int v0 = 0; v0 < nv; v0 += vb) {
        int n = std::min(vb, nv - v0);
        int64_t cnt = (int64_t)n * per;
        pack_real_kernel<<<ew_grid(cnt), 256, 0, st>>>(proj + v0 * per, a, cnt);
        pack_real_kernel<<<ew_grid(cnt), 256, 0, st>>>(meas + v0 * per, b, cnt);
        if ((rc = fft2(a, t, nx, ny, n, -1, st, twx, twy)) || (rc = fft2(b, t, nx, ny, n, -1, st, twx, twy))) break;
        conj_mul_kernel<<<ew_grid(cnt), 256, 0, st>>>(a, b, cnt);
        if ((rc = fft2(a, t, nx, ny, n, +1, st, twx, twy))) break;
        float *dst = cmap_out ? cmap_out + v0 * per : cbuf;
        real_part_kernel<<<ew_grid(cnt), 256, 0, st>>>(a, dst, cnt, 1.0f);
        if (sx && sy) speak_kernel<<<n, 256, 0, st>>>(dst, nx, ny, sx + v0, sy + v0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            rc = hip_fail(e, "cc_views");
            break;
        }
    }
    return rc;
}

// c (optional, [nv][ny][nx]) = correlation maps; sx, sy (optional) = peaks.
int cc_views(const float *proj, const float *meas, int nx, int ny, int nv, float *cmap_out,
             float *sx, float *sy, hipStream_t st) {
    AlignWork w;
    const int vb = views_per_batch(nx, ny, nv);
    const int64_t per = (int64_t)nx * ny;
    int rc;
    if ((rc = w.ensure((size_t)vb * per))) return rc;
    float *cbuf = nullptr;
    if (!cmap_out) AREC_HIP(hipMalloc((void **)&cbuf, sizeof(float) * vb * per));
    rc = cc_batches(proj, meas, nx, ny, nv, cmap_out, sx, sy, st, vb, w.a, w.b, w.t, cbuf, nullptr, nullptr);
    hipStreamSynchronize(st);
    hipFree(cbuf);
    return rc;
}

// ---------------------------------------------------------------------------
// Cross-correlation alignment without allocation.  Its work buffers and
// twiddles live in `scratch`, a static device array of kNoAllocScratchBytes
// that every GPU the library runs on holds until the process ends.  In host
// code `scratch` names a host-side stand-in, so its device address is looked
// up once per device (noalloc_setup, called by arec_create with the device
// current) and only that address is handed to kernels and copies.  Calls
// that overlap on one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(128) << 20;
constexpr int kNoAllocMaxSide = 32768;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static void *g_scratch[kNoAllocMaxDevices];           // device address of scratch, per device
static std::mutex g_scratch_mu;                        // guards g_scratch and g_tw_host
static float2 g_tw_host[2 * kNoAllocMaxSide];          // twiddle staging (no host allocation)

int noalloc_setup(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) {
        set_error("no-allocation scratch: device index beyond the table");
        return AREC_EINVAL;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return 0;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = p;
    return 0;
}

static size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// Bytes of scratch a batch of vb views of nx x ny needs: both lengths'
// twiddles, three complex work buffers and one real correlation map.
static size_t noalloc_bytes(int nx, int ny, int64_t vb) {
    const size_t per = (size_t)nx * ny;
    return align256(sizeof(float2) * ((size_t)nx + ny)) + 3 * align256(sizeof(float2) * per * vb) +
           align256(sizeof(float) * per * vb);
}

int cc_views_noalloc(int device, const float *proj, const float *meas, int nx, int ny, int nv, float *sx, float *sy,
                     hipStream_t st) {
    if (nx > kNoAllocMaxSide || ny > kNoAllocMaxSide || noalloc_bytes(nx, ny, 1) > kNoAllocScratchBytes) {
        set_error("arec_align_cc_noalloc: a view of " + std::to_string(nx) + " x " + std::to_string(ny) +
                  " needs " + std::to_string(noalloc_bytes(nx, ny, 1)) + " bytes of scratch; the fixed scratch holds " +
                  std::to_string(kNoAllocScratchBytes) + " (and sides up to " + std::to_string(kNoAllocMaxSide) + ")");
        return AREC_ESIZE;
    }
    if (device < 0 || device >= kNoAllocMaxDevices) {
        set_error("arec_align_cc_noalloc: device index beyond the scratch table");
        return AREC_EINVAL;
    }
    if (!g_scratch[device]) {  // arec_create's lookup failed: once more, with the context's device current
        AREC_HIP(hipSetDevice(device));
        int rc = noalloc_setup(device);
        if (rc) return rc;
    }
    const size_t per = (size_t)nx * ny;
    const size_t room = kNoAllocScratchBytes - align256(sizeof(float2) * ((size_t)nx + ny)) - 4 * 256;
    int64_t vb = std::max<int64_t>(1, std::min<int64_t>(nv, (int64_t)(room / (28 * per))));
    while (vb > 1 && noalloc_bytes(nx, ny, vb) > kNoAllocScratchBytes) vb--;
    unsigned char *base = static_cast<unsigned char *>(g_scratch[device]);
    float2 *twx = reinterpret_cast<float2 *>(base);
    float2 *twy = twx + nx;
    unsigned char *q = base + align256(sizeof(float2) * ((size_t)nx + ny));
    float2 *a = reinterpret_cast<float2 *>(q);
    float2 *b = reinterpret_cast<float2 *>(q + align256(sizeof(float2) * per * vb));
    float2 *t = reinterpret_cast<float2 *>(q + 2 * align256(sizeof(float2) * per * vb));
    float *cbuf = reinterpret_cast<float *>(q + 3 * align256(sizeof(float2) * per * vb));
    {
        // the same FP32 twiddles as the cached tables, staged through a
        // static host array; the copy completes before the lock is released
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        fill_twiddles(g_tw_host, nx);
        fill_twiddles(g_tw_host + nx, ny);
        AREC_HIP(hipMemcpyAsync(twx, g_tw_host, sizeof(float2) * ((size_t)nx + ny), hipMemcpyHostToDevice, st));
        AREC_HIP(hipStreamSynchronize(st));
    }
    int rc = cc_batches(proj, meas, nx, ny, nv, nullptr, sx, sy, st, (int)vb, a, b, t, cbuf, twx, twy);
    hipStreamSynchronize(st);
    return rc;
}

// data[v] <- fshift2d(data[v], sgn*sx[v], sgn*sy[v]); arecShiftImages uses
// sgn = -1 (arecImage.cpp:664).
int shift_views(float *data, int nx, int ny, int nv, const float *sx, const float *sy, float sgn,
                hipStream_t st) {
    AlignWork w;
    const int vb = views_per_batch(nx, ny, nv);
    const int64_t per = (int64_t)nx * ny;
    int rc;
    if ((rc = w.ensure((size_t)vb * per))) return rc;
    for (int v0 = 0; v0 < nv; v0 += vb) {
        int n = std::min(vb, nv - v0);
        int64_t cnt = (int64_t)n * per;
        pack_real_kernel<<<ew_grid(cnt), 256, 0, st>>>(data + v0 * per, w.a, cnt);
        if ((rc = fft2(w.a, w.t, nx, ny, n, -1, st))) break;
        phase_ramp_kernel<<<ew_grid(cnt), 256, 0, st>>>(w.a, nx, ny, n, sx + v0, sy + v0, sgn);
        if ((rc = fft2(w.a, w.t, nx, ny, n, +1, st))) break;
        real_part_kernel<<<ew_grid(cnt), 256, 0, st>>>(w.a, data + v0 * per, cnt, (float)(nx * ny));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            rc = hip_fail(e, "shift_views");
            break;
        }
    }
    hipStreamSynchronize(st);
    return rc;
}

// ---------------------------------------------------------------------------
// Preprocessing of the measured views (the driver's steps before the first
// reconstruction, AREC3D.cpp:81,105).
// ---------------------------------------------------------------------------

// arecImageTakeLog (arecImage.cpp:358-381): x > 0 -> -log(x), else 0.
__global__ void take_log_kernel(float *__restrict__ x, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float v = x[i];
        x[i] = v > 0.0f ? -logf(v) : 0.0f;
    }
}

// Moisan's boundary image of every view of a batch (makePeriodicSlice,
// imagetools.cpp:26-35), as complex values with zero imaginary part: the
// rows' jumps assigned (row ny-1 last), then the columns' jumps added
// (column nx-1 last), in FP32 -- the reference's order, which matters only
// for ny == 1 or nx == 1.
__global__ void boundary_image_kernel(const float *__restrict__ u, float2 *__restrict__ c, int nx, int ny,
                                      int64_t n) {
    const int64_t per = (int64_t)nx * ny;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = e / per;
        const int k = (int)(e - v * per), iy = k / nx, ix = k - iy * nx;
        const float *__restrict__ uv = u + v * per;
        const int64_t top = ix, bot = (int64_t)(ny - 1) * nx + ix;
        const int64_t left = (int64_t)iy * nx, right = left + nx - 1;
        float val = 0.0f;
        if (iy == ny - 1)
            val = uv[top] - uv[bot];
        else if (iy == 0)
            val = uv[bot] - uv[top];
        if (ix == 0) val += uv[right] - uv[left];
        if (ix == nx - 1) val += uv[left] - uv[right];
        c[e] = make_float2(val, 0.0f);
    }
}

// S = V / (2 cos(2 pi i/nx) + 2 cos(2 pi j/ny) - 4) with S(0,0) = 0
// (imagetools.cpp:53-62: the quotient in double, stored as float).
__global__ void periodic_divide_kernel(float2 *__restrict__ f, int nx, int ny, int64_t n) {
    const double pi = 3.14159265358979323846;
    const int64_t per = (int64_t)nx * ny;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % per), j = k / nx, i = k - j * nx;
        if (k == 0) {
            f[e] = make_float2(0.0f, 0.0f);
            continue;
        }
        const double den = 2 * cos(2.0 * pi * i / nx) + 2 * cos(2.0 * pi * j / ny) - 4;
        f[e] = make_float2((float)(f[e].x / den), (float)(f[e].y / den));
    }
}

// u -= W(d) s with s = real part / (float)(nx ny) (the unnormalised inverse
// transform, imagetools.cpp:68-75), W = exp(-d^2 / (2 * 5^2)), d the distance
// to the nearest border (periodicDecomposition, imagetools.cpp:95-109).
__global__ void periodic_apply_kernel(const float2 *__restrict__ c, float *__restrict__ u, int nx, int ny,
                                      int64_t n, float div) {
    const int64_t per = (int64_t)nx * ny;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int k = (int)(e % per), iy = k / nx, ix = k - iy * nx;
        const int d = min(min(ix, nx - 1 - ix), min(iy, ny - 1 - iy));
        const float s = c[e].x / div;
        const double w = exp(-1.0 * d * d / (2 * 5.0 * 5.0));
        u[e] = (float)((double)u[e] - w * (double)s);
    }
}

// Zeroth and first moments of every view of a stack [nv][ny][nx], in FP64:
// mom[3v..3v+2] = {sum p, sum u*p, sum y*p} over the view's pixels (u, y).
// One workgroup per view; each thread sums a fixed set of columns of every
// row, then a fixed-order tree: the result does not depend on scheduling.
__global__ __launch_bounds__(256) void view_moments_kernel(const float *__restrict__ views, int nx, int ny,
                                                           double *__restrict__ mom) {
    __shared__ double red[3][256];
    const int v = blockIdx.x, tid = threadIdx.x;
    const float *__restrict__ p = views + (int64_t)v * nx * ny;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int y = 0; y < ny; y++) {
        const float *__restrict__ row = p + (int64_t)y * nx;
        double r0 = 0.0, r1 = 0.0;
        for (int u = tid; u < nx; u += 256) {
            const double q = row[u];
            r0 += q;
            r1 += q * u;
        }
        s0 += r0;
        s1 += r1;
        s2 += r0 * y;
    }
    red[0][tid] = s0;
    red[1][tid] = s1;
    red[2][tid] = s2;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h)
            for (int k = 0; k < 3; k++) red[k][tid] += red[k][tid + h];
        __syncthreads();
    }
    if (tid < 3) mom[3 * (int64_t)v + tid] = red[tid][0];
}

int take_log(float *x, int64_t n, hipStream_t st) {
    if (n <= 0) return 0;
    take_log_kernel<<<ew_grid(n), 256, 0, st>>>(x, n);
    AREC_HIP(hipGetLastError());
    return 0;
}

// periodicDecomposition (imagetools.cpp:86-114) of every view, in batches.
int periodic_views(float *data, int nx, int ny, int nv, hipStream_t st) {
    AlignWork w;
    const int vb = views_per_batch(nx, ny, nv);
    const int64_t per = (int64_t)nx * ny;
    int rc;
    if ((rc = w.ensure((size_t)vb * per))) return rc;
    for (int v0 = 0; v0 < nv; v0 += vb) {
        const int n = std::min(vb, nv - v0);
        const int64_t cnt = (int64_t)n * per;
        boundary_image_kernel<<<ew_grid(cnt), 256, 0, st>>>(data + v0 * per, w.a, nx, ny, cnt);
        if ((rc = fft2(w.a, w.t, nx, ny, n, -1, st))) break;
        periodic_divide_kernel<<<ew_grid(cnt), 256, 0, st>>>(w.a, nx, ny, cnt);
        if ((rc = fft2(w.a, w.t, nx, ny, n, +1, st))) break;
        periodic_apply_kernel<<<ew_grid(cnt), 256, 0, st>>>(w.a, data + v0 * per, nx, ny, cnt, (float)(nx * ny));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            rc = hip_fail(e, "periodic_views");
            break;
        }
    }
    hipStreamSynchronize(st);
    return rc;
}

int view_moments(const float *views, int nx, int ny, int nv, double *mom_dev, hipStream_t st) {
    view_moments_kernel<<<nv, 256, 0, st>>>(views, nx, ny, mom_dev);
    AREC_HIP(hipGetLastError());
    return 0;
}

// One view through device copies: n floats of in1 (and of in2, when given)
// are uploaded to a (b); `run` computes into the device buffers (c holds at
// least max(n, n_out, 2) floats) and returns the device pointer whose n_out
// floats are downloaded to `out`.
template <typename F>
static int one_view(const float *in1, const float *in2, int64_t n, float *out, int64_t n_out, const char *what,
                    F run) {
    hipStream_t st;
    AREC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    float *a = nullptr, *b = nullptr, *c = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&a, n * 4) != hipSuccess || (in2 && hipMalloc((void **)&b, n * 4) != hipSuccess) ||
        hipMalloc((void **)&c, std::max<int64_t>(std::max<int64_t>(n, n_out), 2) * 4) != hipSuccess) {
        rc = hip_fail(hipErrorOutOfMemory, what);
    } else {
        hipMemcpyAsync(a, in1, n * 4, hipMemcpyHostToDevice, st);
        if (in2) hipMemcpyAsync(b, in2, n * 4, hipMemcpyHostToDevice, st);
        const float *res = nullptr;
        rc = run(a, b, c, st, &res);
        if (!rc) {
            hipMemcpyAsync(out, res, n_out * 4, hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) rc = hip_fail(hipGetLastError(), what);
        }
    }
    hipFree(a);
    hipFree(b);
    hipFree(c);
    hipStreamDestroy(st);
    return rc;
}

}  // namespace arec

using namespace arec;

extern "C" {

int arec_align_ccorr(arec_ctx *ctx, const float *x_dev, const float *y_dev, int32_t nx, int32_t ny,
                     int32_t nviews, float *c_dev) {
    if (!ctx || !x_dev || !y_dev || !c_dev || nx < 1 || ny < 1 || nviews < 1) {
        set_error("arec_align_ccorr: bad argument");
        return AREC_EINVAL;
    }
    return cc_views(x_dev, y_dev, nx, ny, nviews, c_dev, nullptr, nullptr, ctx->stream);
}

int arec_align_cc(arec_ctx *ctx, const float *proj_dev, const float *meas_dev, int32_t nx, int32_t ny,
                  int32_t nviews, float *sx_dev, float *sy_dev) {
    if (!ctx || !proj_dev || !meas_dev || !sx_dev || !sy_dev || nx < 1 || ny < 1 || nviews < 1) {
        set_error("arec_align_cc: bad argument");
        return AREC_EINVAL;
    }
    return cc_views(proj_dev, meas_dev, nx, ny, nviews, nullptr, sx_dev, sy_dev, ctx->stream);
}

int arec_align_cc_noalloc(arec_ctx *ctx, const float *proj_dev, const float *meas_dev, int32_t nx, int32_t ny,
                          int32_t nviews, float *sx_dev, float *sy_dev) {
    if (!ctx || !proj_dev || !meas_dev || !sx_dev || !sy_dev || nx < 1 || ny < 1 || nviews < 1) {
        set_error("arec_align_cc_noalloc: bad argument");
        return AREC_EINVAL;
    }
    return cc_views_noalloc(ctx->device, proj_dev, meas_dev, nx, ny, nviews, sx_dev, sy_dev, ctx->stream);
}

int arec_align_peak(arec_ctx *ctx, const float *c_dev, int32_t nx, int32_t ny, int32_t nviews,
                    float *sx_dev, float *sy_dev) {
    if (!ctx || !c_dev || !sx_dev || !sy_dev || nx < 1 || ny < 1 || nviews < 1) {
        set_error("arec_align_peak: bad argument");
        return AREC_EINVAL;
    }
    speak_kernel<<<nviews, 256, 0, ctx->stream>>>(c_dev, nx, ny, sx_dev, sy_dev);
    AREC_HIP(hipGetLastError());
    return 0;
}

int arec_take_log(arec_ctx *ctx, float *data_dev, int64_t n) {
    if (!ctx || (!data_dev && n > 0) || n < 0) {
        set_error("arec_take_log: bad argument");
        return AREC_EINVAL;
    }
    return take_log(data_dev, n, ctx->stream);
}

int arec_periodic_decomposition(arec_ctx *ctx, float *views_dev, int32_t nx, int32_t ny, int32_t nviews) {
    if (!ctx || !views_dev || nx < 1 || ny < 1 || nviews < 1) {
        set_error("arec_periodic_decomposition: bad argument");
        return AREC_EINVAL;
    }
    return periodic_views(views_dev, nx, ny, nviews, ctx->stream);
}

int arec_align_shift(arec_ctx *ctx, float *data_dev, int32_t nx, int32_t ny, int32_t nviews,
                     const float *sx_dev, const float *sy_dev, float sign) {
    if (!ctx || !data_dev || !sx_dev || !sy_dev || nx < 1 || ny < 1 || nviews < 1) {
        set_error("arec_align_shift: bad argument");
        return AREC_EINVAL;
    }
    return shift_views(data_dev, nx, ny, nviews, sx_dev, sy_dev, sign, ctx->stream);
}

// Host-pointer versions: arecCCImages (arecImage.cpp:384-434; status -1
// when the stacks differ in size) and arecShiftImages (:653-671).
int arec_h_cc_images(int32_t nx1, int32_t ny1, int32_t nz1, const float *images1, int32_t nx2,
                     int32_t ny2, int32_t nz2, const float *images2, float *sx, float *sy) {
    if (nx1 != nx2 || ny1 != ny2 || nz1 != nz2) {
        set_error("arecCCImages: mismatch in dimension");
        return -1;
    }
    if (!images1 || !images2 || !sx || !sy || nx1 < 1 || ny1 < 1 || nz1 < 1) {
        set_error("arec_h_cc_images: bad argument");
        return AREC_EINVAL;
    }
    int dev = 0;
    hipStream_t st;
    AREC_HIP(hipGetDevice(&dev));
    AREC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    size_t n = (size_t)nx1 * ny1 * nz1;
    float *a = nullptr, *b = nullptr, *s = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&a, n * 4) != hipSuccess || hipMalloc((void **)&b, n * 4) != hipSuccess ||
        hipMalloc((void **)&s, 8 * (size_t)nz1) != hipSuccess) {
        rc = hip_fail(hipErrorOutOfMemory, "arec_h_cc_images");
    } else {
        hipMemcpyAsync(a, images1, n * 4, hipMemcpyHostToDevice, st);
        hipMemcpyAsync(b, images2, n * 4, hipMemcpyHostToDevice, st);
        rc = cc_views(a, b, nx1, ny1, nz1, nullptr, s, s + nz1, st);
        if (!rc) {
            hipMemcpyAsync(sx, s, 4 * (size_t)nz1, hipMemcpyDeviceToHost, st);
            hipMemcpyAsync(sy, s + nz1, 4 * (size_t)nz1, hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) rc = hip_fail(hipGetLastError(), "arec_h_cc_images");
        }
    }
    hipFree(a);
    hipFree(b);
    hipFree(s);
    hipStreamDestroy(st);
    return rc;
}

int arec_h_shift_images(int32_t nx, int32_t ny, int32_t nz, float *data, const float *sx,
                        const float *sy) {
    if (!data || !sx || !sy || nx < 1 || ny < 1 || nz < 1) {
        set_error("arec_h_shift_images: bad argument");
        return AREC_EINVAL;
    }
    hipStream_t st;
    AREC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    size_t n = (size_t)nx * ny * nz;
    float *a = nullptr, *s = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&a, n * 4) != hipSuccess || hipMalloc((void **)&s, 8 * (size_t)nz) != hipSuccess) {
        rc = hip_fail(hipErrorOutOfMemory, "arec_h_shift_images");
    } else {
        hipMemcpyAsync(a, data, n * 4, hipMemcpyHostToDevice, st);
        hipMemcpyAsync(s, sx, 4 * (size_t)nz, hipMemcpyHostToDevice, st);
        hipMemcpyAsync(s + nz, sy, 4 * (size_t)nz, hipMemcpyHostToDevice, st);
        rc = shift_views(a, nx, ny, nz, s, s + nz, -1.0f, st);
        if (!rc) {
            hipMemcpyAsync(data, a, n * 4, hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) rc = hip_fail(hipGetLastError(), "arec_h_shift_images");
        }
    }
    hipFree(a);
    hipFree(s);
    hipStreamDestroy(st);
    return rc;
}

// Host-pointer versions of the reference's per-view alignment helpers
// (align2dstack.h:9,13,24-25, the drop-in's ccorr2d, fshift2d, speak and
// quad2dfit): upload, compute on the GPU with the stack kernels above (one
// view), download.  quad2dfit is 9 values in, 3 out: the same FP64
// arithmetic as the speak kernel's fit, on the host.
int arec_quad2dfit(const float *z, float *xsh, float *ysh, float *peakv) {
    if (!z || !xsh || !ysh || !peakv) {
        set_error("arec_quad2dfit: bad argument");
        return AREC_EINVAL;
    }
    quad2dfit_dev(z, xsh, ysh, peakv);
    return 0;
}

int arec_h_ccorr2d(int32_t nx, int32_t ny, const float *x, const float *y, float *c) {
    if (!x || !y || !c || nx < 1 || ny < 1) {
        set_error("arec_h_ccorr2d: bad argument");
        return AREC_EINVAL;
    }
    const int64_t n = (int64_t)nx * ny;
    return one_view(x, y, n, c, n, "arec_h_ccorr2d",
                    [&](float *a, float *b, float *cm, hipStream_t st, const float **res) {
                        *res = cm;
                        return cc_views(a, b, nx, ny, 1, cm, nullptr, nullptr, st);
                    });
}

int arec_h_speak(int32_t nx, int32_t ny, const float *data, float *xt, float *yt) {
    if (!data || !xt || !yt || nx < 1 || ny < 1) {
        set_error("arec_h_speak: bad argument");
        return AREC_EINVAL;
    }
    float s[2];
    int rc = one_view(data, nullptr, (int64_t)nx * ny, s, 2, "arec_h_speak",
                      [&](float *a, float *, float *o, hipStream_t st, const float **res) {
                          speak_kernel<<<1, 256, 0, st>>>(a, nx, ny, o, o + 1);
                          *res = o;
                          AREC_HIP(hipGetLastError());
                          return 0;
                      });
    if (!rc) {
        *xt = s[0];
        *yt = s[1];
    }
    return rc;
}

int arec_h_fshift2d(int32_t nx, int32_t ny, const float *imagein, float *imageout, float sx, float sy) {
    if (!imagein || !imageout || nx < 1 || ny < 1) {
        set_error("arec_h_fshift2d: bad argument");
        return AREC_EINVAL;
    }
    const int64_t n = (int64_t)nx * ny;
    return one_view(imagein, nullptr, n, imageout, n, "arec_h_fshift2d",
                    [&](float *a, float *, float *o, hipStream_t st, const float **res) {
                        const float shift[2] = {sx, sy};
                        AREC_HIP(hipMemcpyAsync(o, shift, sizeof shift, hipMemcpyHostToDevice, st));
                        AREC_HIP(hipStreamSynchronize(st));  // shift is a stack array
                        *res = a;
                        return shift_views(a, nx, ny, 1, o, o + 1, 1.0f, st);
                    });
}

// Host-pointer versions of the preprocessing (the drop-in's
// arecImageTakeLog and periodicDecomposition): upload, compute, download.
static int host_views_op(float *data, int64_t n, const char *what, int (*op)(float *, hipStream_t, const void *),
                         const void *arg) {
    hipStream_t st;
    AREC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    float *a = nullptr;
    int rc = 0;
    if (hipMalloc((void **)&a, sizeof(float) * std::max<int64_t>(1, n)) != hipSuccess) {
        rc = hip_fail(hipErrorOutOfMemory, what);
    } else {
        hipMemcpyAsync(a, data, sizeof(float) * n, hipMemcpyHostToDevice, st);
        rc = op(a, st, arg);
        if (!rc) {
            hipMemcpyAsync(data, a, sizeof(float) * n, hipMemcpyDeviceToHost, st);
            if (hipStreamSynchronize(st) != hipSuccess) rc = hip_fail(hipGetLastError(), what);
        }
    }
    hipFree(a);
    hipStreamDestroy(st);
    return rc;
}

int arec_h_take_log(float *data, int64_t n) {
    if ((!data && n > 0) || n < 0) {
        set_error("arec_h_take_log: bad argument");
        return AREC_EINVAL;
    }
    if (n == 0) return 0;
    return host_views_op(data, n, "arec_h_take_log",
                         [](float *a, hipStream_t st, const void *arg) {
                             return take_log(a, *static_cast<const int64_t *>(arg), st);
                         },
                         &n);
}

int arec_h_periodic_decomposition(int32_t nx, int32_t ny, int32_t nz, float *data) {
    if (!data || nx < 1 || ny < 1 || nz < 1) {
        set_error("arec_h_periodic_decomposition: bad argument");
        return AREC_EINVAL;
    }
    const int32_t dims[3] = {nx, ny, nz};
    return host_views_op(data, (int64_t)nx * ny * nz, "arec_h_periodic_decomposition",
                         [](float *a, hipStream_t st, const void *arg) {
                             const int32_t *d = static_cast<const int32_t *>(arg);
                             return periodic_views(a, d[0], d[1], d[2], st);
                         },
                         dims);
}

}  // extern "C"
