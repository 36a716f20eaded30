// Parallel-beam projector pair over the compact cylinder, gfx950.
//
// Geometry (AREC3D/AREC3D/arecproject.cpp): every row y of the volume is an
// independent 2-D problem with the same weights, so one operator application
// is a sparse matrix with 3 taps per (angle, voxel column) applied to the ny
// columns of the volume.  Both kernels call tap_weights<>() below, so the
// forward and back projections use bit-identical weights and are exact
// adjoints up to FP32 summation order.
//
//  * Forward (ray-driven gather, no atomics): one wavefront owns one angle and
//    a chunk of 64*RPL detector rows (lanes = rows).  It walks that angle's
//    voxel columns in order of their first detector tap (the forward table),
//    reads each column's rows as one coalesced 256-byte segment and keeps the
//    three detector pixels the column can touch in registers.  A pixel is
//    complete once the walk has passed it and is stored then; every output
//    pixel is written exactly once.
//  * Back (voxel-driven gather, no atomics): one workgroup owns a tile of
//    256 voxel columns inside a 16x16 box times 16 rows, one column per
//    thread.  Per block of 4 angles it stages the tile's detector footprint
//    (<= 32 pixels x 16 rows per angle) in LDS, reads each column's taps from
//    the geometry's tap table (or computes them), and accumulates the 16 rows
//    in registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <thread>

#include "arec_internal.hpp"

namespace arec {

__constant__ float c_kb_lut[1000] = {
#include "kb_lut_f.inc"
};

struct Taps {
    int p0;
    float w0, w1, w2;
};

// piece_wise_integrated (arecproject.h:27-53), the CDF of the unit-area
// trapezoid; the plateau formula applies on |x| < a + 1e-3.
__device__ __forceinline__ double trap_cdf(double x, double a, double b, double ym) {
#pragma clang fp contract(off)
    if (x < -b) return 0.0;
    if (x > b) return 1.0;
    double a_eps = a + 1e-3;
    if (x > -a_eps && x < a_eps) return 0.5 * ym * (b + a) + ym * x;
    if (x < -a) return 0.5 * ym * (x + b) * (x + b) / (b - a);
    if (x > a) return 0.5 * ym * (b - x) * (b - x) / (a - b) + 1.0;
    return 0.0;
}

// interp1_R_MOD (arecproject.cpp:340-346).
__device__ __forceinline__ float kb_interp(float x) {
#pragma clang fp contract(off)
    const float DR = (float)(3.0 / 999);
    x = fabsf(x);
    x = x / DR;
    int ind = (int)x;
    x = x - (float)ind;
    float second = c_kb_lut[ind + 1] * x;
    return (float)((double)c_kb_lut[ind] * (1.0 - (double)x) + (double)second);
}

// Detector taps of the column at coordinates (X, Z) for one angle.
//  SQ  (arecproject.cpp:603-620): xb = (X-r)c + (Z-r)s + r in FP64,
//      iqx = round-half-even(xb), taps iqx-1..iqx+1 weighted by differences
//      of the trapezoid CDF.
//  KB  (:396-403): same xb and iqx, LUT weights at (float)(xb - tap).
//  linear (:73-80): xb rounded to FP32 around the integer centre (int)r+1,
//      taps floor(xb)-1 and floor(xb) (0-based), third weight 0.
template <int M>
__device__ __forceinline__ Taps tap_weights(int X, int Z, const AngleConst &A, double radius,
                                            int xcent) {
#pragma clang fp contract(off)
    Taps t;
    if (M == kLinear) {
        double t1 = (double)(X - xcent) * A.c;
        double t2 = (double)(Z - xcent) * A.s;
        float xb = (float)((t1 + t2) + (double)xcent);
        int iqx = (int)floorf(xb);
        float dipx = xb - (float)iqx;
        t.p0 = iqx - 1;
        t.w0 = (float)(1.0 - (double)dipx);
        t.w1 = dipx;
        t.w2 = 0.0f;
        return t;
    }
    double t1 = ((double)X - radius) * A.c;
    double t2 = ((double)Z - radius) * A.s;
    double xb = (t1 + t2) + radius;
    int iqx = __double2int_rn(xb);
    t.p0 = iqx - 1;
    if (M == kKB) {
        t.w0 = kb_interp((float)(xb - (double)(iqx - 1)));
        t.w1 = kb_interp((float)(xb - (double)iqx));
        t.w2 = kb_interp((float)(xb - (double)(iqx + 1)));
        return t;
    }
    // |xv| <= 1/2 and the footprint half-width b <= 0.7072, so the CDF at
    // xv - 1.5 is exactly 0 and at xv + 1.5 exactly 1: w0 = v1 - 0 and
    // w2 = 1 - v2 are the reference's differences bit for bit.
    double xv = (double)iqx - xb;
    double v1 = trap_cdf(xv - 0.5, A.a, A.b, A.lmax);
    double v2 = trap_cdf(xv + 0.5, A.a, A.b, A.lmax);
    t.w0 = (float)v1;
    t.w1 = (float)(v2 - v1);
    t.w2 = (float)(1.0 - v2);
    return t;
}

// ---------------------------------------------------------------------------
// Forward table: entries of angle a in order of first tap.
// ---------------------------------------------------------------------------
template <int M>
__global__ void build_fwd_table(const int *__restrict__ order, int a0, int na_batch, int nvalid,
                                const int *__restrict__ coord, const AngleConst *__restrict__ ang,
                                double radius, int xcent, uint4 *__restrict__ table,
                                int *__restrict__ err) {
    int ab = blockIdx.y;
    if (ab >= na_batch) return;
    int a = a0 + ab;
    AngleConst A = ang[a];
    const int *ord = order + (size_t)ab * nvalid;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nvalid; i += gridDim.x * blockDim.x) {
        int j = ord[i];
        Taps t = tap_weights<M>(coord[2 * j], coord[2 * j + 1], A, radius, xcent);
        int prev = kTapBase;
        if (i > 0) {
            int jp = ord[i - 1];
            prev = tap_weights<M>(coord[2 * jp], coord[2 * jp + 1], A, radius, xcent).p0;
        }
        int step = t.p0 - prev;
        if (step < 0 || step > kMaxTapStep) {
            atomicOr(err, 1);
            step = 0;
        }
        uint4 e;
        e.x = (unsigned)j | ((unsigned)step << kColBits);
        e.y = __float_as_uint(t.w0);
        e.z = __float_as_uint(t.w1);
        e.w = __float_as_uint(t.w2);
        table[(size_t)a * nvalid + i] = e;
    }
}

// Back projector's tap table: entry ((tile * na) + a) * kTileCols + slot holds
// the taps of the tile's slot-th column at angle a (zeros for an empty slot),
// computed by the same tap_weights<> as everywhere else.
template <int M>
__global__ void build_back_table(const int *__restrict__ tile_cols, int ntiles, int na,
                                 const int *__restrict__ coord, const AngleConst *__restrict__ ang,
                                 double radius, int xcent, uint4 *__restrict__ btab) {
    const int64_t n = (int64_t)ntiles * na * kTileCols;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int slot = (int)(e % kTileCols);
        const int64_t ta = e / kTileCols;
        const int a = (int)(ta % na), tile = (int)(ta / na);
        const int j = tile_cols[(int64_t)tile * kTileCols + slot];
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (j >= 0) {
            Taps t = tap_weights<M>(coord[2 * j], coord[2 * j + 1], ang[a], radius, xcent);
            v = make_uint4((unsigned)t.p0, __float_as_uint(t.w0), __float_as_uint(t.w1), __float_as_uint(t.w2));
        }
        btab[e] = v;
    }
}

// Depth-strip table (design 2) for angles [a0, a0 + nab): entry e of angle a
// (order[e] = column | strip << 22) -> {column | d << 22, w0, w1, w2} with
// d = first tap - (strip * kStripW - 2), which must lie in [0, kStripW + 1].
template <int M>
__global__ void build_strip_table(const uint32_t *__restrict__ order, const int64_t *__restrict__ base, int a0,
                                  int nab, const int *__restrict__ coord, const AngleConst *__restrict__ ang,
                                  double radius, int xcent, uint4 *__restrict__ tab, int *__restrict__ err) {
    const int ab = blockIdx.y;
    if (ab >= nab) return;
    const int a = a0 + ab;
    const int64_t lo = base[a] - base[a0], n = base[a + 1] - base[a];
    const AngleConst A = ang[a];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t o = order[lo + i];
        const int j = (int)(o & (kMaxCols - 1)), sidx = (int)(o >> kColBits);
        Taps t = tap_weights<M>(coord[2 * j], coord[2 * j + 1], A, radius, xcent);
        int d = t.p0 - (sidx * kStripW - 2);
        if (d < 0 || d > kStripW + 1) {
            atomicOr(err, 8);
            d = 0;
            t.w0 = t.w1 = t.w2 = 0.0f;
        }
        tab[base[a] + i] = make_uint4((unsigned)j | ((unsigned)d << kColBits), __float_as_uint(t.w0),
                                      __float_as_uint(t.w1), __float_as_uint(t.w2));
    }
}

// Angle-pair table (design 3) for pairs [u0, u0 + nub): entry e of pair u
// (order[e] = column | strip << 22 | present flags << 30, bit 30 angle 2u,
// bit 31 angle 2u + 1) -> {column | d0 << 22 | d1 << 27, w0[3]}, {w1[3], 0},
// d = first tap - (strip * kStripW - 2) in [0, kStripW + 1] for a present
// angle, kStripGroups (and zero weights) for an absent one.
template <int M>
__global__ void build_pair_table(const uint32_t *__restrict__ order, const int64_t *__restrict__ base, int u0,
                                 int nub, int na, const int *__restrict__ coord, const AngleConst *__restrict__ ang,
                                 double radius, int xcent, uint4 *__restrict__ tab, int *__restrict__ err) {
    const int ub = blockIdx.y;
    if (ub >= nub) return;
    const int u = u0 + ub;
    const int64_t lo = base[u] - base[u0], n = base[u + 1] - base[u];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t o = order[lo + i];
        const int j = (int)(o & (kMaxCols - 1)), sidx = (int)((o >> kColBits) & 255);
        unsigned d[2];
        Taps t[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            t[h] = Taps{0, 0.f, 0.f, 0.f};
            d[h] = kStripGroups;
            const int a = 2 * u + h;
            if ((o >> (30 + h)) & 1u) {
                if (a >= na) {
                    atomicOr(err, 8);
                    continue;
                }
                t[h] = tap_weights<M>(coord[2 * j], coord[2 * j + 1], ang[a], radius, xcent);
                const int dd = t[h].p0 - (sidx * kStripW - 2);
                if (dd < 0 || dd > kStripW + 1) {
                    atomicOr(err, 8);
                    t[h] = Taps{0, 0.f, 0.f, 0.f};
                } else {
                    d[h] = (unsigned)dd;
                }
            }
        }
        uint4 *e = tab + 2 * (base[u] + i);
        e[0] = make_uint4((unsigned)j | (d[0] << kColBits) | (d[1] << (kColBits + 5)), __float_as_uint(t[0].w0),
                          __float_as_uint(t[0].w1), __float_as_uint(t[0].w2));
        e[1] = make_uint4(__float_as_uint(t[1].w0), __float_as_uint(t[1].w1), __float_as_uint(t[1].w2), 0u);
    }
}

// Staged-strip table (design 4) for n entries: entry e (column col[e], angle
// ang[e], meta[e] = slot | d << 8 | strip << 16) -> {slot | d << 8, w0, w1,
// w2}; the device's first tap must give the host's d.  A padding entry
// (column 0xffffffff) gets zero weights.
template <int M>
__global__ void build_staged_table(const uint32_t *__restrict__ col, const uint32_t *__restrict__ angi,
                                   const uint32_t *__restrict__ meta, int64_t n, const int *__restrict__ coord,
                                   const AngleConst *__restrict__ ang, double radius, int xcent, int rows4,
                                   uint4 *__restrict__ tab, int *__restrict__ err) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t m = meta[e];
        const int d = (int)((m >> 8) & 31u), st = (int)(m >> 16);
        Taps t = Taps{0, 0.0f, 0.0f, 0.0f};  // padding entry (column 0xffffffff): zero weights
        if (col[e] != 0xffffffffu) {
            const int j = (int)col[e], a = (int)angi[e];
            t = tap_weights<M>(coord[2 * j], coord[2 * j + 1], ang[a], radius, xcent);
            if (t.p0 - (st * kStripW - 2) != d) {
                atomicOr(err, 16);
                t.w0 = t.w1 = t.w2 = 0.0f;
            }
        }
        tab[e] = make_uint4((m & 0xffu) * (uint32_t)rows4 | (uint32_t)d << kSgOffShift, __float_as_uint(t.w0),
                            __float_as_uint(t.w1), __float_as_uint(t.w2));
    }
}

// Re-encode the staged table's entry words for another column stride (rows
// per lane): slot * from4 + (d << kSgOffShift) -> slot * to4 + (d << ...).
__global__ void restride_staged_table(uint4 *__restrict__ tab, int64_t n, int from4, int to4) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t x = tab[e].x, lo = x & ((1u << kSgOffShift) - 1u);
        tab[e].x = lo / (uint32_t)from4 * (uint32_t)to4 | (x & ~((1u << kSgOffShift) - 1u));
    }
}

// ---------------------------------------------------------------------------
// Forward projector
// ---------------------------------------------------------------------------
constexpr int kFwdBatch = 8;  // table entries whose loads are in flight together

// One wavefront: angle a, rows y0 + lane*RPL + r (r < RPL), detector pixels
// [u_lo, u_hi).  It walks the angle's table entries whose first tap lies in
// [u_lo - 2, u_hi) in tap order, keeping the three pixels base..base+2 in
// registers.  A pixel is complete when the walk moves past it; completed
// pixels go to an LDS strip of kFwdOut columns which is written to global
// memory as 32-byte row segments.  partials (optional) receives the FP64 sum
// of the FP32 squares of the wave's stored pixels.
// Workgroup b of one forward application: angle group b % nag (fastest),
// then detector segment, then row chunk.  Every wave of the workgroup calls
// this with the same b.
// The buffers are __restrict__ parameters (not struct members) so that the
// compiler may keep the batched column loads ahead of the strip stores.
template <int RPL>
__device__ __forceinline__ void forward_body(const float *__restrict__ x, float *__restrict__ p,
                                             const uint4 *__restrict__ table, const int4 *__restrict__ seg,
                                             const int *__restrict__ pixstart, double *__restrict__ partials,
                                             const int npix, const int nvalid, const int nseg, const int segw,
                                             const int nx, const int ny, const int na, const int nchunks,
                                             int b) {
    constexpr int ROWS = 64 * RPL;
    constexpr int LDW = kFwdOut + 1;  // padded strip row (conflict-free column writes)
    __shared__ float s_out[kFwdAngles][ROWS * LDW];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nag = (na + kFwdAngles - 1) / kFwdAngles;
    const int ag = b % nag;
    const int rest = b / nag;
    const int sg = rest % nseg;
    const int chunk = rest / nseg;
    const int a_raw = ag * kFwdAngles + wave;
    const bool active = a_raw < na;  // inactive waves still take part in the barriers
    const int a = active ? a_raw : na - 1;
    const int u_lo = sg * segw, u_hi = min(nx, u_lo + segw);
    const int ybase = chunk * ROWS;
    const int y0 = ybase + lane * RPL;
    float *__restrict__ pa = p + (size_t)a * nx * ny;
    const uint4 *__restrict__ tab = table + (size_t)a * nvalid;
    const int4 sd = seg[(size_t)a * nseg + sg];
    const int i0 = __builtin_amdgcn_readfirstlane(sd.x), i1 = __builtin_amdgcn_readfirstlane(sd.y);
    float *strip = s_out[wave];

    // the window's three pixels (values, never addressed: a pointer to them
    // would place them in scratch memory)
    struct Rows {
        float v[RPL];
    };
    Rows acc0, acc1, acc2;
#pragma unroll
    for (int r = 0; r < RPL; r++) acc0.v[r] = acc1.v[r] = acc2.v[r] = 0.0f;
    double sq = 0.0;
    int base = __builtin_amdgcn_readfirstlane(sd.z);  // window: pixels base..base+2
    int strip_lo = u_lo;                                 // strip holds [strip_lo, +kFwdOut)
    const bool full_rows = (ybase + ROWS <= ny);

    // write the strip [strip_lo, strip_lo + kFwdOut) of the wave's rows
    auto flush_strip = [&]() __attribute__((always_inline)) {
        __threadfence_block();
        for (int e = lane; e < ROWS * kFwdOut; e += 64) {
            int r = e / kFwdOut, c = e % kFwdOut;
            int y = ybase + r, q = strip_lo + c;
            if (y < ny && q < u_hi) pa[(size_t)y * nx + q] = strip[r * LDW + c];
        }
        __threadfence_block();
        strip_lo += kFwdOut;
    };
    // pixel q of the wave's rows is complete with values v
    auto complete = [&](int q, const Rows v) __attribute__((always_inline)) {
        if (q < u_lo || q >= u_hi) return;
        while (q >= strip_lo + kFwdOut) flush_strip();
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            strip[(lane * RPL + r) * LDW + (q - strip_lo)] = v.v[r];
            if (y0 + r < ny) {
                float t = v.v[r] * v.v[r];
                sq += (double)t;
            }
        }
    };
    Rows zero;
#pragma unroll
    for (int r = 0; r < RPL; r++) zero.v[r] = 0.0f;
    // advance the window so that it starts at pixel nb (nb >= base)
    auto advance = [&](int nb) __attribute__((always_inline)) {
        const int step = nb - base;
        if (step == 0) return;
        complete(base, acc0);
        if (step >= 2) complete(base + 1, acc1);
        if (step >= 3) {
            complete(base + 2, acc2);
            for (int q = max(base + 3, u_lo); q < min(nb, u_hi); q++) complete(q, zero);
        }
        // the window's new pixels, every assignment unconditional: stores
        // made in some branches only are merged by the compiler into one
        // store through a selected pointer, which puts the window in scratch
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const float n0 = step == 1 ? acc1.v[r] : step == 2 ? acc2.v[r] : 0.0f;
            const float n1 = step == 1 ? acc2.v[r] : 0.0f;
            acc0.v[r] = n0;
            acc1.v[r] = n1;
            acc2.v[r] = 0.0f;
        }
        base = nb;
    };

    // The workgroup's waves (neighbouring angles) sweep the segment in
    // lockstep: phase k covers the entries whose first tap is below
    // u_lo - 2 + k*kFwdPhase, then all waves meet at a barrier, so that they
    // read nearly the same columns at nearly the same time.
    const int *__restrict__ ps = pixstart + (size_t)a * npix;
    const int nphase = (u_hi - (u_lo - 2) + kFwdPhase - 1) / kFwdPhase;
    int i = i0;
    for (int ph = 1; ph <= nphase; ph++) {
        int thr = u_lo - 2 + ph * kFwdPhase;
        int iend = ph == nphase ? i1 : min(i1, ps[min(npix - 1, max(0, thr - kTapBase))]);
        iend = __builtin_amdgcn_readfirstlane(active ? iend : i);
        for (; i < iend; i += kFwdBatch) {
        uint4 e[kFwdBatch];
#pragma unroll
        for (int k = 0; k < kFwdBatch; k++) {
            if (i + k < iend) {
                e[k] = tab[i + k];
            } else {
                e[k] = make_uint4(0u, 0u, 0u, 0u);  // no column, no step, zero weights
            }
        }
        float v[kFwdBatch][RPL];
#pragma unroll
        for (int k = 0; k < kFwdBatch; k++) {
            const int j = (int)(e[k].x & (kMaxCols - 1));
            const float *__restrict__ xj = x + (size_t)j * ny + y0;
            if (RPL == 4 && full_rows) {
                float4 t = *reinterpret_cast<const float4 *>(xj);
                v[k][0] = t.x;
                v[k][RPL > 1 ? 1 : 0] = t.y;
                v[k][RPL > 2 ? 2 : 0] = t.z;
                v[k][RPL > 3 ? 3 : 0] = t.w;
            } else if (RPL == 2 && full_rows) {
                float2 t = *reinterpret_cast<const float2 *>(xj);
                v[k][0] = t.x;
                v[k][RPL > 1 ? 1 : 0] = t.y;
            } else {
#pragma unroll
                for (int r = 0; r < RPL; r++) v[k][r] = (y0 + r < ny) ? xj[r] : 0.0f;
            }
        }
#pragma unroll
        for (int k = 0; k < kFwdBatch; k++) {
            const int step = (int)(e[k].x >> kColBits);
            if (step) advance(base + step);
            const float w0 = __uint_as_float(e[k].y), w1 = __uint_as_float(e[k].z),
                        w2 = __uint_as_float(e[k].w);
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                acc0.v[r] = fmaf(w0, v[k][r], acc0.v[r]);
                acc1.v[r] = fmaf(w1, v[k][r], acc1.v[r]);
                acc2.v[r] = fmaf(w2, v[k][r], acc2.v[r]);
            }
        }
        }
        i = iend;
        __syncthreads();
    }
    if (!active) return;
    // the last window and the pixels above it up to the segment end
    advance(max(base + 3, u_hi));
    while (strip_lo < u_hi) flush_strip();
    if (partials) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
        if (lane == 0) partials[((size_t)a * nchunks + chunk) * nseg + sg] = sq;
    }
}

// Plain kernel arguments (not a struct): the compiler then keeps the sizes in
// scalar registers (68 VGPRs instead of 89 at four rows per lane).
template <int RPL>
__global__ __launch_bounds__(64 * kFwdAngles) void forward_kernel(
    const float *__restrict__ x, float *__restrict__ p, const uint4 *__restrict__ table,
    const int4 *__restrict__ seg, const int *__restrict__ pixstart, double *__restrict__ partials, int npix,
    int nvalid, int nseg, int segw, int nx, int ny, int na, int nchunks, const int *__restrict__ skip) {
    if (skip && *skip) return;
    // Workgroups with consecutive ids run on different XCDs: give each XCD a
    // contiguous range of (chunk, segment, angle group) triples, angle group
    // fastest, so that the workgroups resident on one XCD sweep the same rows
    // of nearly the same columns (neighbouring angles walk the disk almost
    // together) and share its L2.
    const int nblk = gridDim.x;
    int b = blockIdx.x;
    if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
    forward_body<RPL>(x, p, table, seg, pixstart, partials, npix, nvalid, nseg, segw, nx, ny, na, nchunks, b);
}

// acc[D..D+2] += w * v for one entry of the depth-strip walk (D is the
// entry's first tap relative to the strip start - 2, a constant per group).
template <int D, int RPL>
__device__ __forceinline__ void strip_add(float (&acc)[kStripW + 4][RPL], float w0, float w1, float w2,
                                          const float (&v)[RPL]) {
#pragma unroll
    for (int r = 0; r < RPL; r++) {
        acc[D][r] = fmaf(w0, v[r], acc[D][r]);
        acc[D + 1][r] = fmaf(w1, v[r], acc[D + 1][r]);
        acc[D + 2][r] = fmaf(w2, v[r], acc[D + 2][r]);
    }
}

// The wave's strip table entries [s0, last] stream through a ring of 128
// entries in LDS (slot = entry index mod 128), kStripAhead entries ahead of
// the walk.  Every batch issues the same memory operations in the same
// order: it writes the previous batch's fetched entries to the ring, loads
// its columns' rows, then fetches (lanes 0-7, one entry each) the 8 entries
// kStripAhead past its own first entry.  The fetch is the youngest load when
// the batch waits for its rows, so it stays in flight; it is written one
// batch later, and read about kStripAhead entries later still.  A write
// replaces entries at least 128 - kStripAhead - 8 before the batch, which no
// batch reads again.  With batches of kStripBatch < 8 entries consecutive
// fetches overlap, and an entry is written (with the same value) more than
// once before it is read.
constexpr int kStripAhead = 64;

struct StripRing {
    uint4 *ring;  // this wave's 128 entries in LDS
    const uint4 *__restrict__ t;
    int s0, last;
    int fetched;  // first entry of the fetch in flight
    uint4 pend;   // that fetch (lane l: entry fetched + l, l < 8)
};

__device__ __forceinline__ uint4 strip_entry_load(const StripRing &R, int e) { return R.t[min(e, R.last)]; }

// acc[D..D+2] += w * v for the batch's entries whose first tap is the
// strip start - 2 + D (weights of the other entries are zero).
template <int D, int RPL>
__device__ __forceinline__ void strip_add_batch(float (&acc)[kStripW + 4][RPL], const unsigned (&dk)[kStripBatch],
                                                const float (&w)[kStripBatch][3],
                                                const float (&v)[kStripBatch][RPL]) {
#pragma unroll
    for (int k = 0; k < kStripBatch; k++) {
        const bool on = dk[k] == (unsigned)D;
        strip_add<D, RPL>(acc, on ? w[k][0] : 0.0f, on ? w[k][1] : 0.0f, on ? w[k][2] : 0.0f, v[k]);
    }
}

// strip_add_batch for every offset D in [LO, HI) whose bit is set in
// present, by halving the range: a batch's one or two offsets cost about
// 2 log2(kStripGroups) scalar tests instead of one per offset.
template <int LO, int HI, int RPL>
__device__ __forceinline__ void strip_add_present(float (&acc)[kStripW + 4][RPL], const unsigned present,
                                                  const unsigned (&dk)[kStripBatch],
                                                  const float (&w)[kStripBatch][3],
                                                  const float (&v)[kStripBatch][RPL]) {
    if constexpr (HI - LO == 1) {
        strip_add_batch<LO, RPL>(acc, dk, w, v);
    } else {
        constexpr int MID = (LO + HI) / 2;
        constexpr unsigned ML = ((1u << MID) - 1u) & ~((1u << LO) - 1u);
        constexpr unsigned MH = ((1u << HI) - 1u) & ~((1u << MID) - 1u);
        if (present & ML) strip_add_present<LO, MID, RPL>(acc, present, dk, w, v);
        if (present & MH) strip_add_present<MID, HI, RPL>(acc, present, dk, w, v);
    }
}

// One batch: entries i .. i + kStripBatch - 1 of the strip (indices clamped
// to last so every load is valid; entries past last add zero).  Every entry
// is accumulated for each first-tap offset present in the batch, with the
// weights of the entries at other offsets zeroed, so the batch needs no
// alignment to the offsets (entries are ordered by offset within a bin, so
// a batch usually spans one or two).  FULL: all of the wave's rows lie
// inside the volume.
template <int RPL, bool FULL>
__device__ __forceinline__ void strip_batch(StripRing &R, const int i, const float *__restrict__ x, const int ny,
                                            const int y0, const int lane, float (&acc)[kStripW + 4][RPL]) {
    static_assert(kStripBatch <= 8 && kStripAhead >= 2 * kStripBatch && kStripAhead + 8 <= 128,
                  "strip ring sizes");
    if (lane < 8) R.ring[(R.fetched + lane) & 127] = R.pend;
    // The batch's entries, one 16-byte LDS read each.  Slots past last hold
    // entry last (every fetch clamps its index), so no clamp is needed here;
    // such entries add zero below.
    unsigned ew[kStripBatch][4];
#pragma unroll
    for (int k = 0; k < kStripBatch; k++) {
        const uint4 e = R.ring[(i + k) & 127];
        ew[k][0] = __builtin_amdgcn_readfirstlane(e.x);
        ew[k][1] = __builtin_amdgcn_readfirstlane(e.y);
        ew[k][2] = __builtin_amdgcn_readfirstlane(e.z);
        ew[k][3] = __builtin_amdgcn_readfirstlane(e.w);
    }
    float v[kStripBatch][RPL];
#pragma unroll
    for (int k = 0; k < kStripBatch; k++) {
        const int col = (int)(ew[k][0] & (kMaxCols - 1));
        const float *__restrict__ xj = x + (size_t)col * ny + y0;
        if constexpr (FULL && RPL == 4) {
            const float4 q = *reinterpret_cast<const float4 *>(xj);
            v[k][0] = q.x;
            v[k][RPL > 1 ? 1 : 0] = q.y;
            v[k][RPL > 2 ? 2 : 0] = q.z;
            v[k][RPL > 3 ? 3 : 0] = q.w;
        } else if constexpr (FULL && RPL == 2) {
            const float2 q = *reinterpret_cast<const float2 *>(xj);
            v[k][0] = q.x;
            v[k][RPL > 1 ? 1 : 0] = q.y;
        } else if constexpr (FULL) {
            v[k][0] = xj[0];
        } else {
#pragma unroll
            for (int r = 0; r < RPL; r++) v[k][r] = (y0 + r < ny) ? xj[r] : 0.0f;
        }
    }
    R.fetched = i + kStripAhead;
    R.pend = strip_entry_load(R, R.fetched + (lane & 7));
    unsigned dk[kStripBatch], present = 0;
    float w[kStripBatch][3];
#pragma unroll
    for (int k = 0; k < kStripBatch; k++) {
        dk[k] = i + k <= R.last ? ew[k][0] >> kColBits : (unsigned)kStripGroups;
        present |= 1u << dk[k];
#pragma unroll
        for (int q = 0; q < 3; q++) w[k][q] = __uint_as_float(ew[k][1 + q]);
    }
    strip_add_present<0, kStripGroups, RPL>(acc, present, dk, w, v);
}

// The wave's walk over its strip's entries [o[0], o[nbins]) in batches, with
// a workgroup barrier whenever the walk has passed the end of a depth bin
// (every wave meets nbins barriers; the waves share no data, only the
// lockstep).  The ring starts with entries [s0, s0 + kStripAhead) and the
// fetch of the 8 after them.
template <int RPL, bool FULL>
__device__ __forceinline__ void strip_walk(uint4 *ring, const float *__restrict__ x, const uint4 *__restrict__ t,
                                           const int *__restrict__ o, const int nbins, const bool active,
                                           const int ny, const int y0, const int lane,
                                           float (&acc)[kStripW + 4][RPL]) {
    StripRing R;
    R.ring = ring;
    R.t = t;
    R.s0 = __builtin_amdgcn_readfirstlane(o[0]);
    R.last = __builtin_amdgcn_readfirstlane(o[nbins]) - 1;
    const bool any = active && R.last >= R.s0;
    if (any) {
        static_assert(kStripAhead <= 64, "one entry per lane at the start");
        if (lane < kStripAhead) R.ring[(R.s0 + lane) & 127] = strip_entry_load(R, R.s0 + lane);
        R.fetched = R.s0 + kStripAhead;
        R.pend = strip_entry_load(R, R.fetched + (lane & 7));
    }
    int i = R.s0;
    for (int bin = 0; bin < nbins; bin++) {
        if (any) {
            const int bend = __builtin_amdgcn_readfirstlane(o[bin + 1]);
            for (; i < bend; i += kStripBatch) strip_batch<RPL, FULL>(R, i, x, ny, y0, lane, acc);
        }
        __builtin_amdgcn_s_barrier();
    }
}

// Pixels st*kStripW + q of angle a (rows y0 .. y0 + RPL - 1) from acc[q + 2];
// the FP64 sum of their FP32 squares into partials (optional).
template <int RPL>
__device__ __forceinline__ void strip_store(const float (&acc)[kStripW + 4][RPL], float *__restrict__ p,
                                            double *__restrict__ partials, const int a, const int st,
                                            const int chunk, const int nchunks, const int nstrips, const int nx,
                                            const int ny, const int y0, const int lane) {
    double sq = 0.0;
    const int u0 = st * kStripW;
    float *__restrict__ pa = p + (size_t)a * nx * ny;
    // a whole strip with 16-byte aligned detector rows: four 16-byte stores
    // per row (the same values and FP64 sum order as the pixel loop)
    const bool vec = u0 + kStripW <= nx && (nx & 3) == 0;
#pragma unroll
    for (int r = 0; r < RPL; r++) {
        const int y = y0 + r;
        if (y >= ny) continue;
        if (vec) {
            float4 *row = reinterpret_cast<float4 *>(pa + (size_t)y * nx + u0);
#pragma unroll
            for (int q4 = 0; q4 < kStripW / 4; q4++) {
                const float v0 = acc[4 * q4 + 2][r], v1 = acc[4 * q4 + 3][r], v2 = acc[4 * q4 + 4][r],
                            v3 = acc[4 * q4 + 5][r];
                row[q4] = make_float4(v0, v1, v2, v3);
                const float s0 = v0 * v0, s1 = v1 * v1, s2 = v2 * v2, s3 = v3 * v3;
                sq += (double)s0;
                sq += (double)s1;
                sq += (double)s2;
                sq += (double)s3;
            }
            continue;
        }
#pragma unroll
        for (int q = 0; q < kStripW; q++) {
            if (u0 + q < nx) {
                const float val = acc[q + 2][r];
                pa[(size_t)y * nx + u0 + q] = val;
                const float s2 = val * val;
                sq += (double)s2;
            }
        }
    }
    if (partials) {
#pragma unroll
        for (int off2 = 32; off2 > 0; off2 >>= 1) sq += __shfl_xor(sq, off2, 64);
        if (lane == 0) partials[((size_t)a * nchunks + chunk) * nstrips + st] = sq;
    }
}

// Forward projection, depth-strip design.  Workgroup b: angle group b % nag
// (kStripAngles waves, one angle each), then detector strip, then row
// chunk.  Each wave walks its angle's entries of the strip bin by bin in
// depth along the ray; the waves meet at a barrier after every bin, so
// neighbouring angles read nearly the same columns at nearly the same time.
// The strip's pixels stay in registers (an entry's first-tap offset selects
// a fixed register set); every pixel of the strip is written once at the
// end.
template <int RPL>
__device__ __forceinline__ void forward_strip_body(const float *__restrict__ x, float *__restrict__ p,
                                                   const uint4 *__restrict__ tab,
                                                   const int64_t *__restrict__ base,
                                                   const int *__restrict__ off, double *__restrict__ partials,
                                                   const int nstrips, const int nbins, const int nx, const int ny,
                                                   const int na, const int nchunks, int b) {
    constexpr int ROWS = 64 * RPL;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nag = (na + kStripAngles - 1) / kStripAngles;
    const int ag = b % nag;
    const int rest = b / nag;
    const int st = rest % nstrips;
    const int chunk = rest / nstrips;
    const int a_raw = ag * kStripAngles + wave;
    const bool active = a_raw < na;
    const int a = active ? a_raw : na - 1;
    const int ybase = chunk * ROWS;
    const int y0 = ybase + lane * RPL;
    const bool full_rows = (ybase + ROWS <= ny);
    const uint4 *__restrict__ t = tab + base[a];
    const int *__restrict__ o = off + (size_t)a * ((size_t)nstrips * nbins + 1) + (size_t)st * nbins;
    float acc[kStripW + 4][RPL];
#pragma unroll
    for (int i = 0; i < kStripW + 4; i++)
#pragma unroll
        for (int r = 0; r < RPL; r++) acc[i][r] = 0.0f;
    __shared__ uint4 s_ring[kStripAngles][2 * 64];
    if (full_rows)
        strip_walk<RPL, true>(s_ring[wave], x, t, o, nbins, active, ny, y0, lane, acc);
    else
        strip_walk<RPL, false>(s_ring[wave], x, t, o, nbins, active, ny, y0, lane, acc);
    if (!active) return;
    strip_store<RPL>(acc, p, partials, a, st, chunk, nchunks, nstrips, nx, ny, y0, lane);
}

// At least 5 waves per SIMD (93 VGPRs at 4 rows per lane instead of 108 at 4
// waves; 0.8% faster).
template <int RPL>
__global__ __launch_bounds__(64 * kStripAngles) __attribute__((amdgpu_waves_per_eu(5, 8))) void forward_strip_kernel(
    const float *__restrict__ x, float *__restrict__ p, const uint4 *__restrict__ tab,
    const int64_t *__restrict__ base, const int *__restrict__ off, double *__restrict__ partials, int nstrips,
    int nbins, int nx, int ny, int na, int nchunks, const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int nblk = gridDim.x;
    int b = blockIdx.x;
    if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
    forward_strip_body<RPL>(x, p, tab, base, off, partials, nstrips, nbins, nx, ny, na, nchunks, b);
}

// Forward projection, depth strips with angle pairs (design 3): as design 2,
// but each wave walks one pair of angles through the strip with one table
// whose entries carry both angles' offsets and weights, so each column's
// rows are loaded once for the two angles (the union of the two angles'
// columns is about 1.1 times either).  The ring holds 128 two-uint4
// entries; each batch's fetch of 8 entries uses lanes 0-15, one uint4 each.
struct PairRing {
    uint4 *ring;  // this wave's 128 entries (256 uint4) in LDS
    const uint4 *__restrict__ t;
    int s0, last;
    int fetched;  // first entry of the fetch in flight
    uint4 pend;   // lane l < 16: half l & 1 of entry fetched + l / 2
};

__device__ __forceinline__ uint4 pair_half_load(const PairRing &R, int e, int h) {
    return R.t[2 * (int64_t)min(e, R.last) + h];
}

// Word q (0-7) of entry idx (s0 <= idx <= last, in the ring), wave-uniform.
__device__ __forceinline__ unsigned pair_ring_word(const PairRing &R, const int idx, const int q) {
    const unsigned *r = reinterpret_cast<const unsigned *>(R.ring);
    return __builtin_amdgcn_readfirstlane(r[8 * (idx & 127) + q]);
}

template <int RPL, bool FULL>
__device__ __forceinline__ void pair_batch(PairRing &R, const int i, const float *__restrict__ x, const int ny,
                                           const int y0, const int lane, float (&acc0)[kStripW + 4][RPL],
                                           float (&acc1)[kStripW + 4][RPL]) {
    if (lane < 16) R.ring[2 * ((R.fetched + (lane >> 1)) & 127) + (lane & 1)] = R.pend;
    float v[kStripBatch][RPL];
#pragma unroll
    for (int k = 0; k < kStripBatch; k++) {
        const int col = (int)(pair_ring_word(R, min(i + k, R.last), 0) & (kMaxCols - 1));
        const float *__restrict__ xj = x + (size_t)col * ny + y0;
        if constexpr (FULL && RPL == 2) {
            const float2 q = *reinterpret_cast<const float2 *>(xj);
            v[k][0] = q.x;
            v[k][RPL > 1 ? 1 : 0] = q.y;
        } else if constexpr (FULL && RPL == 1) {
            v[k][0] = xj[0];
        } else {
#pragma unroll
            for (int r = 0; r < RPL; r++) v[k][r] = (y0 + r < ny) ? xj[r] : 0.0f;
        }
    }
    R.fetched = i + kStripAhead;
    R.pend = pair_half_load(R, R.fetched + ((lane & 15) >> 1), lane & 1);
    unsigned d0[kStripBatch], d1[kStripBatch], p0 = 0, p1 = 0;
    float w0[kStripBatch][3], w1[kStripBatch][3];
#pragma unroll
    for (int k = 0; k < kStripBatch; k++) {
        const int e = min(i + k, R.last);
        const unsigned word = pair_ring_word(R, e, 0);
        const bool in = i + k <= R.last;
        d0[k] = in ? (word >> kColBits) & 31u : (unsigned)kStripGroups;
        d1[k] = in ? word >> (kColBits + 5) : (unsigned)kStripGroups;
        p0 |= 1u << d0[k];
        p1 |= 1u << d1[k];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            w0[k][q] = __uint_as_float(pair_ring_word(R, e, 1 + q));
            w1[k][q] = __uint_as_float(pair_ring_word(R, e, 4 + q));
        }
    }
    strip_add_present<0, kStripGroups, RPL>(acc0, p0, d0, w0, v);
    strip_add_present<0, kStripGroups, RPL>(acc1, p1, d1, w1, v);
}

template <int RPL, bool FULL>
__device__ __forceinline__ void pair_walk(uint4 *ring, const float *__restrict__ x, const uint4 *__restrict__ t,
                                          const int *__restrict__ o, const int nbins, const bool active,
                                          const int ny, const int y0, const int lane,
                                          float (&acc0)[kStripW + 4][RPL], float (&acc1)[kStripW + 4][RPL]) {
    PairRing R;
    R.ring = ring;
    R.t = t;
    R.s0 = __builtin_amdgcn_readfirstlane(o[0]);
    R.last = __builtin_amdgcn_readfirstlane(o[nbins]) - 1;
    const bool any = active && R.last >= R.s0;
    if (any) {
        static_assert(kStripAhead <= 64, "one entry per lane at the start");
        if (lane < kStripAhead) {
            const int slot = 2 * ((R.s0 + lane) & 127);
            R.ring[slot] = pair_half_load(R, R.s0 + lane, 0);
            R.ring[slot + 1] = pair_half_load(R, R.s0 + lane, 1);
        }
        R.fetched = R.s0 + kStripAhead;
        R.pend = pair_half_load(R, R.fetched + ((lane & 15) >> 1), lane & 1);
    }
    int i = R.s0;
    for (int bin = 0; bin < nbins; bin++) {
        if (any) {
            const int bend = __builtin_amdgcn_readfirstlane(o[bin + 1]);
            for (; i < bend; i += kStripBatch) pair_batch<RPL, FULL>(R, i, x, ny, y0, lane, acc0, acc1);
        }
        __builtin_amdgcn_s_barrier();
    }
}

// Workgroup b: group of kStripAngles pairs b % ngroups (one per wave), then
// detector strip, then row chunk.
template <int RPL>
__device__ __forceinline__ void forward_pair_body(const float *__restrict__ x, float *__restrict__ p,
                                                  const uint4 *__restrict__ tab, const int64_t *__restrict__ base,
                                                  const int *__restrict__ off, double *__restrict__ partials,
                                                  const int nstrips, const int nbins, const int nx, const int ny,
                                                  const int na, const int nchunks, int b) {
    constexpr int ROWS = 64 * RPL;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int nu = (na + 1) / 2;
    const int ngroups = (nu + kStripAngles - 1) / kStripAngles;
    const int grp = b % ngroups;
    const int rest = b / ngroups;
    const int st = rest % nstrips;
    const int chunk = rest / nstrips;
    const int u_raw = grp * kStripAngles + wave;
    const bool active = u_raw < nu;
    const int u = active ? u_raw : nu - 1;
    const int ybase = chunk * ROWS;
    const int y0 = ybase + lane * RPL;
    const bool full_rows = (ybase + ROWS <= ny);
    const uint4 *__restrict__ t = tab + 2 * base[u];
    const int *__restrict__ o = off + (size_t)u * ((size_t)nstrips * nbins + 1) + (size_t)st * nbins;
    float acc0[kStripW + 4][RPL], acc1[kStripW + 4][RPL];
#pragma unroll
    for (int i = 0; i < kStripW + 4; i++)
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            acc0[i][r] = 0.0f;
            acc1[i][r] = 0.0f;
        }
    __shared__ uint4 s_ring[kStripAngles][4 * 64];
    if (full_rows)
        pair_walk<RPL, true>(s_ring[wave], x, t, o, nbins, active, ny, y0, lane, acc0, acc1);
    else
        pair_walk<RPL, false>(s_ring[wave], x, t, o, nbins, active, ny, y0, lane, acc0, acc1);
    if (!active) return;
    strip_store<RPL>(acc0, p, partials, 2 * u, st, chunk, nchunks, nstrips, nx, ny, y0, lane);
    if (2 * u + 1 < na) strip_store<RPL>(acc1, p, partials, 2 * u + 1, st, chunk, nchunks, nstrips, nx, ny, y0, lane);
}

template <int RPL>
__global__ __launch_bounds__(64 * kStripAngles) void forward_pair_kernel(
    const float *__restrict__ x, float *__restrict__ p, const uint4 *__restrict__ tab,
    const int64_t *__restrict__ base, const int *__restrict__ off, double *__restrict__ partials, int nstrips,
    int nbins, int nx, int ny, int na, int nchunks, const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int nblk = gridDim.x;
    int b = blockIdx.x;
    if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
    forward_pair_body<RPL>(x, p, tab, base, off, partials, nstrips, nbins, nx, ny, na, nchunks, b);
}

// Forward projection, staged strips (design 4).  Workgroup b: angle group
// G = b % ngroups (kSgAngles waves, one angle each), then strip s, then row
// chunk.  The group's stage list of (G, s) is walked chunk by chunk: the
// kSgChunk columns x 64*RPL rows of the next chunk are loaded (each wave
// one column in kSgAngles, one coalesced 64*RPL*4-byte segment) while the
// current chunk is read from LDS by every wave's entries, then stored to
// the other LDS buffer; one barrier per chunk (a buffer a wave writes in
// iteration k was last read before the barrier that ended iteration k - 1).
// Each entry adds its column's rows, read from LDS, to the strip's pixel
// registers as in design 2 (batches of kStripBatch entries, every offset
// present in the batch accumulated with the others' weights zeroed); the
// entries of a (chunk, angle) are ordered by offset, and a chunk's columns
// lie along one band of the central angle's rays, so a batch mostly holds
// one offset.  Table entries are wave-uniform (scalar loads).
constexpr int kSgThreads = 64 * kSgAngles;

// acc[D..D+2] += w * v for every entry of a batch, for the one offset D whose
// bit is set in present (the batch's real entries all start at D; entries
// staged past the range end have zero weights, so none needs a mask).
// Sequential bit tests, as strip_add_present: an if/else on D would let the
// compiler merge the leaves' stores into one store through a selected
// pointer, which puts acc in scratch memory.
template <int LO, int HI, int RPL>
__device__ __forceinline__ void strip_add_single(float (&acc)[kStripW + 4][RPL], const unsigned present,
                                                 const float (&w)[kStripBatch][3],
                                                 const float (&v)[kStripBatch][RPL]) {
    if constexpr (HI - LO == 1) {
#pragma unroll
        for (int k = 0; k < kStripBatch; k++) strip_add<LO, RPL>(acc, w[k][0], w[k][1], w[k][2], v[k]);
    } else {
        constexpr int MID = (LO + HI) / 2;
        constexpr unsigned ML = ((1u << MID) - 1u) & ~((1u << LO) - 1u);
        constexpr unsigned MH = ((1u << HI) - 1u) & ~((1u << MID) - 1u);
        if (present & ML) strip_add_single<LO, MID, RPL>(acc, present, w, v);
        if (present & MH) strip_add_single<MID, HI, RPL>(acc, present, w, v);
    }
}

// Lane l gets lane (l & ~3) + k's value (quad_perm [k,k,k,k]); the compiler
// folds it into the consuming add as a DPP operand.
// A wave-uniform pointer to global memory, read back from the first lane
// into scalar registers and typed as global: an address built from it plus
// a lane offset uses the scalar-base form of the load, and a pointer read
// from a struct in memory (generic to the compiler, so accessed with flat
// instructions, which also count against the LDS wait counter) is accessed
// with global ones.  For uniform pointers only.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *scalar_ptr(T *p) {
    const uint64_t u = (uint64_t)p;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return (__attribute__((address_space(1))) T *)(((uint64_t)hi << 32) | lo);
}
template <typename T>
__device__ __forceinline__ T *as_global(T *p) {
    return (T *)scalar_ptr(p);
}

template <int K>
__device__ __forceinline__ unsigned quad_bcast_u32(const unsigned v) {
    static_assert(K >= 0 && K < 4, "quad lane");
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, K * 0x55, 0xf, 0xf, false);
}

// a_t[r] += w_t(entry k) * v[k][r] for the batch's four entries k and the
// taps t (3, or the linear pair's 2): entry k's weights are lane (l & ~3) +
// k's, read through a quad_perm DPP operand of v_fmac_f32 (the compiler keeps
// such a broadcast as a separate v_mov_b32_dpp).  Entry by entry, so every
// accumulator sums in entry order.  One asm statement per batch (two at 4
// rows per lane, for the operand count), each opening with the two wait
// states a DPP operand needs after a VALU write (the weights come from an
// LDS read; s_nop 1 covers any copy); inside a statement no DPP operand is
// written.
#define AREC_QP(K) " quad_perm:[" #K "," #K "," #K "," #K "] row_mask:0xf bank_mask:0xf"
#define AREC_TAP(T, K, R) "v_fmac_f32_dpp %[a" #T #R "], %[w" #T "], %[v" #K #R "]" AREC_QP(K) "\n\t"
#define AREC_E3(K, R) AREC_TAP(0, K, R) AREC_TAP(1, K, R) AREC_TAP(2, K, R)
#define AREC_E2(K, R) AREC_TAP(0, K, R) AREC_TAP(1, K, R)
#define AREC_R1(E, K) E(K, 0)
#define AREC_R2(E, K) E(K, 0) E(K, 1)
#define AREC_R4(E, K) E(K, 0) E(K, 1) E(K, 2) E(K, 3)
#define AREC_A3(R) [a0##R] "+v"(a0[R]), [a1##R] "+v"(a1[R]), [a2##R] "+v"(a2[R])
#define AREC_A2(R) [a0##R] "+v"(a0[R]), [a1##R] "+v"(a1[R])
#define AREC_W3 [w0] "v"(w0), [w1] "v"(w1), [w2] "v"(w2)
#define AREC_W2 [w0] "v"(w0), [w1] "v"(w1)
#define AREC_V(K, R) [v##K##R] "v"(v[K][R])
#define AREC_V2(K) AREC_V(K, 0), AREC_V(K, 1)
#define AREC_V4(K) AREC_V(K, 0), AREC_V(K, 1), AREC_V(K, 2), AREC_V(K, 3)
template <int RPL, int TAPS>
__device__ __forceinline__ void fma_batch(float (&a0)[RPL], float (&a1)[RPL], float (&a2)[RPL], const float w0,
                                          const float w1, const float w2, const float (&v)[kStripBatch][RPL]) {
    static_assert(kStripBatch == 4 && (RPL == 1 || RPL == 2 || RPL == 4) && (TAPS == 2 || TAPS == 3),
                  "quad entries, rows per lane, taps");
    if constexpr (TAPS == 3 && RPL == 1) {
        asm volatile("s_nop 1\n\t" AREC_R1(AREC_E3, 0) AREC_R1(AREC_E3, 1) AREC_R1(AREC_E3, 2) AREC_R1(AREC_E3, 3)
                     : AREC_A3(0)
                     : AREC_W3, AREC_V(0, 0), AREC_V(1, 0), AREC_V(2, 0), AREC_V(3, 0));
    } else if constexpr (TAPS == 3 && RPL == 2) {
        asm volatile("s_nop 1\n\t" AREC_R2(AREC_E3, 0) AREC_R2(AREC_E3, 1) AREC_R2(AREC_E3, 2) AREC_R2(AREC_E3, 3)
                     : AREC_A3(0), AREC_A3(1)
                     : AREC_W3, AREC_V2(0), AREC_V2(1), AREC_V2(2), AREC_V2(3));
    } else if constexpr (TAPS == 3) {
        asm volatile("s_nop 1\n\t" AREC_R4(AREC_E3, 0) AREC_R4(AREC_E3, 1)
                     : AREC_A3(0), AREC_A3(1), AREC_A3(2), AREC_A3(3)
                     : AREC_W3, AREC_V4(0), AREC_V4(1));
        asm volatile("s_nop 1\n\t" AREC_R4(AREC_E3, 2) AREC_R4(AREC_E3, 3)
                     : AREC_A3(0), AREC_A3(1), AREC_A3(2), AREC_A3(3)
                     : AREC_W3, AREC_V4(2), AREC_V4(3));
    } else if constexpr (RPL == 1) {
        asm volatile("s_nop 1\n\t" AREC_R1(AREC_E2, 0) AREC_R1(AREC_E2, 1) AREC_R1(AREC_E2, 2) AREC_R1(AREC_E2, 3)
                     : AREC_A2(0)
                     : AREC_W2, AREC_V(0, 0), AREC_V(1, 0), AREC_V(2, 0), AREC_V(3, 0));
    } else if constexpr (RPL == 2) {
        asm volatile("s_nop 1\n\t" AREC_R2(AREC_E2, 0) AREC_R2(AREC_E2, 1) AREC_R2(AREC_E2, 2) AREC_R2(AREC_E2, 3)
                     : AREC_A2(0), AREC_A2(1)
                     : AREC_W2, AREC_V2(0), AREC_V2(1), AREC_V2(2), AREC_V2(3));
    } else {
        asm volatile("s_nop 1\n\t" AREC_R4(AREC_E2, 0) AREC_R4(AREC_E2, 1) AREC_R4(AREC_E2, 2) AREC_R4(AREC_E2, 3)
                     : AREC_A2(0), AREC_A2(1), AREC_A2(2), AREC_A2(3)
                     : AREC_W2, AREC_V4(0), AREC_V4(1), AREC_V4(2), AREC_V4(3));
    }
}
#undef AREC_V4
#undef AREC_V2
#undef AREC_V
#undef AREC_W2
#undef AREC_W3
#undef AREC_A2
#undef AREC_A3
#undef AREC_R4
#undef AREC_R2
#undef AREC_R1
#undef AREC_E2
#undef AREC_E3
#undef AREC_TAP
#undef AREC_QP

// acc[D..D+2] += w(entry k) * v[k] for the batch's four entries (lane group
// l & ~3 holds them), for the one offset D whose bit is set in present: a
// binary search over [LO, HI), one path through it (present has exactly one
// bit, inside the range).
template <int LO, int HI, int RPL, int TAPS>
__device__ __forceinline__ void strip_add_quad(float (&acc)[kStripW + 4][RPL], const unsigned present,
                                               const float w0, const float w1, const float w2,
                                               const float (&v)[kStripBatch][RPL]) {
    if constexpr (HI - LO == 1) {
        fma_batch<RPL, TAPS>(acc[LO], acc[LO + 1], acc[LO + 2], w0, w1, w2, v);
    } else {
        constexpr int MID = (LO + HI) / 2;
        constexpr unsigned ML = ((1u << MID) - 1u) & ~((1u << LO) - 1u);
        constexpr unsigned MH = ((1u << HI) - 1u) & ~((1u << MID) - 1u);
        if (present & ML) strip_add_quad<LO, MID, RPL, TAPS>(acc, present, w0, w1, w2, v);
        if (present & MH) strip_add_quad<MID, HI, RPL, TAPS>(acc, present, w0, w1, w2, v);
    }
}

// The walk over chunks [K0, K1) of one (group, strip) for the wave's rows;
// FULL: all 64*RPL rows lie inside the volume and the row count is a
// multiple of RPL (so every lane's RPL rows of a column are aligned for one
// vector load); otherwise each row is loaded on its own, bounds-checked.
// Per chunk, staged into LDS
// one chunk ahead: the kSgChunk columns' rows (an empty slot, stage -1,
// stages column 0: no entry reads it) and each wave's entries of the chunk
// (at most kSgEnt, whole batches of one offset each; lane l fetches entry
// l straight into LDS).  With whole chunks at 4 rows per lane a column goes
// from global memory straight to its LDS slot too (one
// global_load_lds_dwordx4 per column: 64 lanes x 16 bytes), else through
// registers (at 2 rows per lane one 8-byte load per lane measured faster
// than two direct 4-byte loads: 15.9 vs 17.0 ms at 512^3/720).  A batch reads its four
// entries from LDS (the same address in every lane), takes its offset from
// the first, and adds each entry's column rows with the entry's weights.
template <int RPL, bool FULL, int TAPS>
__device__ __forceinline__ void staged_walk(float (*__restrict__ col0)[64 * RPL], float (*__restrict__ col1)[64 * RPL],
                                            uint4 *__restrict__ ent0, uint4 *__restrict__ ent1,
                                            const float *__restrict__ x, const uint4 *__restrict__ tab,
                                            const int *__restrict__ stage, const int64_t *__restrict__ off,
                                            const int64_t K0, const int64_t K1, const int wave, const int lane,
                                            const int ny, const int y0, float (&acc)[kStripW + 4][RPL]) {
    constexpr int kPer = kSgChunk / kSgAngles;  // columns each wave stages per chunk
    static_assert((kPer & (kPer - 1)) == 0 && kPer <= 64, "slot lanes");
    static_assert(kPer * kSgAngles == kSgChunk, "chunk staging pattern");
    static_assert(kSgEnt <= 64 && kSgEnt % kStripBatch == 0, "one entry per lane, whole batches");
    constexpr bool kDma = FULL && RPL == 4;
    float stg[kDma ? 1 : kPer][RPL];
    int nent_next;  // its entry count (a multiple of kStripBatch)
    // Chunk K's stage-list slots and entry bounds: lane c < kPer holds the
    // index of the column this wave stages into slot wave + c * kSgAngles,
    // lanes 0 and 1 the entry range.  Below 4 rows per lane they are fetched
    // one chunk before the columns, so that they have landed when the
    // columns and entries are requested (no dependent round trip); at 4 rows
    // per lane the registers to hold them ahead are not there (they spill,
    // and a spill reload waits for the direct loads in flight), so they are
    // fetched just before.
    constexpr bool kAhead = RPL < 4;
    int jv = 0;
    int64_t bv = 0;
    auto fetch_idx = [&](const int64_t K) __attribute__((always_inline)) {
        // scalar bases plus small lane offsets: no 64-bit lane pointers
        // kept across the chunk loop (at 4 rows per lane they spill)
        const auto sb = scalar_ptr(stage + (K * kSgChunk + wave));
        const auto ob = scalar_ptr(off + (K * kSgAngles + wave));
        jv = sb[(unsigned)(lane & (kPer - 1)) * kSgAngles];
        bv = ob[(unsigned)(lane & 1)];
    };
    auto load_stage = [&](float (*__restrict__ cb)[64 * RPL], uint4 *__restrict__ eb) __attribute__((always_inline)) {
        int64_t e0 = 0;
        auto entry_range = [&]() __attribute__((always_inline)) {
            e0 = (int64_t)(unsigned)__builtin_amdgcn_readlane((int)bv, 0) |
                 ((int64_t)__builtin_amdgcn_readlane((int)(bv >> 32), 0) << 32);
            nent_next = __builtin_amdgcn_readlane((int)bv, 1) - (int)e0;
        };
        // fetched just before (4 rows per lane): the entry range first, so
        // that waiting for it waits for no column load
        if constexpr (!kAhead) entry_range();
#pragma unroll
        for (int c = 0; c < kPer; c++) {
            const int j = max(0, __builtin_amdgcn_readlane(jv, c));
            // the column's scalar base plus the lane's row offset
            const float *__restrict__ xj = (const float *)scalar_ptr(x + (size_t)j * ny) + (unsigned)y0;
            if constexpr (kDma) {
                __builtin_amdgcn_global_load_lds(xj, &cb[wave + c * kSgAngles][0], 16, 0, 0);
            } else if constexpr (FULL && RPL == 2) {
                const float2 q = *reinterpret_cast<const float2 *>(xj);
                stg[c][0] = q.x;
                stg[c][RPL > 1 ? 1 : 0] = q.y;
            } else if constexpr (FULL) {
                stg[c][0] = xj[0];
            } else {
#pragma unroll
                for (int r = 0; r < RPL; r++) stg[c][r] = y0 + r < ny ? xj[r] : 0.0f;
            }
        }
        if constexpr (kAhead) entry_range();
        if (lane < nent_next) __builtin_amdgcn_global_load_lds((void *)(tab + e0 + lane), eb, 16, 0, 0);
    };
    auto store_stage = [&](float (*__restrict__ cb)[64 * RPL]) __attribute__((always_inline)) {
        if constexpr (!kDma) {
#pragma unroll
            for (int c = 0; c < kPer; c++) {
                float *d = &cb[wave + c * kSgAngles][lane * RPL];
                if constexpr (RPL == 4) {
                    *reinterpret_cast<float4 *>(d) = make_float4(stg[c][0], stg[c][RPL > 1 ? 1 : 0],
                                                                 stg[c][RPL > 2 ? 2 : 0], stg[c][RPL > 3 ? 3 : 0]);
                } else if constexpr (RPL == 2) {
                    *reinterpret_cast<float2 *>(d) = make_float2(stg[c][0], stg[c][RPL > 1 ? 1 : 0]);
                } else {
                    d[0] = stg[c][0];
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): the direct loads land before the barrier
    };

    // chunk K from buffers (cb, eb) while chunk K + 1 goes to (nb, ne).  The
    // two buffers are separate restrict pointers and alternate statically
    // (the chunk loop is unrolled by two), so the compiler can tell that the
    // reads of one do not wait for the direct loads into the other.
    int nent = 0;
    auto step = [&](const int64_t K, const float (*__restrict__ cb)[64 * RPL], const uint4 *__restrict__ eb,
                    float (*__restrict__ nb)[64 * RPL], uint4 *__restrict__ ne) __attribute__((always_inline)) {
        const bool more = K + 1 < K1;
        if (more) {
            if (!kAhead) fetch_idx(K + 1);
            load_stage(nb, ne);
            if (kAhead && K + 2 < K1) fetch_idx(K + 2);
        }
        // lane l's byte address of its rows in slot 0 of this buffer
        const char *lb = reinterpret_cast<const char *>(&cb[0][lane * RPL]);
        for (int i = 0; i < nent; i += kStripBatch) {
            // lane l holds entry i + l % 4; entry k reaches every lane of a
            // quad through a quad_perm [k,k,k,k] operand
            const uint4 q = eb[i + (lane & (kStripBatch - 1))];
            const unsigned xd = (unsigned)__builtin_amdgcn_readfirstlane((int)q.x);
            const unsigned d = (xd >> kSgOffShift) & 31u;
            const char *lbd = lb - (d << kSgOffShift);  // entry words hold slot offset + (d << kSgOffShift)
            float v[kStripBatch][RPL];
            auto col_rows = [&](const float *src, float(&vk)[RPL]) __attribute__((always_inline)) {
                if constexpr (RPL == 4) {
                    const float4 c = *reinterpret_cast<const float4 *>(src);
                    vk[0] = c.x;
                    vk[RPL > 1 ? 1 : 0] = c.y;
                    vk[RPL > 2 ? 2 : 0] = c.z;
                    vk[RPL > 3 ? 3 : 0] = c.w;
                } else if constexpr (RPL == 2) {
                    const float2 c = *reinterpret_cast<const float2 *>(src);
                    vk[0] = c.x;
                    vk[RPL > 1 ? 1 : 0] = c.y;
                } else {
                    vk[0] = src[0];
                }
            };
            static_assert(kStripBatch == 4, "one entry per lane of a quad");
            col_rows(reinterpret_cast<const float *>(lbd + quad_bcast_u32<0>(q.x)), v[0]);
            col_rows(reinterpret_cast<const float *>(lbd + quad_bcast_u32<1>(q.x)), v[1]);
            col_rows(reinterpret_cast<const float *>(lbd + quad_bcast_u32<2>(q.x)), v[2]);
            col_rows(reinterpret_cast<const float *>(lbd + quad_bcast_u32<3>(q.x)), v[3]);
            unsigned present = 1u << d;
            // opaque to the optimiser: knowing it is one bit, it turns the
            // bit tests into a switch whose leaves copy every accumulator
            // (at 1 row per lane past 128 VGPRs, into scratch)
            asm volatile("" : "+s"(present));
            strip_add_quad<0, kStripGroups, RPL, TAPS>(acc, present, __uint_as_float(q.y), __uint_as_float(q.z),
                                                 __uint_as_float(q.w), v);
        }
        if (more) {
            store_stage(nb);
            nent = nent_next;
        }
        __syncthreads();
    };

    if (K0 < K1) {
        fetch_idx(K0);
        load_stage(col0, ent0);
        if (kAhead && K0 + 1 < K1) fetch_idx(K0 + 1);
        store_stage(col0);
        nent = nent_next;
    }
    __syncthreads();
    for (int64_t K = K0; K < K1; K += 2) {
        step(K, col0, ent0, col1, ent1);
        if (K + 1 < K1) step(K + 1, col1, ent1, col0, ent0);
    }
}

// Workgroup b of one staged-strip forward: angle group b % ngroups (one
// angle per wave), then detector strip, then row chunk.
template <int RPL, int TAPS>
__device__ __forceinline__ void forward_staged_body(const float *__restrict__ x, float *__restrict__ p,
                                                    const uint4 *__restrict__ tab, const int *__restrict__ stage,
                                                    const int64_t *__restrict__ chunk_base,
                                                    const int64_t *__restrict__ off, double *__restrict__ partials,
                                                    const int nstrips, const int nx, const int ny, const int na,
                                                    const int nchunks, const int b) {
    constexpr int ROWS = 64 * RPL;
    __shared__ __align__(16) float s_col[2][kSgChunk][ROWS];
    __shared__ uint4 s_ent[2][kSgAngles][kSgEnt];
    const int ngroups = (na + kSgAngles - 1) / kSgAngles;
    const int G = b % ngroups, rest = b / ngroups, st = rest % nstrips, chunk = rest / nstrips;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int a = G * kSgAngles + wave;
    const bool active = a < na;
    const int ybase = chunk * ROWS, y0 = ybase + lane * RPL;
    const int64_t K0 = chunk_base[(int64_t)G * nstrips + st], K1 = chunk_base[(int64_t)G * nstrips + st + 1];
    float acc[kStripW + 4][RPL];
#pragma unroll
    for (int i = 0; i < kStripW + 4; i++)
#pragma unroll
        for (int r = 0; r < RPL; r++) acc[i][r] = 0.0f;
    // whole chunk with every column's rows aligned to RPL floats: vector loads
    if (ybase + ROWS <= ny && (ny & (RPL - 1)) == 0)
        staged_walk<RPL, true, TAPS>(s_col[0], s_col[1], s_ent[0][wave], s_ent[1][wave], x, tab, stage, off, K0, K1, wave,
                               lane, ny, y0, acc);
    else
        staged_walk<RPL, false, TAPS>(s_col[0], s_col[1], s_ent[0][wave], s_ent[1][wave], x, tab, stage, off, K0, K1, wave,
                                lane, ny, y0, acc);
    if (!active) return;
    strip_store<RPL>(acc, p, partials, a, st, chunk, nchunks, nstrips, nx, ny, y0, lane);
}

// At least 4 waves per SIMD, so that two workgroups share a CU.  With a
// multiple of 8 workgroups each XCD takes a contiguous range of them
// (neighbouring angle groups, which read the same columns).
template <int RPL, int TAPS>
__global__ __launch_bounds__(kSgThreads) __attribute__((amdgpu_waves_per_eu(RPL == 2 ? 6 : 4, 8))) void forward_staged_kernel(
    const float *__restrict__ x, float *__restrict__ p, const uint4 *__restrict__ tab, const int *__restrict__ stage,
    const int64_t *__restrict__ chunk_base, const int64_t *__restrict__ off, double *__restrict__ partials,
    int nstrips, int nx, int ny, int na, int nchunks, const int *__restrict__ skip) {
    if (skip && *skip) return;
    const int nblk = gridDim.x;
    int b = blockIdx.x;
    if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
    forward_staged_body<RPL, TAPS>(x, p, tab, stage, chunk_base, off, partials, nstrips, nx, ny, na, nchunks, b);
}

// The staged-strip forward of many problems in one launch (one grid, no
// loop: a persistent loop kept the body's invariants in registers and
// halved occupancy).  Problem k owns workgroups [8 blk8, 8 blk8 + 8 m) for a
// padded count 8 m >= nblk; workgroups go round-robin over the 8 XCDs, so
// local index l of the problem sits on XCD l & 7 and is remapped as in the
// single launch: each XCD takes a contiguous eighth of every problem, and
// all XCDs move through the problems together.  Padding workgroups return.
template <int RPL, int TAPS>
__global__ __launch_bounds__(kSgThreads) __attribute__((amdgpu_waves_per_eu(RPL == 2 ? 6 : 4, 8))) void forward_staged_batch_kernel(
    const BatchFwd *__restrict__ probs, const int *__restrict__ owner) {
    const int k = owner[blockIdx.x >> 3];
    const BatchFwd q = probs[k];
    if (q.skip && *as_global(q.skip)) return;
    const int l = (int)((int64_t)blockIdx.x - 8 * q.blk8);
    const int m = ((q.nblk + 7) >> 3);  // workgroups per XCD
    const int b = (l & 7) * m + (l >> 3);
    if (b >= q.nblk) return;
    forward_staged_body<RPL, TAPS>(as_global(q.x), as_global(q.p), as_global(q.tab), as_global(q.stage),
                             as_global(q.chunk_base), as_global(q.off), nullptr, q.nstrips, q.nx, q.ny, q.na,
                             q.nchunks, b);
}

// Index of the problem that owns work item `it`: the last k with
// item0(k) <= it (item0 ascending).  Uniform across the workgroup.  The
// item0 values of the first kBatchLdsProbs problems are staged in LDS once
// per workgroup (stage_item0), so the search costs no global round trips
// for batches of that size.
constexpr int kBatchLdsProbs = 256;

template <typename P>
__device__ __forceinline__ void stage_item0(const P *__restrict__ probs, int nprob, int *s_item0) {
    for (int i = threadIdx.x; i < min(nprob, kBatchLdsProbs); i += blockDim.x) s_item0[i] = probs[i].item0;
    __syncthreads();
}

template <typename P>
__device__ __forceinline__ int owner_of(const P *__restrict__ probs, const int *s_item0, int nprob, int it) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        const int i0 = mid < kBatchLdsProbs ? s_item0[mid] : probs[mid].item0;
        if (i0 <= it) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Back projector
// ---------------------------------------------------------------------------
constexpr int kBpThreads = kTileCols;  // one voxel column per thread
// Rows per workgroup and angles per LDS pass, measured at 512^3/720 with the
// pipelined loop below: 32 rows x 2 angles 12.5 ms (128 VGPRs), 32 x 4
// 12.9 ms, 16 x 4 14.3 ms, 16 x 8 14.9 ms, 64 x 2 14.3 ms, 24 x 4 15.8 ms,
// 8 x 4 17.3 ms; the unpipelined loop took 20.8 ms at 16 x 4.
// The shape is a template parameter (kBpShapeRows / kBpShapeAngles, chosen
// per geometry); the LDS row stride is the row count + 4 (16-byte aligned).
constexpr int kBpWin = 32;  // staged pixels per (angle)

template <int M>
__device__ __forceinline__ double corner_xb(int X, int Z, const AngleConst &A, double radius,
                                            int xcent) {
#pragma clang fp contract(off)
    if (M == kLinear) return ((double)(X - xcent) * A.c + (double)(Z - xcent) * A.s) + (double)xcent;
    return (((double)X - radius) * A.c + ((double)Z - radius) * A.s) + radius;
}

// x[j][y] = sum_a sum_k w_k(a,j) p[a][y][p0+k] for the tile's columns (one
// per thread) and ROWS rows.  Per block of ANGB angles the tile's footprint
// (<= 32 pixels x ROWS rows per angle) is staged in LDS with rows fastest, so one
// 16-byte LDS read gives 4 rows of one tap.  Epilogue by mode:
//  kBpStore      out = acc
//  kBpSirtGrad   out = bvec - acc (FP32), partial += sum out*(double)out
//  kBpStoreNorm  out = acc, partial += sum (double)(acc*acc)
template <int M, int MODE, int ROWS, int ANGB>
__device__ __forceinline__ void backward_body(const float *__restrict__ p, float *__restrict__ out,
                                              const float *__restrict__ bvec, const uint4 *__restrict__ btab,
                                              const int *__restrict__ tile_cols,
                                              const int *__restrict__ tile_org, const int *__restrict__ coord,
                                              const AngleConst *__restrict__ ang, double *__restrict__ partials,
                                              int *__restrict__ err, const double radius, const int xcent,
                                              const int nx, const int ny, const int na, const int nrowblk, int b) {
    constexpr int kBpRows = ROWS, kBpAngBlk = ANGB, kBpRowPad = ROWS + 4;
    // Software pipeline over blocks of kBpAngBlk angles, one barrier per
    // block: while block k is accumulated from s_p[k & 1], the footprints of
    // block k + 1 are in flight to registers (then stored to s_p[(k+1) & 1])
    // and the windows of block k + 2 are computed into s_u0[(k+2) % 3].  A
    // buffer written in iteration k was last read before the barrier that
    // ended iteration k - 1.
    __shared__ __align__(16) float s_p[2][kBpAngBlk][kBpWin][kBpRowPad];
    __shared__ int s_u0[3][kBpAngBlk];
    __shared__ double s_red[kBpThreads / 64];
    constexpr int kStage = kBpRows * kBpWin / kBpThreads;  // staged values per thread and angle
    static_assert(kStage * kBpThreads == kBpRows * kBpWin, "staging pattern");

    const int tid = threadIdx.x;
    const int tile = b / nrowblk;
    const int rb = b % nrowblk;
    const int ybase = rb * kBpRows;
    const int j = tile_cols[tile * kTileCols + tid];
    const int X0 = tile_org[2 * tile], Z0 = tile_org[2 * tile + 1];
    int cX = 0, cZ = 0;
    if (j >= 0) {
        cX = coord[2 * j];
        cZ = coord[2 * j + 1];
    }
    float4 acc[kBpRows / 4];
#pragma unroll
    for (int q = 0; q < kBpRows / 4; q++) acc[q] = make_float4(0.f, 0.f, 0.f, 0.f);

    // footprint windows of block a0's angles (threads < kBpAngBlk)
    auto windows = [&](const int a0, int *su) __attribute__((always_inline)) {
        if (tid < min(kBpAngBlk, na - a0)) {
            AngleConst A = ang[a0 + tid];
            double lo = 1e300, hi = -1e300;
#pragma unroll
            for (int cz = 0; cz < 2; cz++)
#pragma unroll
                for (int cx = 0; cx < 2; cx++) {
                    double v = corner_xb<M>(X0 + cx * (kTileBox - 1), Z0 + cz * (kTileBox - 1), A,
                                            radius, xcent);
                    lo = fmin(lo, v);
                    hi = fmax(hi, v);
                }
            int u0 = (int)floor(lo) - 2;
            if ((int)floor(hi) + 3 - u0 >= kBpWin) atomicOr(err, 2);
            su[tid] = u0;
        }
    };
    // block a0's footprints (rows fastest in LDS) into registers, then LDS
    float stg[kBpAngBlk][kStage];
    auto load_stage = [&](const int a0, const int *su) __attribute__((always_inline)) {
        const int nb = min(kBpAngBlk, na - a0);
#pragma unroll
        for (int ab = 0; ab < kBpAngBlk; ab++) {
            const int u0 = ab < nb ? su[ab] : 0;
            const float *pa = p + (size_t)(a0 + (ab < nb ? ab : 0)) * ny * nx;
#pragma unroll
            for (int q = 0; q < kStage; q++) {
                const int e = q * kBpThreads + tid;
                const int row = e >> 5, w = e & (kBpWin - 1);
                const int y = ybase + row, u = u0 + w;
                float v = 0.0f;
                if (ab < nb && y < ny && u >= 0 && u < nx) v = pa[(size_t)y * nx + u];
                stg[ab][q] = v;
            }
        }
    };
    auto store_stage = [&](float (*sp)[kBpWin][kBpRowPad]) __attribute__((always_inline)) {
#pragma unroll
        for (int ab = 0; ab < kBpAngBlk; ab++)
#pragma unroll
            for (int q = 0; q < kStage; q++) {
                const int e = q * kBpThreads + tid;
                sp[ab][e & (kBpWin - 1)][e >> 5] = stg[ab][q];
            }
    };

    windows(0, s_u0[0]);
    if (kBpAngBlk < na) windows(kBpAngBlk, s_u0[1]);
    __syncthreads();
    load_stage(0, s_u0[0]);
    store_stage(s_p[0]);
    __syncthreads();

    int k = 0;
    for (int a0 = 0; a0 < na; a0 += kBpAngBlk, k++) {
        const int nb = min(kBpAngBlk, na - a0);
        const int cur = k & 1, u_cur = k % 3;
        const bool more = a0 + kBpAngBlk < na;
        // this column's taps for block k (registers): from the tap table
        // (one coalesced 16-byte load per thread and angle) or computed
        Taps t[kBpAngBlk];
#pragma unroll
        for (int ab = 0; ab < kBpAngBlk; ab++) {
            t[ab] = Taps{0, 0.f, 0.f, 0.f};
            if (j >= 0 && ab < nb) {
                if (btab) {
                    const uint4 e = btab[((size_t)tile * na + a0 + ab) * kTileCols + tid];
                    t[ab] = Taps{(int)e.x, __uint_as_float(e.y), __uint_as_float(e.z), __uint_as_float(e.w)};
                } else {
                    t[ab] = tap_weights<M>(cX, cZ, ang[a0 + ab], radius, xcent);
                }
            }
        }
        if (more) load_stage(a0 + kBpAngBlk, s_u0[(k + 1) % 3]);
        if (a0 + 2 * kBpAngBlk < na) windows(a0 + 2 * kBpAngBlk, s_u0[(k + 2) % 3]);
        if (j >= 0) {
#pragma unroll
            for (int ab = 0; ab < kBpAngBlk; ab++) {
                if (ab >= nb) break;
                int o = t[ab].p0 - s_u0[u_cur][ab];
                if (o < 0 || o > kBpWin - 3) {  // outside the staged window
                    atomicOr(err, 4);
                    continue;
                }
                const float w0 = t[ab].w0, w1 = t[ab].w1, w2 = t[ab].w2;
                const float4 *c0 = reinterpret_cast<const float4 *>(&s_p[cur][ab][o][0]);
                const float4 *c1 = reinterpret_cast<const float4 *>(&s_p[cur][ab][o + 1][0]);
                const float4 *c2 = reinterpret_cast<const float4 *>(&s_p[cur][ab][o + 2][0]);
                if constexpr (M == kLinear) {
                    // two taps: the linear pair's third weight is zero
#pragma unroll
                    for (int q = 0; q < kBpRows / 4; q++) {
                        float4 v0 = c0[q], v1 = c1[q];
                        float4 a = acc[q];
                        a.x = fmaf(w1, v1.x, fmaf(w0, v0.x, a.x));
                        a.y = fmaf(w1, v1.y, fmaf(w0, v0.y, a.y));
                        a.z = fmaf(w1, v1.z, fmaf(w0, v0.z, a.z));
                        a.w = fmaf(w1, v1.w, fmaf(w0, v0.w, a.w));
                        acc[q] = a;
                    }
                    continue;
                }
#pragma unroll
                for (int q = 0; q < kBpRows / 4; q++) {
                    float4 v0 = c0[q], v1 = c1[q], v2 = c2[q];
                    float4 a = acc[q];
                    a.x = fmaf(w2, v2.x, fmaf(w1, v1.x, fmaf(w0, v0.x, a.x)));
                    a.y = fmaf(w2, v2.y, fmaf(w1, v1.y, fmaf(w0, v0.y, a.y)));
                    a.z = fmaf(w2, v2.z, fmaf(w1, v1.z, fmaf(w0, v0.z, a.z)));
                    a.w = fmaf(w2, v2.w, fmaf(w1, v1.w, fmaf(w0, v0.w, a.w)));
                    acc[q] = a;
                }
            }
        }
        if (more) store_stage(s_p[cur ^ 1]);
        __syncthreads();
    }

    double part = 0.0;
    if (j >= 0 && ybase + kBpRows <= ny && (ny & 3) == 0) {
        // whole tile with 16-byte aligned column rows: vector loads and
        // stores; the same FP32 values and FP64 sum order as below
        float4 *o = reinterpret_cast<float4 *>(out + (size_t)j * ny + ybase);
        const float4 *bv = bvec ? reinterpret_cast<const float4 *>(bvec + (size_t)j * ny + ybase) : nullptr;
#pragma unroll
        for (int q = 0; q < kBpRows / 4; q++) {
            float v[4] = {acc[q].x, acc[q].y, acc[q].z, acc[q].w};
            if (MODE == kBpSirtGrad) {
                const float4 b4 = bv[q];
                const float bb[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    v[r] = bb[r] - v[r];
                    part += (double)v[r] * (double)v[r];
                }
            } else if (MODE == kBpStoreNorm) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    float s2 = v[r] * v[r];
                    part += (double)s2;
                }
            }
            o[q] = make_float4(v[0], v[1], v[2], v[3]);
        }
    } else if (j >= 0) {
        float *o = out + (size_t)j * ny + ybase;
        const float *bv = bvec ? bvec + (size_t)j * ny + ybase : nullptr;
#pragma unroll
        for (int q = 0; q < kBpRows / 4; q++) {
            float v[4] = {acc[q].x, acc[q].y, acc[q].z, acc[q].w};
#pragma unroll
            for (int r = 0; r < 4; r++) {
                int yy = 4 * q + r;
                if (ybase + yy < ny) {
                    float u = v[r];
                    if (MODE == kBpSirtGrad) {
                        u = bv[yy] - u;
                        part += (double)u * (double)u;
                    } else if (MODE == kBpStoreNorm) {
                        float s2 = u * u;
                        part += (double)s2;
                    }
                    o[yy] = u;
                }
            }
        }
    }
    if (MODE != kBpStore) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off, 64);
        if ((tid & 63) == 0) s_red[tid >> 6] = part;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < kBpThreads / 64; w++) s += s_red[w];
            partials[b] = s;
        }
        __syncthreads();  // s_red is reused by the next item of a persistent workgroup
    }
}

template <int M, int MODE, int S>
__device__ __forceinline__ void backward_item(const BpArgs &B, int b) {
    backward_body<M, MODE, kBpShapeRows[S], kBpShapeAngles[S]>(B.p, B.out, B.bvec, B.btab, B.tile_cols, B.tile_org, B.coord, B.ang, B.partials,
                           B.err, B.radius, B.xcent, B.nx, B.ny, B.na, B.nrowblk, b);
}

template <int M, int MODE, int S>
__global__ __launch_bounds__(kBpThreads) void backward_kernel(BpArgs B, const int *__restrict__ skip) {
    if (skip && *skip) return;
    // Workgroup ids go round-robin over the XCDs, so each XCD takes the same
    // row blocks of neighbouring tiles, which share detector pixels (cheaper
    // than grouping a tile's row blocks on one XCD: 42 vs 61 GB fetched per
    // launch at 512^3/720).
    backward_item<M, MODE, S>(B, blockIdx.x);
}

// Persistent back projector over many problems: workgroups take work items
// (one back-projection workgroup's job of one problem) from a queue until it
// is empty.  *counter must be 0 at launch.
template <int M, int MODE, int S>
__global__ __launch_bounds__(kBpThreads) __attribute__((amdgpu_waves_per_eu(S == 2 ? 2 : 4, 8))) void backward_batch_kernel(const BatchBp *__restrict__ probs, int nprob,
                                                                   int total, int *__restrict__ counter) {
    __shared__ int s_item;
    __shared__ int s_item0[kBatchLdsProbs];
    stage_item0(probs, nprob, s_item0);
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= total) break;
        const int k = __builtin_amdgcn_readfirstlane(owner_of(probs, s_item0, nprob, it));
        const int *skip = probs[k].skip;
        if (skip && *as_global(skip)) continue;
        BpArgs a = probs[k].a;
        a.p = as_global(a.p);
        a.out = as_global(a.out);
        a.bvec = as_global(a.bvec);
        a.btab = as_global(a.btab);
        a.tile_cols = as_global(a.tile_cols);
        a.tile_org = as_global(a.tile_org);
        a.coord = as_global(a.coord);
        a.ang = as_global(a.ang);
        a.partials = as_global(a.partials);
        a.err = as_global(a.err);
        backward_item<M, MODE, S>(a, it - probs[k].item0);
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int64_t forward_partial_count(const Geometry &g) {
    if (g.fwd_design >= 2) return (int64_t)g.na * ((g.ny + 64 * g.strip_rpl - 1) / (64 * g.strip_rpl)) * g.nstrips;
    return (int64_t)g.na * g.fwd_chunks * g.nseg;
}

int64_t backward_partial_count(const Geometry &g) { return backward_items(g); }

int64_t forward_items(const Geometry &g) {
    return (int64_t)((g.na + kFwdAngles - 1) / kFwdAngles) * g.nseg * g.fwd_chunks;
}

int64_t backward_items(const Geometry &g) {
    const int rows = kBpShapeRows[g.bp_shape];
    int nrowblk = (g.ny + rows - 1) / rows;
    return (int64_t)g.ntiles * nrowblk;
}

FwdArgs fwd_args(const Geometry &g, const float *x, float *p, double *partials) {
    FwdArgs A;
    A.x = x;
    A.p = p;
    A.table = g.d_fwd;
    A.seg = reinterpret_cast<const int4 *>(g.d_seg);
    A.pixstart = g.d_pixstart;
    A.partials = partials;
    A.npix = g.npix;
    A.nvalid = g.nvalid;
    A.nseg = g.nseg;
    A.segw = g.segw;
    A.nx = g.nx;
    A.ny = g.ny;
    A.na = g.na;
    A.nchunks = g.fwd_chunks;
    return A;
}

BpArgs bp_args(const Geometry &g, const float *p, float *x, const float *bvec, double *partials) {
    BpArgs B;
    B.p = p;
    B.out = x;
    B.bvec = bvec;
    B.btab = g.d_btab;
    B.tile_cols = g.d_tile_cols;
    B.tile_org = g.d_tile_org;
    B.coord = g.d_coord;
    B.ang = g.d_ang;
    B.partials = partials;
    B.err = g.d_err;
    B.radius = g.radius;
    B.xcent = g.xcent;
    B.nx = g.nx;
    B.ny = g.ny;
    B.na = g.na;
    B.nrowblk = (g.ny + kBpShapeRows[g.bp_shape] - 1) / kBpShapeRows[g.bp_shape];
    return B;
}

template <typename T>
static int upload(T **dst, const std::vector<T> &src, hipStream_t st) {
    AREC_HIP(hipMalloc((void **)dst, sizeof(T) * std::max<size_t>(1, src.size())));
    if (!src.empty())
        AREC_HIP(hipMemcpyAsync(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, st));
    return 0;
}

static void release_staged_tables(Geometry &g);

void geometry_free_device(Geometry &g) {
    hipFree(g.d_ang);
    hipFree(g.d_coord);
    hipFree(g.d_tile_cols);
    hipFree(g.d_tile_org);
    hipFree(g.d_fwd);
    hipFree(g.d_btab);
    g.d_btab = nullptr;
    hipFree(g.d_strip);
    hipFree(g.d_strip_base);
    hipFree(g.d_strip_off);
    g.d_strip = nullptr;
    g.d_strip_base = nullptr;
    g.d_strip_off = nullptr;
    release_staged_tables(g);
    hipFree(g.d_err);
    hipFree(g.d_seg);
    hipFree(g.d_pixstart);
    g.d_seg = nullptr;
    g.d_pixstart = nullptr;
    g.d_ang = nullptr;
    g.d_coord = nullptr;
    g.d_tile_cols = nullptr;
    g.d_tile_org = nullptr;
    g.d_fwd = nullptr;
    g.d_err = nullptr;
    g.device_ready = false;
}

// Design 1's table (the detector sweep's entries of every angle): the order
// on the host in batches of angles (a stable counting sort by first tap, which
// also fills pixstart), the weights on the device.  It is released while
// another design is selected (their tables replace it) and rebuilt, with the
// same contents, when design 1 is selected again.
static int build_sweep_table(Geometry &g, hipStream_t st) {
    size_t entries = (size_t)g.na * std::max(1, g.nvalid);
    AREC_HIP(hipMalloc((void **)&g.d_fwd, sizeof(uint4) * entries));
    int batch = (int)std::max<int64_t>(1, std::min<int64_t>(g.na, (int64_t)(64 << 20) / std::max(1, g.nvalid)));
    std::vector<int> order;
    int *d_order = nullptr;
    AREC_HIP(hipMalloc((void **)&d_order, sizeof(int) * (size_t)batch * std::max(1, g.nvalid)));
    for (int a0 = 0; a0 < g.na && g.nvalid > 0; a0 += batch) {
        int a1 = std::min(g.na, a0 + batch);
        geometry_sort_angles(g, a0, a1, order);
        AREC_HIP(hipStreamSynchronize(st));  // previous batch done with d_order
        AREC_HIP(hipMemcpyAsync(d_order, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, st));
        dim3 grid((g.nvalid + 255) / 256, a1 - a0);
        if (grid.x > 1024) grid.x = 1024;
        if (g.method == kSQ)
            build_fwd_table<kSQ><<<grid, 256, 0, st>>>(d_order, a0, a1 - a0, g.nvalid, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_fwd, g.d_err);
        else if (g.method == kKB)
            build_fwd_table<kKB><<<grid, 256, 0, st>>>(d_order, a0, a1 - a0, g.nvalid, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_fwd, g.d_err);
        else
            build_fwd_table<kLinear><<<grid, 256, 0, st>>>(d_order, a0, a1 - a0, g.nvalid, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_fwd, g.d_err);
        AREC_HIP(hipGetLastError());
    }
    AREC_HIP(hipStreamSynchronize(st));
    hipFree(d_order);
    return 0;
}

static void release_sweep_table(Geometry &g) {
    hipFree(g.d_fwd);
    g.d_fwd = nullptr;
}

// Design 1's table back after another design's was released (or failed to
// build); the device's self-check flag reports disagreeing taps.
static int ensure_sweep_table(Geometry &g, hipStream_t st) {
    if (g.d_fwd) return 0;
    AREC_HIP(hipMemsetAsync(g.d_err, 0, sizeof(int), st));
    int rc = build_sweep_table(g, st);
    if (rc) {
        release_sweep_table(g);
        return rc;
    }
    int herr = 0;
    AREC_HIP(hipMemcpy(&herr, g.d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        set_error("forward table: device taps disagree with the host ordering");
        return AREC_EGEOM;
    }
    return 0;
}

int geometry_upload(Geometry &g, hipStream_t st, std::string *err) {
    int rc;
    if ((rc = upload(&g.d_ang, g.ang, st))) return rc;
    if ((rc = upload(&g.d_coord, g.coord, st))) return rc;
    if ((rc = upload(&g.d_tile_cols, g.tile_cols, st))) return rc;
    if ((rc = upload(&g.d_tile_org, g.tile_org, st))) return rc;
    AREC_HIP(hipMalloc((void **)&g.d_err, sizeof(int)));
    AREC_HIP(hipMemsetAsync(g.d_err, 0, sizeof(int), st));
    geometry_plan_forward(g, 0, 0);
    g.npix = g.nx + 8;
    g.pixstart.assign((size_t)g.na * g.npix, 0);
    if ((rc = build_sweep_table(g, st))) return rc;
    // AREC_NO_BACK_TABLE forces the on-the-fly taps (same results; tests use it)
    const int64_t btab_entries = (int64_t)g.ntiles * g.na * kTileCols;
    if (g.ntiles > 0 && btab_entries * (int64_t)sizeof(uint4) <= kMaxBackTableBytes &&
        !getenv("AREC_NO_BACK_TABLE")) {
        AREC_HIP(hipMalloc((void **)&g.d_btab, sizeof(uint4) * btab_entries));
        dim3 bg((unsigned)std::min<int64_t>((btab_entries + 255) / 256, 65536));
        if (g.method == kSQ)
            build_back_table<kSQ><<<bg, 256, 0, st>>>(g.d_tile_cols, g.ntiles, g.na, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_btab);
        else if (g.method == kKB)
            build_back_table<kKB><<<bg, 256, 0, st>>>(g.d_tile_cols, g.ntiles, g.na, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_btab);
        else
            build_back_table<kLinear><<<bg, 256, 0, st>>>(g.d_tile_cols, g.ntiles, g.na, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_btab);
        AREC_HIP(hipGetLastError());
    }
    geometry_segments(g);
    if ((rc = upload(&g.d_seg, g.seg, st))) return rc;
    if ((rc = upload(&g.d_pixstart, g.pixstart, st))) return rc;
    AREC_HIP(hipStreamSynchronize(st));
    int herr = 0;
    AREC_HIP(hipMemcpy(&herr, g.d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        *err = "forward table: device taps disagree with the host ordering";
        return AREC_EGEOM;
    }
    g.device_ready = true;
    return 0;
}

int geometry_replan_forward(Geometry &g, int rpl, int min_groups, hipStream_t st) {
    if (!g.device_ready) {
        set_error("geometry not uploaded");
        return AREC_EINVAL;
    }
    if (!geometry_plan_forward(g, rpl, min_groups)) {
        set_error("forward plan: rows per lane must be 1, 2 or 4 and divide the row count");
        return AREC_EINVAL;
    }
    geometry_segments(g);
    AREC_HIP(hipStreamSynchronize(st));
    hipFree(g.d_seg);
    g.d_seg = nullptr;
    return upload(&g.d_seg, g.seg, st);
}

static int build_pair_tables(Geometry &g, hipStream_t st);

static int build_strip_tables(Geometry &g, hipStream_t st) {
    if (g.strip_unit == 2) return build_pair_tables(g, st);
    geometry_strip_count(g);
    const int64_t total = g.strip_base[g.na];
    if (total * (int64_t)sizeof(uint4) > kMaxStripTableBytes) {
        set_error("depth-strip forward table too large for this geometry");
        return AREC_ESIZE;
    }
    int rc;
    AREC_HIP(hipMalloc((void **)&g.d_strip, sizeof(uint4) * std::max<int64_t>(1, total)));
    if ((rc = upload(&g.d_strip_base, g.strip_base, st))) return rc;
    AREC_HIP(hipMemsetAsync(g.d_err, 0, sizeof(int), st));
    // angles in batches of <= 64 M entries: order on the host, weights on the device
    std::vector<uint32_t> order;
    uint32_t *d_order = nullptr;
    int64_t cap = 0;
    for (int a0 = 0; a0 < g.na;) {
        int a1 = a0 + 1;
        while (a1 < g.na && g.strip_base[a1 + 1] - g.strip_base[a0] <= (int64_t)(64 << 20)) a1++;
        geometry_strip_angles(g, a0, a1, order);
        AREC_HIP(hipStreamSynchronize(st));  // previous batch done with d_order
        if ((int64_t)order.size() > cap) {
            hipFree(d_order);
            cap = (int64_t)order.size();
            AREC_HIP(hipMalloc((void **)&d_order, sizeof(uint32_t) * std::max<int64_t>(1, cap)));
        }
        AREC_HIP(hipMemcpyAsync(d_order, order.data(), sizeof(uint32_t) * order.size(), hipMemcpyHostToDevice, st));
        int64_t most = 0;
        for (int a = a0; a < a1; a++) most = std::max(most, g.strip_base[a + 1] - g.strip_base[a]);
        dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>((most + 255) / 256, 1024)), a1 - a0);
        if (g.method == kSQ)
            build_strip_table<kSQ><<<grid, 256, 0, st>>>(d_order, g.d_strip_base, a0, a1 - a0, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_strip, g.d_err);
        else if (g.method == kKB)
            build_strip_table<kKB><<<grid, 256, 0, st>>>(d_order, g.d_strip_base, a0, a1 - a0, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_strip, g.d_err);
        else
            build_strip_table<kLinear><<<grid, 256, 0, st>>>(d_order, g.d_strip_base, a0, a1 - a0, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_strip, g.d_err);
        AREC_HIP(hipGetLastError());
        a0 = a1;
    }
    AREC_HIP(hipStreamSynchronize(st));
    hipFree(d_order);
    if ((rc = upload(&g.d_strip_off, g.strip_off, st))) return rc;
    AREC_HIP(hipStreamSynchronize(st));
    int herr = 0;
    AREC_HIP(hipMemcpy(&herr, g.d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        set_error("depth-strip table: device taps disagree with the host ordering");
        return AREC_EGEOM;
    }
    return 0;
}

// Design 3's table: the pair ordering on the host, the weights on the device.
static int build_pair_tables(Geometry &g, hipStream_t st) {
    if ((g.nx + kStripW - 1) / kStripW > 256) {
        set_error("angle-pair forward table: at most 256 strips (detector width 4096)");
        return AREC_ESIZE;
    }
    geometry_pair_count(g);
    const int nu = (g.na + 1) / 2;
    const int64_t total = g.strip_base[nu];
    if (2 * total * (int64_t)sizeof(uint4) > kMaxStripTableBytes) {
        set_error("angle-pair forward table too large for this geometry");
        return AREC_ESIZE;
    }
    int rc;
    AREC_HIP(hipMalloc((void **)&g.d_strip, 2 * sizeof(uint4) * std::max<int64_t>(1, total)));
    if ((rc = upload(&g.d_strip_base, g.strip_base, st))) return rc;
    AREC_HIP(hipMemsetAsync(g.d_err, 0, sizeof(int), st));
    std::vector<uint32_t> order;
    uint32_t *d_order = nullptr;
    int64_t cap = 0;
    for (int u0 = 0; u0 < nu;) {
        int u1 = u0 + 1;
        while (u1 < nu && g.strip_base[u1 + 1] - g.strip_base[u0] <= (int64_t)(64 << 20)) u1++;
        geometry_pair_units(g, u0, u1, order);
        AREC_HIP(hipStreamSynchronize(st));  // previous batch done with d_order
        if ((int64_t)order.size() > cap) {
            hipFree(d_order);
            cap = (int64_t)order.size();
            AREC_HIP(hipMalloc((void **)&d_order, sizeof(uint32_t) * std::max<int64_t>(1, cap)));
        }
        AREC_HIP(hipMemcpyAsync(d_order, order.data(), sizeof(uint32_t) * order.size(), hipMemcpyHostToDevice, st));
        int64_t most = 0;
        for (int u = u0; u < u1; u++) most = std::max(most, g.strip_base[u + 1] - g.strip_base[u]);
        dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>((most + 255) / 256, 1024)), u1 - u0);
        if (g.method == kSQ)
            build_pair_table<kSQ><<<grid, 256, 0, st>>>(d_order, g.d_strip_base, u0, u1 - u0, g.na, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_strip, g.d_err);
        else if (g.method == kKB)
            build_pair_table<kKB><<<grid, 256, 0, st>>>(d_order, g.d_strip_base, u0, u1 - u0, g.na, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_strip, g.d_err);
        else
            build_pair_table<kLinear><<<grid, 256, 0, st>>>(d_order, g.d_strip_base, u0, u1 - u0, g.na, g.d_coord, g.d_ang, g.radius, g.xcent, g.d_strip, g.d_err);
        AREC_HIP(hipGetLastError());
        u0 = u1;
    }
    AREC_HIP(hipStreamSynchronize(st));
    hipFree(d_order);
    if ((rc = upload(&g.d_strip_off, g.strip_off, st))) return rc;
    AREC_HIP(hipStreamSynchronize(st));
    int herr = 0;
    AREC_HIP(hipMemcpy(&herr, g.d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        set_error("angle-pair table: device taps disagree with the host ordering");
        return AREC_EGEOM;
    }
    return 0;
}

// Design 4's tables: the staged ordering on the host, in batches of angle
// groups of at most 32 M entries, and the weights on the device.
static int build_staged_tables(Geometry &g, const int rpl, hipStream_t st) {
    if (g.na > 65535 || (g.nx + kStripW - 1) / kStripW > 65535) {
        set_error("staged-strip forward table: at most 65535 angles and strips");
        return AREC_ESIZE;
    }
    const auto tc0 = std::chrono::steady_clock::now();
    if (geometry_staged_count(g) > kSgEnt) {
        set_error("staged-strip forward table: a chunk holds more entries for one angle than the kernel stages");
        return AREC_ESIZE;
    }
    if (getenv("AREC_PROGRESS"))
        fprintf(stderr, "[staged table] %lld entries counted in %.1f s\n", (long long)g.sg_group_ent.back(),
                std::chrono::duration<double>(std::chrono::steady_clock::now() - tc0).count());
    const int ns = g.nstrips, ngroups = (g.na + kSgAngles - 1) / kSgAngles;
    const int64_t nchunk = g.sg_chunk_base.back(), total = g.sg_group_ent.back();
    // Limited by the device memory free now (design 1's table is released
    // before this build), leaving room for a solve: a SIRT's work vectors (two
    // volumes and a projection stack) and 1 GiB.
    const int64_t bytes =
        (total + kStripBatch) * (int64_t)sizeof(uint4) + nchunk * (int64_t)(kSgChunk * 4 + kSgAngles * 8);
    size_t free_b = 0, total_b = 0;
    AREC_HIP(hipMemGetInfo(&free_b, &total_b));
    const int64_t reserve = (int64_t)sizeof(float) * (2 * g.nnz + g.nproj) + (1ll << 30);
    if (bytes + reserve > (int64_t)free_b) {
        set_error("staged-strip forward table (" + std::to_string(bytes >> 20) + " MiB) does not fit in the " +
                  std::to_string((int64_t)free_b >> 20) + " MiB of free device memory with room for a solve");
        return AREC_ESIZE;
    }
    int rc;
    AREC_HIP(hipMalloc((void **)&g.d_sg_tab, sizeof(uint4) * (size_t)(total + kStripBatch)));
    AREC_HIP(hipMemsetAsync(g.d_sg_tab + total, 0, sizeof(uint4) * kStripBatch, st));
    AREC_HIP(hipMalloc((void **)&g.d_sg_stage, sizeof(int) * (size_t)std::max<int64_t>(1, nchunk * kSgChunk)));
    AREC_HIP(hipMalloc((void **)&g.d_sg_off, sizeof(int64_t) * (size_t)(nchunk * kSgAngles + 1)));
    if ((rc = upload(&g.d_sg_chunk_base, g.sg_chunk_base, st))) return rc;
    AREC_HIP(hipMemsetAsync(g.d_err, 0, sizeof(int), st));
    uint32_t *d_tmp = nullptr;
    int64_t cap = 0;
    // progress lines on stderr when AREC_PROGRESS is set (the host ordering
    // of the full slab's table takes minutes)
    const bool progress = getenv("AREC_PROGRESS") != nullptr;
    // batches of angle groups: up to 32 M entries, and at least one group per
    // host thread (the groups of a batch are ordered in parallel) within
    // 512 M entries (6 GB of staging on each side)
    const int min_groups = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<int> bounds{0};
    int64_t most_ent = 0, most_chunks = 0;
    while (bounds.back() < ngroups) {
        const int G0 = bounds.back();
        int G1 = G0 + 1;
        while (G1 < ngroups && (g.sg_group_ent[G1 + 1] - g.sg_group_ent[G0] <= (int64_t)(32 << 20) ||
                                (G1 - G0 < min_groups && g.sg_group_ent[G1 + 1] - g.sg_group_ent[G0] <= (int64_t)(512 << 20))))
            G1++;
        most_ent = std::max(most_ent, g.sg_group_ent[G1] - g.sg_group_ent[G0]);
        most_chunks = std::max(most_chunks, g.sg_chunk_base[(size_t)G1 * ns] - g.sg_chunk_base[(size_t)G0 * ns]);
        bounds.push_back(G1);
    }
    // two host staging sets, each sized for the largest batch and written
    // once per batch by the ordering threads (no zero fill): batch b is
    // ordered into set b & 1 while a helper thread uploads batch b - 1 from
    // the other set and runs its table kernel (one upload at a time, so one
    // device buffer)
    const int nset = bounds.size() > 2 ? 2 : 1;
    std::unique_ptr<uint32_t[]> h_ent[2];
    std::unique_ptr<int[]> h_stage[2];
    std::unique_ptr<int64_t[]> h_off[2];
    for (int k = 0; k < nset; k++) {
        h_ent[k].reset(new uint32_t[(size_t)std::max<int64_t>(1, 3 * most_ent)]);
        h_stage[k].reset(new int[(size_t)std::max<int64_t>(1, most_chunks * kSgChunk)]);
        h_off[k].reset(new int64_t[(size_t)std::max<int64_t>(1, most_chunks * kSgAngles)]);
    }
    cap = std::max<int64_t>(1, most_ent);
    AREC_HIP(hipMalloc((void **)&d_tmp, 3 * sizeof(uint32_t) * (size_t)cap));
    int dev = 0;
    AREC_HIP(hipGetDevice(&dev));
    std::atomic<int> up_err{0};  // the first failing call's hipError_t
    auto upload = [&](const size_t bi, const int k) {
        hipError_t e = hipSetDevice(dev);
        const int G0 = bounds[bi], G1 = bounds[bi + 1];
        const int64_t nchunks_b = g.sg_chunk_base[(size_t)G1 * ns] - g.sg_chunk_base[(size_t)G0 * ns];
        const int64_t n = g.sg_group_ent[G1] - g.sg_group_ent[G0];
        const int64_t K0 = g.sg_chunk_base[(size_t)G0 * ns], E0 = g.sg_group_ent[G0];
        const uint32_t *col = h_ent[k].get(), *angi = col + most_ent, *meta = angi + most_ent;
        if (e == hipSuccess)
            e = hipMemcpyAsync(g.d_sg_stage + K0 * kSgChunk, h_stage[k].get(),
                               sizeof(int) * (size_t)(nchunks_b * kSgChunk), hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(g.d_sg_off + K0 * kSgAngles, h_off[k].get(),
                               sizeof(int64_t) * (size_t)(nchunks_b * kSgAngles), hipMemcpyHostToDevice, st);
        if (e == hipSuccess && n > 0) {
            e = hipMemcpyAsync(d_tmp, col, 4 * (size_t)n, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipMemcpyAsync(d_tmp + cap, angi, 4 * (size_t)n, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) e = hipMemcpyAsync(d_tmp + 2 * cap, meta, 4 * (size_t)n, hipMemcpyHostToDevice, st);
            if (e == hipSuccess) {
                dim3 grid((unsigned)std::min<int64_t>((n + 255) / 256, 65536));
                if (g.method == kSQ)
                    build_staged_table<kSQ><<<grid, 256, 0, st>>>(d_tmp, d_tmp + cap, d_tmp + 2 * cap, n, g.d_coord, g.d_ang, g.radius, g.xcent, 256 * rpl, g.d_sg_tab + E0, g.d_err);
                else if (g.method == kKB)
                    build_staged_table<kKB><<<grid, 256, 0, st>>>(d_tmp, d_tmp + cap, d_tmp + 2 * cap, n, g.d_coord, g.d_ang, g.radius, g.xcent, 256 * rpl, g.d_sg_tab + E0, g.d_err);
                else
                    build_staged_table<kLinear><<<grid, 256, 0, st>>>(d_tmp, d_tmp + cap, d_tmp + 2 * cap, n, g.d_coord, g.d_ang, g.radius, g.xcent, 256 * rpl, g.d_sg_tab + E0, g.d_err);
                e = hipGetLastError();
            }
        }
        if (e == hipSuccess) e = hipStreamSynchronize(st);  // the staging set and d_tmp are free again
        if (e != hipSuccess) {
            int z = 0;
            up_err.compare_exchange_strong(z, (int)e);
        }
    };
    std::thread up;
    double t_host = 0.0, t_copy = 0.0;  // seconds: host ordering, waits for the uploads
    auto now = []() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
    for (size_t bi = 0; bi + 1 < bounds.size(); bi++) {
        const int G0 = bounds[bi], G1 = bounds[bi + 1];
        const int k = (int)(bi % (size_t)nset);
        if (nset == 1 && up.joinable()) up.join();  // one set: its upload must be done first
        double t0 = now();
        uint32_t *col = h_ent[k].get(), *angi = col + most_ent, *meta = angi + most_ent;
        const bool ok = geometry_staged_groups(g, G0, G1, h_stage[k].get(), h_off[k].get(), col, angi, meta);
        double t1 = now();
        t_host += t1 - t0;
        if (up.joinable()) up.join();  // the previous batch (the other set) is uploaded
        t_copy += now() - t1;
        if (!ok || up_err.load()) {
            hipFree(d_tmp);
            if (!ok) set_error("staged-strip table: the host ordering's two passes disagree");
            else set_error(std::string("staged-strip table upload: ") + hipGetErrorString((hipError_t)up_err.load()));
            return ok ? AREC_EHIP : AREC_EGEOM;
        }
        up = std::thread(upload, bi, k);
        if (progress)
            fprintf(stderr, "[staged table] angle groups %d of %d ordered and uploaded (%.1f s ordering, %.1f s uploads)\n",
                    G1, ngroups, t_host, t_copy);
    }
    {
        const double t1 = now();
        if (up.joinable()) up.join();
        t_copy += now() - t1;
    }
    if (up_err.load()) {
        hipFree(d_tmp);
        set_error(std::string("staged-strip table upload: ") + hipGetErrorString((hipError_t)up_err.load()));
        return AREC_EHIP;
    }
    std::vector<StagedOrder>().swap(g.sg_order);  // every kept order has been used
    AREC_HIP(hipMemcpyAsync(g.d_sg_off + nchunk * kSgAngles, &g.sg_group_ent.back(), sizeof(int64_t),
                            hipMemcpyHostToDevice, st));
    AREC_HIP(hipStreamSynchronize(st));
    hipFree(d_tmp);
    int herr = 0;
    AREC_HIP(hipMemcpy(&herr, g.d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        set_error("staged-strip table: device taps disagree with the host ordering");
        return AREC_EGEOM;
    }
    return 0;
}

static void release_staged_tables(Geometry &g) {
    hipFree(g.d_sg_tab);
    hipFree(g.d_sg_stage);
    hipFree(g.d_sg_off);
    hipFree(g.d_sg_chunk_base);
    g.d_sg_tab = nullptr;
    g.d_sg_stage = nullptr;
    g.d_sg_off = nullptr;
    g.d_sg_chunk_base = nullptr;
    g.sg_rows4 = 0;
    std::vector<int64_t>().swap(g.sg_chunk_base);
    std::vector<int64_t>().swap(g.sg_group_ent);
    std::vector<StagedOrder>().swap(g.sg_order);
}

// Frees the depth-strip tables (device and host).
static void release_strip_tables(Geometry &g) {
    hipFree(g.d_strip);
    hipFree(g.d_strip_base);
    hipFree(g.d_strip_off);
    g.d_strip = nullptr;
    g.d_strip_base = nullptr;
    g.d_strip_off = nullptr;
    std::vector<int64_t>().swap(g.strip_base);
    std::vector<int>().swap(g.strip_off);
}

// Designs 2-4 replace design 1's table while selected (it is released before
// their tables are built, and rebuilt with the same contents when design 1
// is selected again or their build fails); design 1's segment plan and the
// back projector's data stay.  Only one of the tables of designs 2-4 exists
// at a time; switching between designs 2 and 3 rebuilds theirs.
int geometry_set_forward(Geometry &g, int design, int rpl, int min_groups, hipStream_t st) {
    if (!g.device_ready) {
        set_error("forward design: the geometry is not uploaded");
        return AREC_EINVAL;
    }
    if (design < 1 || design > 4) {
        set_error("forward design must be 1 (sweep), 2 (depth strips), 3 (depth strips, angle pairs) or 4 "
                  "(staged strips)");
        return AREC_EINVAL;
    }
    if (design == 1) {
        int rc = geometry_replan_forward(g, rpl, min_groups, st);
        if (rc) return rc;
        AREC_HIP(hipStreamSynchronize(st));
        release_staged_tables(g);
        release_strip_tables(g);
        if ((rc = ensure_sweep_table(g, st))) return rc;
        g.fwd_design = 1;
        return 0;
    }
    if (design == 4) {
        // any row count: whole row chunks take vector loads only where the
        // rows per lane divide the row count (column starts aligned)
        if (rpl == 0) rpl = g.ny >= 256 ? 4 : g.ny >= 128 ? 2 : 1;
        if (rpl != 1 && rpl != 2 && rpl != 4) {
            set_error("forward plan: rows per lane must be 1, 2 or 4");
            return AREC_EINVAL;
        }
    } else {
        if (rpl == 0)
            rpl = design == 3 ? ((g.ny >= 128 && g.ny % 2 == 0) ? 2 : 1)
                              : (g.ny >= 256 && g.ny % 4 == 0) ? 4 : (g.ny >= 128 && g.ny % 2 == 0) ? 2 : 1;
        if (!fwd_rpl_allowed(g.ny, rpl) || (design == 3 && rpl == 4)) {
            set_error("forward plan: rows per lane must be 1, 2 or 4 (design 3: 1 or 2) and divide the row count");
            return AREC_EINVAL;
        }
    }
    const int unit = design == 3 ? 2 : 1;
    AREC_HIP(hipStreamSynchronize(st));
    if (design == 4) {
        release_strip_tables(g);
    } else {
        release_staged_tables(g);
        if (g.d_strip && g.strip_unit != unit) release_strip_tables(g);
    }
    release_sweep_table(g);
    int rc = 0;
    if (design == 4 && !g.d_sg_tab) {
        rc = build_staged_tables(g, rpl, st);
        if (!rc) g.sg_rows4 = 256 * rpl;
    } else if (design == 4 && g.sg_rows4 != 256 * rpl) {
        // entry words hold the slot's byte offset for the current rows per lane
        const int64_t n = g.sg_group_ent.back() + kStripBatch;
        dim3 grid((unsigned)std::min<int64_t>((n + 255) / 256, 65536));
        restride_staged_table<<<grid, 256, 0, st>>>(g.d_sg_tab, n, g.sg_rows4, 256 * rpl);
        AREC_HIP(hipGetLastError());
        g.sg_rows4 = 256 * rpl;
    } else if (design != 4 && !g.d_strip) {
        g.strip_unit = unit;
        rc = build_strip_tables(g, st);
    }
    if (rc) {
        hipStreamSynchronize(st);
        release_staged_tables(g);
        release_strip_tables(g);
        g.fwd_design = 1;  // design 1 again: its table comes back
        const int rc1 = ensure_sweep_table(g, st);
        if (rc1) return rc1;
        return rc;
    }
    g.strip_rpl = rpl;  // design 1's plan keeps its own rows per lane and chunks
    g.fwd_design = design;
    return 0;
}

int launch_forward(const Geometry &g, const float *x, float *p, double *partials, hipStream_t st,
                   const int *skip) {
    if (g.nvalid == 0) return fill(p, 0.0f, g.nproj, st, skip);
    if (g.fwd_design == 1 && !g.d_fwd) {
        set_error("forward projection: design 1's table is missing");
        return AREC_EINVAL;
    }
    if (g.fwd_design == 4) {
        const int rpl = g.strip_rpl, nchunks = (g.ny + 64 * rpl - 1) / (64 * rpl);
        const int ngroups = (g.na + kSgAngles - 1) / kSgAngles;
        dim3 grid((unsigned)((int64_t)ngroups * g.nstrips * nchunks));
        // the linear pair's entries have two taps (the third weight is zero)
#define AREC_FWDG(R, T)                                                                                           \
    forward_staged_kernel<R, T><<<grid, kSgThreads, 0, st>>>(x, p, g.d_sg_tab, g.d_sg_stage, g.d_sg_chunk_base,    \
                                                             g.d_sg_off, partials, g.nstrips, g.nx, g.ny, g.na, \
                                                             nchunks, skip)
        const bool two = g.method == kLinear;
        if (rpl == 4) {
            if (two) AREC_FWDG(4, 2); else AREC_FWDG(4, 3);
        } else if (rpl == 2) {
            if (two) AREC_FWDG(2, 2); else AREC_FWDG(2, 3);
        } else {
            if (two) AREC_FWDG(1, 2); else AREC_FWDG(1, 3);
        }
#undef AREC_FWDG
        AREC_HIP(hipGetLastError());
        return 0;
    }
    if (g.fwd_design == 3) {
        const int rpl = g.strip_rpl, nchunks = (g.ny + 64 * rpl - 1) / (64 * rpl);
        const int ngroups = ((g.na + 1) / 2 + kStripAngles - 1) / kStripAngles;
        dim3 grid((unsigned)((int64_t)ngroups * g.nstrips * nchunks));
        if (rpl == 2)
            forward_pair_kernel<2><<<grid, 64 * kStripAngles, 0, st>>>(x, p, g.d_strip, g.d_strip_base, g.d_strip_off,
                                                                      partials, g.nstrips, g.nbins, g.nx, g.ny, g.na,
                                                                      nchunks, skip);
        else
            forward_pair_kernel<1><<<grid, 64 * kStripAngles, 0, st>>>(x, p, g.d_strip, g.d_strip_base, g.d_strip_off,
                                                                      partials, g.nstrips, g.nbins, g.nx, g.ny, g.na,
                                                                      nchunks, skip);
        AREC_HIP(hipGetLastError());
        return 0;
    }
    if (g.fwd_design == 2) {
        const int rpl = g.strip_rpl, nchunks = (g.ny + 64 * rpl - 1) / (64 * rpl);
        dim3 grid((unsigned)((int64_t)((g.na + kStripAngles - 1) / kStripAngles) * g.nstrips * nchunks));
#define AREC_FWDS(R)                                                                                       \
    forward_strip_kernel<R><<<grid, 64 * kStripAngles, 0, st>>>(x, p, g.d_strip, g.d_strip_base, g.d_strip_off, \
                                                              partials, g.nstrips, g.nbins, g.nx, g.ny, g.na,  \
                                                              nchunks, skip)
        if (rpl == 4)
            AREC_FWDS(4);
        else if (rpl == 2)
            AREC_FWDS(2);
        else
            AREC_FWDS(1);
#undef AREC_FWDS
        AREC_HIP(hipGetLastError());
        return 0;
    }
    dim3 grid((unsigned)forward_items(g));
    const FwdArgs A = fwd_args(g, x, p, partials);
#define AREC_FWD(R)                                                                                         \
    forward_kernel<R><<<grid, 64 * kFwdAngles, 0, st>>>(A.x, A.p, A.table, A.seg, A.pixstart, A.partials, \
                                                        A.npix, A.nvalid, A.nseg, A.segw, A.nx, A.ny, A.na, \
                                                        A.nchunks, skip)
    if (g.fwd_rpl == 4)
        AREC_FWD(4);
    else if (g.fwd_rpl == 2)
        AREC_FWD(2);
    else
        AREC_FWD(1);
#undef AREC_FWD
    AREC_HIP(hipGetLastError());
    return 0;
}

template <int M, int S>
static void bp_dispatch_shape(int mode, dim3 grid, hipStream_t st, const BpArgs &B, const int *skip) {
    if (mode == kBpSirtGrad)
        backward_kernel<M, kBpSirtGrad, S><<<grid, kBpThreads, 0, st>>>(B, skip);
    else if (mode == kBpStoreNorm)
        backward_kernel<M, kBpStoreNorm, S><<<grid, kBpThreads, 0, st>>>(B, skip);
    else
        backward_kernel<M, kBpStore, S><<<grid, kBpThreads, 0, st>>>(B, skip);
}

template <int M>
static void bp_dispatch(int mode, int shape, dim3 grid, hipStream_t st, const BpArgs &B, const int *skip) {
    static_assert(kBpShapes == 3, "one branch per shape");
    if (shape == 1)
        bp_dispatch_shape<M, 1>(mode, grid, st, B, skip);
    else if (shape == 2)
        bp_dispatch_shape<M, 2>(mode, grid, st, B, skip);
    else
        bp_dispatch_shape<M, 0>(mode, grid, st, B, skip);
}

int geometry_set_back(Geometry &g, int shape) {
    if (shape < 0 || shape >= kBpShapes) {
        set_error("back projector: unknown tile shape");
        return AREC_EINVAL;
    }
    g.bp_shape = shape;
    return 0;
}

int64_t forward_batch_item(const Geometry &g, const float *x, float *p, const int *skip, int64_t blk8,
                           BatchFwd &d) {
    const int rpl = g.strip_rpl, nchunks = (g.ny + 64 * rpl - 1) / (64 * rpl);
    const int64_t nblk = (int64_t)((g.na + kSgAngles - 1) / kSgAngles) * g.nstrips * nchunks;
    d.x = x;
    d.p = p;
    d.tab = g.d_sg_tab;
    d.stage = g.d_sg_stage;
    d.chunk_base = g.d_sg_chunk_base;
    d.off = g.d_sg_off;
    d.skip = skip;
    d.nstrips = g.nstrips;
    d.nx = g.nx;
    d.ny = g.ny;
    d.na = g.na;
    d.nchunks = nchunks;
    d.nblk = (int)nblk;
    d.blk8 = blk8;
    return (nblk + 7) / 8 * 8;
}

int launch_forward_batch(int method, int rpl, const BatchFwd *d_probs, const int *d_owner, int64_t nblk,
                         hipStream_t st) {
    if (nblk <= 0) return 0;
    if (nblk % 8 || nblk > INT32_MAX) {
        set_error("batched forward projection: workgroup count must be a multiple of 8 below 2^31");
        return AREC_ESIZE;
    }
    dim3 grid((unsigned)nblk);
    const bool two = method == kLinear;  // two taps per entry
#define AREC_FWDB(R, T) forward_staged_batch_kernel<R, T><<<grid, kSgThreads, 0, st>>>(d_probs, d_owner)
    if (rpl == 4) {
        if (two) AREC_FWDB(4, 2); else AREC_FWDB(4, 3);
    } else if (rpl == 2) {
        if (two) AREC_FWDB(2, 2); else AREC_FWDB(2, 3);
    } else if (rpl == 1) {
        if (two) AREC_FWDB(1, 2); else AREC_FWDB(1, 3);
    } else {
        set_error("batched forward projection: rows per lane must be 1, 2 or 4");
        return AREC_EINVAL;
    }
    AREC_HIP(hipGetLastError());
#undef AREC_FWDB
    return 0;
}

int launch_backward(const Geometry &g, const float *p, float *x, int mode, const float *bvec,
                    double *partials, hipStream_t st, const int *skip) {
    // columns that never contribute (linear pair's last two) stay zero
    if (g.nvalid < g.nrays) {
        int rc = fill(x + (size_t)g.nvalid * g.ny, 0.0f, (int64_t)(g.nrays - g.nvalid) * g.ny, st, skip);
        if (rc) return rc;
    }
    if (g.ntiles == 0) return 0;
    dim3 grid((unsigned)backward_items(g));
    BpArgs B = bp_args(g, p, x, bvec, partials);
    if (g.method == kSQ)
        bp_dispatch<kSQ>(mode, g.bp_shape, grid, st, B, skip);
    else if (g.method == kKB)
        bp_dispatch<kKB>(mode, g.bp_shape, grid, st, B, skip);
    else
        bp_dispatch<kLinear>(mode, g.bp_shape, grid, st, B, skip);
    AREC_HIP(hipGetLastError());
    return 0;
}

// The queue counter to zero by a kernel (a memset node would not be safe in
// a captured graph).
__global__ void zero_counter_kernel(int *c) { *c = 0; }

// Persistent grid: as many workgroups as can be resident on the device.
template <typename K>
static int resident_grid(K kernel, int threads, int *out) {
    int dev = 0, ncu = 0, per_cu = 0;
    AREC_HIP(hipGetDevice(&dev));
    AREC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    AREC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
    *out = std::max(1, ncu * std::max(1, per_cu));
    return 0;
}

template <int M, int MODE, int S>
static int bp_batch(const BatchBp *d_probs, int nprob, int total, int *d_counter, hipStream_t st) {
    int grid = 0, rc;
    if ((rc = resident_grid(backward_batch_kernel<M, MODE, S>, kBpThreads, &grid))) return rc;
    grid = std::min(grid, total);
    backward_batch_kernel<M, MODE, S><<<grid, kBpThreads, 0, st>>>(d_probs, nprob, total, d_counter);
    return 0;
}

template <int M, int S>
static int bp_batch_mode(int mode, const BatchBp *d_probs, int nprob, int total, int *d_counter,
                         hipStream_t st) {
    if (mode == kBpSirtGrad) return bp_batch<M, kBpSirtGrad, S>(d_probs, nprob, total, d_counter, st);
    if (mode == kBpStoreNorm) return bp_batch<M, kBpStoreNorm, S>(d_probs, nprob, total, d_counter, st);
    return bp_batch<M, kBpStore, S>(d_probs, nprob, total, d_counter, st);
}

template <int M>
static int bp_batch_shape(int mode, int shape, const BatchBp *d_probs, int nprob, int total, int *d_counter,
                          hipStream_t st) {
    if (shape == 1) return bp_batch_mode<M, 1>(mode, d_probs, nprob, total, d_counter, st);
    if (shape == 2) return bp_batch_mode<M, 2>(mode, d_probs, nprob, total, d_counter, st);
    return bp_batch_mode<M, 0>(mode, d_probs, nprob, total, d_counter, st);
}

int launch_backward_batch(int method, int mode, int shape, const BatchBp *d_probs, int nprob, int64_t total,
                          int *d_counter, hipStream_t st) {
    if (nprob <= 0 || total <= 0) return 0;
    if (total > INT32_MAX) {
        set_error("batched back projection: too many work items");
        return AREC_ESIZE;
    }
    zero_counter_kernel<<<1, 1, 0, st>>>(d_counter);
    int rc;
    if (method == kSQ) rc = bp_batch_shape<kSQ>(mode, shape, d_probs, nprob, (int)total, d_counter, st);
    else if (method == kKB) rc = bp_batch_shape<kKB>(mode, shape, d_probs, nprob, (int)total, d_counter, st);
    else rc = bp_batch_shape<kLinear>(mode, shape, d_probs, nprob, (int)total, d_counter, st);
    if (rc) return rc;
    AREC_HIP(hipGetLastError());
    return 0;
}

}  // namespace arec
