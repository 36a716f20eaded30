// Learned warm start: a 3-level 3-D U-Net (16/32/64 channels, 3x3x3
// convolutions, ReLU, BF16 weights and activations, FP32 accumulation) with
// hand-written gfx950 kernels for inference and training.  It maps a
// few-iteration SIRT volume to an estimate of a converged SIRT volume; SIRT
// then continues from it.  (No counterpart in the reference.)
//
// Tensors are channels-last [D][H][W][C] (D = detector rows, H = iz, W = ix
// of the cylinder's bounding cube), BF16 stored as raw 16-bit patterns.
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x16bf16_1k:
//  * forward:     out[s,co] = sum_{tap,ci} in[s+off(tap), ci] W[tap][co][ci]
//                 (M = 16 positions along W, N = 16 output channels, K = 16
//                 input channels per instruction)
//  * backward data: the same kernel on dZ with W'[t][ci][co] = W[26-t][co][ci]
//  * backward weights: dW[tap][co][ci] = sum_s dZ[s,co] in[s+off(tap),ci]
//                 (M = co, N = ci, K = positions), rows staged in LDS, one
//                 workgroup per (tap, chunk of rows), partials summed in a
//                 fixed order.
// The decoder's concatenation with the 2x nearest-upsampled coarser tensor is
// fused into the convolution's input addressing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "arec_internal.hpp"

namespace arec {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));
typedef uint16_t bf16_t;

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ bf16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16_t)(u >> 16);
}

constexpr int kCvThreads = 256;  // 4 waves x 16 positions = 64 positions per workgroup
constexpr int kWgChunks = 128;   // spatial chunks of the weight-gradient reduction

struct Dims {
    int D, H, W;
};

// ---------------------------------------------------------------------------
// 3x3x3 convolution, forward form, on MFMA.
// Input channels [0, C1) come from in1 (dims d), [C1, C1+C2) from in2 at half
// resolution (nearest upsampling: position s reads in2[s/2]).
// A workgroup computes 64 positions along W (16 per wave) in conv_rows(COUT)
// consecutive H rows: each weight operand loaded for a tap serves every row,
// which divides the weight traffic by the row count and keeps 3 x (rows + NB)
// loads in flight per (tap row, channel block).
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int conv_rows(int cout) { return cout >= 64 ? 2 : 4; }

template <int COUT>
__global__ __launch_bounds__(kCvThreads) void conv3_mfma_kernel(
    const bf16_t *__restrict__ in1, int C1, const bf16_t *__restrict__ in2, int C2,
    const bf16_t *__restrict__ w, const float *__restrict__ bias, bf16_t *__restrict__ out, Dims d,
    int relu) {
    constexpr int NB = COUT / 16, R = conv_rows(COUT);
    const int Ct = C1 + C2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wtiles = (d.W + 63) / 64, hblocks = (d.H + R - 1) / R;
    const int wt = blockIdx.x % wtiles;
    const int h0 = ((blockIdx.x / wtiles) % hblocks) * R;
    const int z = blockIdx.x / (wtiles * hblocks);
    const int pbase = wt * 64 + wave * 16;
    const int pa = pbase + (lane & 15);  // this lane's A-operand position
    const int kq = (lane >> 4) * 4;      // this lane's 4 input channels within a 16-block
    f4v acc[R][NB];
#pragma unroll
    for (int r = 0; r < R; r++)
#pragma unroll
        for (int nb = 0; nb < NB; nb++) acc[r][nb] = f4v{0.f, 0.f, 0.f, 0.f};
    const int H2 = d.H >> 1, W2 = d.W >> 1;
    // Per (dz, dh) row of taps and 16-channel block, the operands of its three
    // dw taps for all rows are loaded first and multiplied after.  The MFMAs
    // of an output run in the order (tap row, channel block, dw tap): with
    // one block (Ct = 16) that is the tap order, otherwise the FP32 sums are
    // reassociated.
    const bf8v zero8 = __builtin_bit_cast(bf8v, u4v{0u, 0u, 0u, 0u});
    for (int row = 0; row < 9; row++) {
        const int dz = row / 3 - 1, dh = row % 3 - 1;
        const int zz = z + dz;
        if (zz < 0 || zz >= d.D) continue;  // uniform
        // 32-channel blocks on v_mfma_f32_16x16x32_bf16: lane l supplies
        // channels 8 (l / 16) .. + 7 of the block (16 bytes per operand load)
        int cb = 0;
        for (; cb + 32 <= Ct; cb += 32) {
            const int ci = cb + (lane >> 4) * 8;  // 8 channels within one 16-group: in1 or in2
            bf8v a[3][R], b[3][NB];
#pragma unroll
            for (int t = 0; t < 3; t++) {
                const int pp = pa + t - 1;
                const bool pv = pp >= 0 && pp < d.W;
                const bf16_t *wt_tap = w + (size_t)(row * 3 + t) * COUT * Ct;
#pragma unroll
                for (int nb = 0; nb < NB; nb++) {
                    const int co = nb * 16 + (lane & 15);
                    b[t][nb] = *reinterpret_cast<const bf8v *>(wt_tap + (size_t)co * Ct + ci);
                }
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int hh = h0 + r + dh;
                    a[t][r] = zero8;
                    if (pv && hh >= 0 && hh < d.H) {
                        const bf16_t *src;
                        if (ci < C1)
                            src = in1 + (((size_t)zz * d.H + hh) * d.W + pp) * C1 + ci;
                        else
                            src = in2 + (((size_t)(zz >> 1) * H2 + (hh >> 1)) * W2 + (pp >> 1)) * C2 + (ci - C1);
                        a[t][r] = *reinterpret_cast<const bf8v *>(src);
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int nb = 0; nb < NB; nb++)
                        acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t][r], b[t][nb], acc[r][nb], 0, 0, 0);
        }
        // a remaining 16-channel block on v_mfma_f32_16x16x16bf16_1k
        for (; cb < Ct; cb += 16) {
            const int ci = cb + kq;  // the block lies in in1 or in in2 (C1 is a multiple of 16)
            s4v a[3][R], b[3][NB];
#pragma unroll
            for (int t = 0; t < 3; t++) {
                const int pp = pa + t - 1;
                const bool pv = pp >= 0 && pp < d.W;
                const bf16_t *wt_tap = w + (size_t)(row * 3 + t) * COUT * Ct;
#pragma unroll
                for (int nb = 0; nb < NB; nb++) {
                    const int co = nb * 16 + (lane & 15);
                    b[t][nb] = *reinterpret_cast<const s4v *>(wt_tap + (size_t)co * Ct + ci);
                }
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int hh = h0 + r + dh;
                    a[t][r] = s4v{0, 0, 0, 0};
                    if (pv && hh >= 0 && hh < d.H) {
                        const bf16_t *src;
                        if (ci < C1)
                            src = in1 + (((size_t)zz * d.H + hh) * d.W + pp) * C1 + ci;
                        else
                            src = in2 + (((size_t)(zz >> 1) * H2 + (hh >> 1)) * W2 + (pp >> 1)) * C2 + (ci - C1);
                        a[t][r] = *reinterpret_cast<const s4v *>(src);
                    }
                }
            }
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int nb = 0; nb < NB; nb++)
                        acc[r][nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t][r], b[t][nb], acc[r][nb], 0, 0, 0);
        }
    }
    // lane holds rows 4*(lane/16)+i (positions) of column lane%16 (channel)
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int h = h0 + r;
        if (h >= d.H) continue;
#pragma unroll
        for (int nb = 0; nb < NB; nb++) {
            const int co = nb * 16 + (lane & 15);
            const float bv = bias ? bias[co] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int p = pbase + (lane >> 4) * 4 + i;
                if (p < d.W) {
                    float v = acc[r][nb][i] + bv;
                    if (relu) v = fmaxf(v, 0.f);
                    out[(((size_t)z * d.H + h) * d.W + p) * COUT + co] = f2bf(v);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The full-resolution 16-output convolutions (16 skip channels, plus 32
// nearest-upsampled channels in the decoder), with the input staged in LDS.
// The generic kernel above loads every A operand from global memory: each
// input value is fetched for 27 taps x several workgroups through the
// vector memory path, which bounds it far below the MFMA rate.  Here every
// A operand is read from LDS, where the input tile was copied with 16-byte
// loads (zeros outside the volume).  A staged slice is 6 H rows x 66 W
// positions (64 + halo) of the 16 skip channels, and 4 x 34 half-resolution
// positions of the 32 upsampled ones.  Channel quads are the slower LDS
// index, so the 16 lanes of a quad read 16 consecutive 8-byte words.  The
// MFMAs of an output run in the order (tap row, dw tap, channel block).  The
// upsampled part needs even D, H and W: then every position outside the
// volume maps to a half-resolution position outside it too, staged as zero.
// ---------------------------------------------------------------------------
constexpr int kC16Rows = 4;                 // H rows per workgroup (conv_rows(16))
constexpr int kC16Pos = 66;                 // staged W positions: 64 + halo
constexpr int kC16Hs = kC16Rows + 2;        // staged H rows
constexpr int kC16Pos2 = 34;                // half-resolution W positions
constexpr int kC16Hs2 = kC16Rows / 2 + 2;   // half-resolution H rows

// The kernel walks D: a workgroup owns 64 positions along W x 4 H rows and
// a run of kC16Zrun consecutive D slices.  Its LDS holds a ring of three
// staged input slices (the 16 skip channels) and, with upsampled channels, a
// ring of two half-resolution slices; per output slice one new input slice
// (and every other slice a new half-resolution one) is loaded -- issued into
// registers before the slice's MFMAs and stored into the ring after them --
// so each input slice row is fetched once per run instead of three times,
// and the fetch overlaps the arithmetic.  Without upsampled channels the 27
// taps' weights stay in registers for the whole run.
constexpr int kC16Zrun = 16;

template <int C2>
__global__ __launch_bounds__(kCvThreads) void conv3_c16z_kernel(const bf16_t *__restrict__ in1,
                                                                  const bf16_t *__restrict__ in2,
                                                                  const bf16_t *__restrict__ w,
                                                                  const float *__restrict__ bias,
                                                                  bf16_t *__restrict__ out, Dims d, int relu) {
    static_assert(C2 == 0 || C2 == 32, "upsampled channels: none or 32");
    constexpr int Ct = 16 + C2;
    constexpr int kSlice = kC16Hs * kC16Pos * 2;                  // 16-byte halves per staged slice
    constexpr int kPer = (kSlice + kCvThreads - 1) / kCvThreads;  // per thread
    constexpr int kSlice2 = kC16Hs2 * kC16Pos2 * (C2 / 8);        // 16-byte quarters per half-resolution slice
    constexpr int kPer2 = C2 ? (kSlice2 + kCvThreads - 1) / kCvThreads : 1;
    __shared__ uint2 s1[3][kC16Hs][4][kC16Pos];
    // half-resolution ring (the upsampled channels): slot of half slice zh is zh & 1
    __shared__ uint2 s2[C2 ? 2 : 1][kC16Hs2][C2 ? C2 / 4 : 1][kC16Pos2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wtiles = (d.W + 63) / 64, hblocks = (d.H + kC16Rows - 1) / kC16Rows;
    const int wt = blockIdx.x % wtiles;
    const int h0 = ((blockIdx.x / wtiles) % hblocks) * kC16Rows;
    const int zA = (blockIdx.x / (wtiles * hblocks)) * kC16Zrun, zE = min(zA + kC16Zrun, d.D);
    const int w0 = wt * 64;
    const int p = lane & 15, q = lane >> 4;

    uint4 pre[kPer];
    auto fetch = [&](const int zz) __attribute__((always_inline)) {  // slice zz into registers
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int idx = tid + k * kCvThreads;
            const int half = idx & 1, pi = idx >> 1;
            const int ws = pi % kC16Pos, hs = pi / kC16Pos;
            const int hh = h0 - 1 + hs, ww = w0 - 1 + ws;
            pre[k] = make_uint4(0u, 0u, 0u, 0u);
            if (idx < kSlice && zz >= 0 && zz < d.D && hh >= 0 && hh < d.H && ww >= 0 && ww < d.W)
                pre[k] = *reinterpret_cast<const uint4 *>(in1 + (((size_t)zz * d.H + hh) * d.W + ww) * 16 + half * 8);
        }
    };
    auto put = [&](const int slot) __attribute__((always_inline)) {  // registers into ring slot
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int idx = tid + k * kCvThreads;
            if (idx < kSlice) {
                const int half = idx & 1, pi = idx >> 1;
                const int ws = pi % kC16Pos, hs = pi / kC16Pos;
                s1[slot][hs][2 * half][ws] = make_uint2(pre[k].x, pre[k].y);
                s1[slot][hs][2 * half + 1][ws] = make_uint2(pre[k].z, pre[k].w);
            }
        }
    };
    // half-resolution slice zh (upsampled channels, 8 per 16-byte quarter)
    const int D2 = d.D >> 1, H2 = d.H >> 1, W2 = d.W >> 1;
    const int hb2 = (h0 - 1) >> 1, wb2 = (w0 - 1) >> 1;  // arithmetic shifts
    uint4 pre2[kPer2];
    auto fetch2 = [&](const int zh) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < kPer2; k++) {
            const int idx = tid + k * kCvThreads;
            const int qq = idx % (C2 / 8 ? C2 / 8 : 1), pi = idx / (C2 / 8 ? C2 / 8 : 1);
            const int ws = pi % kC16Pos2, hs = pi / kC16Pos2;
            const int hh = hb2 + hs, ww = wb2 + ws;
            pre2[k] = make_uint4(0u, 0u, 0u, 0u);
            if (C2 && idx < kSlice2 && zh >= 0 && zh < D2 && hh >= 0 && hh < H2 && ww >= 0 && ww < W2)
                pre2[k] = *reinterpret_cast<const uint4 *>(in2 + (((size_t)zh * H2 + hh) * W2 + ww) * C2 + qq * 8);
        }
    };
    auto put2 = [&](const int zh) __attribute__((always_inline)) {
        if constexpr (C2 > 0) {
            const int slot = zh & 1;
#pragma unroll
            for (int k = 0; k < kPer2; k++) {
                const int idx = tid + k * kCvThreads;
                if (idx < kSlice2) {
                    const int qq = idx % (C2 / 8), pi = idx / (C2 / 8);
                    const int ws = pi % kC16Pos2, hs = pi / kC16Pos2;
                    s2[slot][hs][2 * qq][ws] = make_uint2(pre2[k].x, pre2[k].y);
                    s2[slot][hs][2 * qq + 1][ws] = make_uint2(pre2[k].z, pre2[k].w);
                }
            }
        }
    };
    // C2 = 0: the 27 taps' B operands (channel quad q of output channel p)
    // stay in registers; with upsampled channels they are loaded per tap
    s4v b[C2 ? 1 : 27];
    if constexpr (C2 == 0) {
#pragma unroll
        for (int tap = 0; tap < 27; tap++) b[tap] = *reinterpret_cast<const s4v *>(w + (size_t)tap * 256 + p * 16 + q * 4);
    }
    const float bv = bias ? bias[lane & 15] : 0.f;

    // ring slot of slice zz: (zz - zA + 1) mod 3; of half slice zh: zh & 1.
    // Output slice z reads half slices (z - 1) >> 1 and (z + 1) >> 1; the
    // next one, (z + 2) >> 1, is new when z is even.
    fetch(zA - 1);
    put(0);
    fetch(zA);
    put(1);
    if constexpr (C2 > 0) {
        fetch2((zA - 1) >> 1);
        put2((zA - 1) >> 1);
        if (((zA + 1) >> 1) != ((zA - 1) >> 1)) {
            fetch2((zA + 1) >> 1);
            put2((zA + 1) >> 1);
        }
    }
    fetch(zA + 1);
    bool have2 = false;  // pre2 holds half slice (z + 2) >> 1
    for (int z = zA; z < zE; z++) {
        const int k = z - zA;
        put((k + 2) % 3);  // slice z + 1
        if constexpr (C2 > 0) {
            if (have2) put2((z + 1) >> 1);  // its slot held (z - 3) >> 1, read last by slice z - 1
        }
        __syncthreads();
        fetch(z + 2);      // in flight during the MFMAs (zeros past the volume)
        if constexpr (C2 > 0) {
            have2 = !(z & 1);  // (z + 2) >> 1 differs from (z + 1) >> 1 for even z
            if (have2) fetch2((z + 2) >> 1);
        }
        f4v acc[kC16Rows];
#pragma unroll
        for (int r = 0; r < kC16Rows; r++) acc[r] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int dz = 0; dz < 3; dz++) {
            const int zz = z - 1 + dz;
            if (zz < 0 || zz >= d.D) continue;  // uniform
            const int slot = (k + dz) % 3, slot2 = (zz >> 1) & 1;
            auto tap_row = [&](const int dh) __attribute__((always_inline)) {
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    const int tap = (dz * 3 + dh) * 3 + t;
                    const int ws = wave * 16 + p + t;
                    s4v bt = b[0], bu[C2 ? C2 / 16 : 1];
                    if constexpr (C2 == 0) {
                        bt = b[tap];
                    } else {
                        const bf16_t *wtap = w + (size_t)tap * 16 * Ct + (size_t)p * Ct + q * 4;
                        bt = *reinterpret_cast<const s4v *>(wtap);
#pragma unroll
                        for (int j = 0; j < C2 / 16; j++) bu[j] = *reinterpret_cast<const s4v *>(wtap + 16 + j * 16);
                    }
#pragma unroll
                    for (int r = 0; r < kC16Rows; r++) {
                        const s4v a1 = __builtin_bit_cast(s4v, s1[slot][r + dh][q][ws]);
                        acc[r] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, bt, acc[r], 0, 0, 0);
                        if constexpr (C2 > 0) {
                            const int hs2 = ((h0 + r + dh - 1) >> 1) - hb2;
                            const int ws2 = ((w0 + ws - 1) >> 1) - wb2;
#pragma unroll
                            for (int j = 0; j < C2 / 16; j++) {
                                const s4v a2 = __builtin_bit_cast(s4v, s2[slot2][hs2][j * 4 + q][ws2]);
                                acc[r] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, bu[j], acc[r], 0, 0, 0);
                            }
                        }
                    }
                }
            };
            if constexpr (C2 == 0) {
                tap_row(0);
                tap_row(1);
                tap_row(2);
            } else {
                // one tap row at a time (unrolled, the operands of all nine
                // taps would need over 256 VGPRs)
#pragma unroll 1
                for (int dh = 0; dh < 3; dh++) tap_row(dh);
            }
        }
#pragma unroll
        for (int r = 0; r < kC16Rows; r++) {
            const int h = h0 + r;
            if (h >= d.H) continue;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int pw = w0 + wave * 16 + (lane >> 4) * 4 + i;
                if (pw < d.W) {
                    float v = acc[r][i] + bv;
                    if (relu) v = fmaxf(v, 0.f);
                    out[(((size_t)z * d.H + h) * d.W + pw) * 16 + (lane & 15)] = f2bf(v);
                }
            }
        }
        __syncthreads();  // every wave is done with slice z - 1 before its slot takes z + 2
    }
}

// W'[t][ci][co] = W[26-t][co][ci]: the backward-data convolution's weights.
__global__ void weight_transpose_kernel(const bf16_t *__restrict__ w, bf16_t *__restrict__ wt, int Cout,
                                        int Cin) {
    int n = 27 * Cout * Cin;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        int t = e / (Cout * Cin), rem = e % (Cout * Cin);
        int ci = rem / Cout, co = rem % Cout;  // output layout [t][ci][co]
        wt[e] = w[((size_t)(26 - t) * Cout + co) * Cin + ci];
    }
}

// ---------------------------------------------------------------------------
// Weight gradient on MFMA: dW[tap][co][ci] partial sums over a chunk of rows.
// Grid: 27 taps x kWgChunks; each workgroup stages, per row, the dZ row
// [co][W] and the shifted input row [ci][W+2] in LDS (transposed so that an
// operand's 4 positions are contiguous).
// ---------------------------------------------------------------------------
constexpr int kWgMaxW = 128;
constexpr int kWgMaxC = 96;

__global__ __launch_bounds__(kCvThreads) void conv3_wgrad_kernel(
    const bf16_t *__restrict__ in1, int C1, const bf16_t *__restrict__ in2, int C2,
    const bf16_t *__restrict__ dz, int Cout, Dims d, float *__restrict__ part) {
    __shared__ bf16_t s_dz[64][kWgMaxW];
    __shared__ bf16_t s_in[kWgMaxC][kWgMaxW + 8];
    const int Ct = C1 + C2;
    const int tap = blockIdx.x % 27, chunk = blockIdx.x / 27;
    const int dzt = tap / 9 - 1, dht = (tap / 3) % 3 - 1, dwt = tap % 3 - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int mb = Cout / 16, nbk = Ct / 16, ntiles = mb * nbk;
    const int rows = d.D * d.H, rpc = rows / kWgChunks;
    const int H2 = d.H >> 1, W2 = d.W >> 1;
    // tiles of this wave: t = wave, wave+4, ... (at most 6 for 64 x 96)
    f4v acc[6];
#pragma unroll
    for (int k = 0; k < 6; k++) acc[k] = f4v{0.f, 0.f, 0.f, 0.f};
    for (int r = chunk * rpc; r < (chunk + 1) * rpc; r++) {
        const int z = r / d.H, h = r % d.H;
        const int zz = z + dzt, hh = h + dht;
        const bool rowv = zz >= 0 && zz < d.D && hh >= 0 && hh < d.H;
        __syncthreads();
        // stage dZ row transposed: s_dz[co][w]
        for (int e = threadIdx.x; e < Cout * d.W; e += kCvThreads) {
            int wp = e / Cout, co = e % Cout;
            s_dz[co][wp] = dz[(((size_t)z * d.H + h) * d.W + wp) * Cout + co];
        }
        // stage input row (zz, hh) at positions w + dwt, stored at index w+1
        for (int e = threadIdx.x; e < Ct * (d.W + 2); e += kCvThreads) {
            int wq = e / Ct - 1, ci = e % Ct;  // wq in [-1, W]
            bf16_t v = 0;
            if (rowv && wq >= 0 && wq < d.W) {
                if (ci < C1)
                    v = in1[(((size_t)zz * d.H + hh) * d.W + wq) * C1 + ci];
                else
                    v = in2[(((size_t)(zz >> 1) * H2 + (hh >> 1)) * W2 + (wq >> 1)) * C2 + (ci - C1)];
            }
            s_in[ci][wq + 1] = v;
        }
        __syncthreads();
        if (!rowv) continue;
        for (int k0 = 0; k0 < d.W; k0 += 16) {
            const int kq = k0 + (lane >> 4) * 4;
#pragma unroll
            for (int k = 0; k < 6; k++) {
                const int t = wave + 4 * k;
                if (t >= ntiles) break;
                const int mi = t / nbk, ni = t % nbk;
                s4v a, b;
                const int co = mi * 16 + (lane & 15), ci = ni * 16 + (lane & 15);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    a[i] = (short)s_dz[co][kq + i];
                    b[i] = (short)s_in[ci][kq + i + dwt + 1];
                }
                acc[k] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[k], 0, 0, 0);
            }
        }
    }
    // acc[k]: lane holds rows (co) 4*(lane/16)+i of column (ci) lane%16
    float *pp = part + ((size_t)chunk * 27 + tap) * Cout * Ct;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int t = wave + 4 * k;
        if (t >= ntiles) break;
        const int mi = t / nbk, ni = t % nbk;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int co = mi * 16 + (lane >> 4) * 4 + i, ci = ni * 16 + (lane & 15);
            pp[(size_t)co * Ct + ci] = acc[k][i];
        }
    }
}

// dW[e] = sum over chunks of part[chunk][e] (fixed order); n = 27*Cout*Cin
__global__ void reduce_chunks_kernel(const float *__restrict__ part, float *__restrict__ dw, int n,
                                     int nchunks) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int c = 0; c < nchunks; c++) s += part[(size_t)c * n + e];
        dw[e] = s;
    }
}

// db[co] = sum_s dZ[s,co]: one workgroup per channel, fixed order
__global__ void bias_grad_kernel(const bf16_t *__restrict__ dzv, int C, int64_t S, float *__restrict__ db) {
    __shared__ float s[256];
    const int co = blockIdx.x;
    float acc = 0.f;
    for (int64_t i = threadIdx.x; i < S; i += blockDim.x) acc += bf2f(dzv[i * C + co]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[co] = s[0];
}

// ---------------------------------------------------------------------------
// First layer (1 input channel, FP32 input), VALU
// ---------------------------------------------------------------------------
__global__ void conv_first_kernel(const float *__restrict__ x, const float *__restrict__ w,
                                  const float *__restrict__ bias, bf16_t *__restrict__ out, Dims d,
                                  float scale) {
    const int64_t S = (int64_t)d.D * d.H * d.W;
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < S; s += (int64_t)gridDim.x * blockDim.x) {
        int wq = (int)(s % d.W);
        int h = (int)((s / d.W) % d.H);
        int z = (int)(s / ((int64_t)d.W * d.H));
        float v[27];
#pragma unroll
        for (int t = 0; t < 27; t++) {
            int zz = z + t / 9 - 1, hh = h + (t / 3) % 3 - 1, ww = wq + t % 3 - 1;
            v[t] = (zz >= 0 && zz < d.D && hh >= 0 && hh < d.H && ww >= 0 && ww < d.W)
                       ? x[((int64_t)zz * d.H + hh) * d.W + ww] * scale
                       : 0.f;
        }
        for (int co = 0; co < 16; co++) {
            float a = bias[co];
#pragma unroll
            for (int t = 0; t < 27; t++) a = fmaf(w[t * 16 + co], v[t], a);
            out[s * 16 + co] = f2bf(fmaxf(a, 0.f));
        }
    }
}

// dW_first[t][co] partials per block, db_first too (last 16 entries)
__global__ __launch_bounds__(256) void conv_first_wgrad_kernel(const float *__restrict__ x,
                                                              const bf16_t *__restrict__ dzv, Dims d,
                                                              float scale, float *__restrict__ part) {
    __shared__ float s_acc[27 * 16 + 16];
    for (int e = threadIdx.x; e < 27 * 16 + 16; e += blockDim.x) s_acc[e] = 0.f;
    __syncthreads();
    const int64_t S = (int64_t)d.D * d.H * d.W;
    float acc[27 * 16 / 16];  // per thread: one output channel, 27 taps
    const int co = threadIdx.x & 15;
    for (int t = 0; t < 27; t++) acc[t] = 0.f;
    float accb = 0.f;
    for (int64_t s = blockIdx.x * (int64_t)(blockDim.x / 16) + (threadIdx.x >> 4); s < S;
         s += (int64_t)gridDim.x * (blockDim.x / 16)) {
        int wq = (int)(s % d.W);
        int h = (int)((s / d.W) % d.H);
        int z = (int)(s / ((int64_t)d.W * d.H));
        float g = bf2f(dzv[s * 16 + co]);
        accb += g;
#pragma unroll
        for (int t = 0; t < 27; t++) {
            int zz = z + t / 9 - 1, hh = h + (t / 3) % 3 - 1, ww = wq + t % 3 - 1;
            float v = (zz >= 0 && zz < d.D && hh >= 0 && hh < d.H && ww >= 0 && ww < d.W)
                          ? x[((int64_t)zz * d.H + hh) * d.W + ww] * scale
                          : 0.f;
            acc[t] = fmaf(g, v, acc[t]);
        }
    }
    // combine the 16 thread-groups of this block in a fixed order
    for (int grp = 0; grp < (int)blockDim.x / 16; grp++) {
        if ((int)(threadIdx.x >> 4) == grp) {
            for (int t = 0; t < 27; t++) s_acc[t * 16 + co] += acc[t];
            s_acc[27 * 16 + co] += accb;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < 27 * 16 + 16; e += blockDim.x) part[(size_t)blockIdx.x * (27 * 16 + 16) + e] = s_acc[e];
}

// ---------------------------------------------------------------------------
// Last layer: y = x + b + sum_c w[c] T[s,c] (FP32 out), and its backward
// ---------------------------------------------------------------------------
__global__ void conv_last_kernel(const bf16_t *__restrict__ t, const float *__restrict__ w, float b,
                                 const float *__restrict__ xin, float *__restrict__ y, int64_t S,
                                 float inv_scale) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < S; s += (int64_t)gridDim.x * blockDim.x) {
        // the voxel's 16 channels as two 16-byte loads
        const uint4 *tv = reinterpret_cast<const uint4 *>(t + s * 16);
        const uint4 lo = tv[0], hi = tv[1];
        const unsigned u[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        float a = b;
#pragma unroll
        for (int c = 0; c < 16; c++) a = fmaf(w[c], bf2f((bf16_t)((c & 1) ? u[c >> 1] >> 16 : u[c >> 1] & 0xffffu)), a);
        y[s] = xin[s] + a * inv_scale;
    }
}

// dT[s,c] = w[c] dR[s]; partial sums of dW[c] = sum_s T[s,c] dR[s] and db
__global__ __launch_bounds__(256) void conv_last_bwd_kernel(const bf16_t *__restrict__ t,
                                                           const float *__restrict__ w,
                                                           const float *__restrict__ dy, int64_t S,
                                                           float inv_scale, bf16_t *__restrict__ dt,
                                                           float *__restrict__ part) {
    __shared__ float s_acc[256][17];
    float acc[17];
    for (int c = 0; c < 17; c++) acc[c] = 0.f;
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < S; s += (int64_t)gridDim.x * blockDim.x) {
        float g = dy[s] * inv_scale;  // dL/dR
        for (int c = 0; c < 16; c++) {
            acc[c] = fmaf(bf2f(t[s * 16 + c]), g, acc[c]);
            dt[s * 16 + c] = f2bf(w[c] * g);
        }
        acc[16] += g;
    }
    for (int c = 0; c < 17; c++) s_acc[threadIdx.x][c] = acc[c];
    __syncthreads();
    if (threadIdx.x < 17) {
        float s = 0.f;
        for (int i = 0; i < 256; i++) s += s_acc[i][threadIdx.x];
        part[(size_t)blockIdx.x * 17 + threadIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------
// Elementwise and resampling kernels
// ---------------------------------------------------------------------------
#define GS(i, n) for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// dZ = dOut * (act > 0)
__global__ void relu_mask_kernel(const bf16_t *__restrict__ dout, const bf16_t *__restrict__ act,
                                 bf16_t *__restrict__ dzv, int64_t n) {
    GS(i, n) dzv[i] = bf2f(act[i]) > 0.f ? dout[i] : (bf16_t)0;
}

// 2x2x2 average pooling, C a multiple of 8: a thread per (output position,
// 8 channels) with 16-byte loads and stores, one workgroup per 256 such
// items of one output (D, H) row, coordinates without 64-bit divisions.
// Each channel's 8 values are summed in (dz, dy, dx) order.
__global__ __launch_bounds__(256) void avgpool8_kernel(const bf16_t *__restrict__ in, bf16_t *__restrict__ out,
                                                       Dims d, int C) {
    const int H2 = d.H / 2, W2 = d.W / 2, G = C / 8;
    const int per_row = W2 * G, rb = (per_row + 255) / 256;
    const int row = (int)(blockIdx.x / (unsigned)rb);
    const int t = (int)(blockIdx.x - (unsigned)row * rb) * 256 + (int)threadIdx.x;
    if (t >= per_row) return;
    const int x2 = t / G, g = t - x2 * G;
    const int z2 = row / H2, y2 = row - z2 * H2;
    float a[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int z = 2 * z2 + (k >> 2), y = 2 * y2 + ((k >> 1) & 1), x = 2 * x2 + (k & 1);
        const uint4 v = *reinterpret_cast<const uint4 *>(in + (((int64_t)z * d.H + y) * d.W + x) * C + g * 8);
        const unsigned u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            a[2 * j] += bf2f((bf16_t)(u[j] & 0xffffu));
            a[2 * j + 1] += bf2f((bf16_t)(u[j] >> 16));
        }
    }
    unsigned pk[4];
#pragma unroll
    for (int j = 0; j < 4; j++)
        pk[j] = (unsigned)f2bf(0.125f * a[2 * j]) | ((unsigned)f2bf(0.125f * a[2 * j + 1]) << 16);
    *reinterpret_cast<uint4 *>(out + (((int64_t)z2 * H2 + y2) * W2 + x2) * C + g * 8) =
        make_uint4(pk[0], pk[1], pk[2], pk[3]);
}

// dS = dpart + avgpool backward of dP (each fine voxel gets dP/8)
__global__ void avgpool_bwd_add_kernel(const bf16_t *__restrict__ dp, const bf16_t *__restrict__ dpart,
                                       int cpart, int coff, bf16_t *__restrict__ ds, Dims d, int C) {
    const int64_t n = (int64_t)d.D * d.H * d.W * C;
    const int H2 = d.H / 2, W2 = d.W / 2;
    GS(e, n) {
        int c = (int)(e % C);
        int64_t s = e / C;
        int x = (int)(s % d.W), y = (int)((s / d.W) % d.H), z = (int)(s / ((int64_t)d.W * d.H));
        float v = 0.125f * bf2f(dp[(((int64_t)(z / 2) * H2 + y / 2) * W2 + x / 2) * C + c]);
        if (dpart) v += bf2f(dpart[s * cpart + coff + c]);
        ds[e] = f2bf(v);
    }
}

// gradient of the upsampled half of a concatenated input: sum of the 8 fine
// voxels' channels [c0, c0+C2) of dIn (Ct channels) -> coarse dLow (C2)
__global__ void upsample_bwd_kernel(const bf16_t *__restrict__ din, int Ct, int c0, int C2,
                                    bf16_t *__restrict__ dlow, Dims d) {
    const int D2 = d.D / 2, H2 = d.H / 2, W2 = d.W / 2;
    const int64_t n = (int64_t)D2 * H2 * W2 * C2;
    GS(e, n) {
        int c = (int)(e % C2);
        int64_t s = e / C2;
        int x2 = (int)(s % W2), y2 = (int)((s / W2) % H2), z2 = (int)(s / ((int64_t)W2 * H2));
        float a = 0.f;
        for (int k = 0; k < 8; k++) {
            int z = 2 * z2 + (k >> 2), y = 2 * y2 + ((k >> 1) & 1), x = 2 * x2 + (k & 1);
            a += bf2f(din[(((int64_t)z * d.H + y) * d.W + x) * Ct + c0 + c]);
        }
        dlow[e] = f2bf(a);
    }
}

// loss partials: sum (y-t)^2 per block (FP64); dY = 2(y-t)/n
__global__ __launch_bounds__(256) void mse_kernel(const float *__restrict__ y, const float *__restrict__ t,
                                                 float *__restrict__ dy, int64_t n, double *__restrict__ part) {
    __shared__ double s[256];
    double a = 0.0;
    const float k = 2.0f / (float)n;
    GS(i, n) {
        float e = y[i] - t[i];
        a += (double)e * e;
        dy[i] = k * e;
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// Adam on FP32 master parameters; BF16 copies for the MFMA convolutions.
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, bf16_t *__restrict__ pb, int64_t n, float lr, float b1,
                            float b2, float c1, float c2, float eps) {
    GS(i, n) {
        float gi = g[i];
        float mi = b1 * m[i] + (1.f - b1) * gi;
        float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        float pi = p[i] - upd;
        p[i] = pi;
        if (pb) pb[i] = f2bf(pi);
    }
}

__global__ void to_bf16_kernel(const float *__restrict__ a, bf16_t *__restrict__ b, int64_t n) {
    GS(i, n) b[i] = f2bf(a[i]);
}

// cylinder [nrays][ny] <-> cube [ny][n][n] (outside the mask: 0): element
// j * ny + y of the cylinder volume <-> cube voxel (y, iz, ix) of ray j.
// Cylinder to cube through 64-ray x 64-row tiles in LDS: each ray's rows
// are read as one contiguous run and each cube row is written by
// neighbouring rays (a direct element-wise copy scatters one of the two
// sides: 2.2 ms at 512^3).  256 threads; tiles with rays fastest.
constexpr int kCubeTile = 64;
__global__ __launch_bounds__(256) void cyl_to_cube_kernel(const float *__restrict__ x, const int *__restrict__ cord,
                                                          int nrays, int ny, int n, float *__restrict__ cube) {
    __shared__ float t[kCubeTile][kCubeTile + 1];
    const int nrt = (nrays + kCubeTile - 1) / kCubeTile, nyt = (ny + kCubeTile - 1) / kCubeTile;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int tile = blockIdx.x; tile < nrt * nyt; tile += gridDim.x) {
        const int j0 = (tile % nrt) * kCubeTile, y0 = (tile / nrt) * kCubeTile;
        for (int r = w; r < kCubeTile; r += 4) {
            const int j = j0 + r, y = y0 + lane;
            t[r][lane] = (j < nrays && y < ny) ? x[(int64_t)j * ny + y] : 0.0f;
        }
        __syncthreads();
        const int j = j0 + lane;
        if (j < nrays) {
            const int64_t c0 = (int64_t)cord[2 * j + 1] * n + cord[2 * j];
            for (int yy = w; yy < kCubeTile && y0 + yy < ny; yy += 4)
                cube[(int64_t)(y0 + yy) * n * n + c0] = t[lane][yy];
        }
        __syncthreads();
    }
}

// Cube to cylinder through the same tiles: cube rows read by neighbouring
// rays, each ray's rows written as one contiguous run.
__global__ __launch_bounds__(256) void cube_to_cyl_kernel(const float *__restrict__ cube, const int *__restrict__ cord,
                                                          int nrays, int ny, int n, float *__restrict__ x) {
    __shared__ float t[kCubeTile][kCubeTile + 1];
    const int nrt = (nrays + kCubeTile - 1) / kCubeTile, nyt = (ny + kCubeTile - 1) / kCubeTile;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int tile = blockIdx.x; tile < nrt * nyt; tile += gridDim.x) {
        const int j0 = (tile % nrt) * kCubeTile, y0 = (tile / nrt) * kCubeTile;
        const int j = j0 + lane;
        if (j < nrays) {
            const int64_t c0 = (int64_t)cord[2 * j + 1] * n + cord[2 * j];
            for (int yy = w; yy < kCubeTile && y0 + yy < ny; yy += 4) t[lane][yy] = cube[(int64_t)(y0 + yy) * n * n + c0];
        }
        __syncthreads();
        for (int r = w; r < kCubeTile; r += 4) {
            const int jr = j0 + r, y = y0 + lane;
            if (jr < nrays && y < ny) x[(int64_t)jr * ny + y] = t[r][lane];
        }
        __syncthreads();
    }
}

static dim3 ew(int64_t n) { return dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384))); }

// ---------------------------------------------------------------------------
// Host orchestration
// ---------------------------------------------------------------------------
static int launch_conv3(const bf16_t *in1, int C1, const bf16_t *in2, int C2, const bf16_t *w,
                        const float *bias, bf16_t *out, int Cout, Dims d, int relu, hipStream_t st) {
    if (C1 % 16 || C2 % 16) {
        set_error("conv3: channel counts must be multiples of 16");
        return AREC_EINVAL;
    }
    const int R = conv_rows(Cout);
    dim3 grid((unsigned)(((d.W + 63) / 64) * ((d.H + R - 1) / R) * d.D));
    // 16 full-resolution channels to 16 outputs (with 32 upsampled channels
    // only where D, H and W are even): the LDS-staged kernel
    if (Cout == 16 && C1 == 16 && (C2 == 0 || (C2 == 32 && !(d.D & 1) && !(d.H & 1) && !(d.W & 1)))) {
        dim3 gz((unsigned)(((d.W + 63) / 64) * ((d.H + kC16Rows - 1) / kC16Rows) * ((d.D + kC16Zrun - 1) / kC16Zrun)));
        if (C2 == 0)
            conv3_c16z_kernel<0><<<gz, kCvThreads, 0, st>>>(in1, nullptr, w, bias, out, d, relu);
        else
            conv3_c16z_kernel<32><<<gz, kCvThreads, 0, st>>>(in1, in2, w, bias, out, d, relu);
        AREC_HIP(hipGetLastError());
        return 0;
    }
    switch (Cout) {
    case 16: conv3_mfma_kernel<16><<<grid, kCvThreads, 0, st>>>(in1, C1, in2, C2, w, bias, out, d, relu); break;
    case 32: conv3_mfma_kernel<32><<<grid, kCvThreads, 0, st>>>(in1, C1, in2, C2, w, bias, out, d, relu); break;
    case 48: conv3_mfma_kernel<48><<<grid, kCvThreads, 0, st>>>(in1, C1, in2, C2, w, bias, out, d, relu); break;
    case 64: conv3_mfma_kernel<64><<<grid, kCvThreads, 0, st>>>(in1, C1, in2, C2, w, bias, out, d, relu); break;
    case 96: conv3_mfma_kernel<96><<<grid, kCvThreads, 0, st>>>(in1, C1, in2, C2, w, bias, out, d, relu); break;
    default: set_error("conv3: unsupported channel count"); return AREC_EINVAL;
    }
    AREC_HIP(hipGetLastError());
    return 0;
}

struct Layer {
    int cin1, cin2, cout;  // input split (skip, upsampled) and outputs
    int level;             // 0 full, 1 half, 2 quarter resolution
    int64_t woff, boff;    // offsets into the parameter vector
};

struct UNet {
    arec_ctx *ctx = nullptr;
    Dims d0{0, 0, 0};
    std::vector<Layer> L;  // 10 MFMA convolutions
    int64_t nparam = 0, wfirst = 0, bfirst = 0, wlast = 0, blast = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // FP32 params, grads, Adam moments
    bf16_t *pb = nullptr;                                          // BF16 copy of the parameters
    bf16_t *wt = nullptr;                                          // transposed weights (scratch)
    float *part = nullptr;                                         // reduction partials
    double *lpart = nullptr;
    // activations (BF16), per resolution level
    std::vector<bf16_t *> act;
    // gradient buffers (training only): three per resolution level plus the
    // concatenated-input gradients of the two decoder levels
    bf16_t *g0[3] = {nullptr, nullptr, nullptr};  // 16 ch, full resolution
    bf16_t *g1[3] = {nullptr, nullptr, nullptr};  // 32 ch, half
    bf16_t *g2[3] = {nullptr, nullptr, nullptr};  // 64 ch, quarter
    bf16_t *gin7 = nullptr;                       // 48 ch, full (skip 0..15 | up 16..47)
    bf16_t *gin5 = nullptr;                       // 96 ch, half (skip 0..31 | up 32..95)
    float *dy = nullptr;
    int64_t step = 0;
    float scale = 1.f;  // input scaling: network sees x * scale
    std::vector<void *> allocs;
    ~UNet() {
        for (void *q : allocs) hipFree(q);
    }
    template <typename T>
    int alloc(T **q, int64_t n) {
        void *r = nullptr;
        AREC_HIP(hipMalloc(&r, sizeof(T) * std::max<int64_t>(1, n)));
        allocs.push_back(r);
        *q = (T *)r;
        return 0;
    }
    Dims dims(int level) const { return Dims{d0.D >> level, d0.H >> level, d0.W >> level}; }
    int64_t vox(int level) const {
        Dims d = dims(level);
        return (int64_t)d.D * d.H * d.W;
    }
};

// activation slots
enum { T1A, S1, P1, T2A, S2, P2, T3A, B3, T4A, T4B, T5A, T5B, NACT };

static void unet_layers(UNet &n) {
    // {cin1, cin2, cout, level}
    const int spec[10][4] = {{16, 0, 16, 0},  {16, 0, 32, 1}, {32, 0, 32, 1}, {32, 0, 64, 2},
                             {64, 0, 64, 2},  {32, 64, 32, 1}, {32, 0, 32, 1}, {16, 32, 16, 0},
                             {16, 0, 16, 0},  {0, 0, 0, 0}};
    int64_t off = 0;
    n.wfirst = off;
    off += 27 * 16;
    n.bfirst = off;
    off += 16;
    for (int i = 0; i < 9; i++) {
        Layer l{spec[i][0], spec[i][1], spec[i][2], spec[i][3], 0, 0};
        l.woff = off;
        off += 27 * (int64_t)l.cout * (l.cin1 + l.cin2);
        l.boff = off;
        off += l.cout;
        n.L.push_back(l);
    }
    n.wlast = off;
    off += 16;
    n.blast = off;
    off += 1;
    n.nparam = off;
}

static int unet_alloc(UNet &n, Dims d, bool training) {
    n.d0 = d;
    unet_layers(n);
    int rc;
    if ((rc = n.alloc(&n.p, n.nparam)) || (rc = n.alloc(&n.pb, n.nparam)) || (rc = n.alloc(&n.g, n.nparam)) ||
        (rc = n.alloc(&n.m, n.nparam)) || (rc = n.alloc(&n.v, n.nparam)))
        return rc;
    const int lvl[NACT] = {0, 0, 1, 1, 1, 2, 2, 2, 1, 1, 0, 0};
    const int ch[NACT] = {16, 16, 16, 32, 32, 32, 64, 64, 32, 32, 16, 16};
    n.act.assign(NACT, nullptr);
    for (int i = 0; i < NACT; i++)
        if ((rc = n.alloc(&n.act[i], n.vox(lvl[i]) * ch[i]))) return rc;
    if (training) {
        for (int k = 0; k < 3; k++)
            if ((rc = n.alloc(&n.g0[k], n.vox(0) * 16)) || (rc = n.alloc(&n.g1[k], n.vox(1) * 32)) ||
                (rc = n.alloc(&n.g2[k], n.vox(2) * 64)))
                return rc;
        if ((rc = n.alloc(&n.gin7, n.vox(0) * 48)) || (rc = n.alloc(&n.gin5, n.vox(1) * 96))) return rc;
    }
    if ((rc = n.alloc(&n.wt, 27 * 96 * 64)) || (rc = n.alloc(&n.dy, n.vox(0))) ||
        (rc = n.alloc(&n.lpart, 16384)))
        return rc;
    int64_t npart = training ? (int64_t)kWgChunks * 27 * 64 * 96 : 16;
    npart = std::max<int64_t>(npart, 4096 * (27 * 16 + 16));
    if ((rc = n.alloc(&n.part, npart))) return rc;
    (void)training;
    return 0;
}

// He-style initialisation from a seeded generator; the last layer starts at
// zero so the untrained network returns its input.
static void unet_init_params(UNet &n, uint64_t seed, std::vector<float> &h) {
    h.assign(n.nparam, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> N01(0.f, 1.f);
    for (int t = 0; t < 27 * 16; t++) h[n.wfirst + t] = N01(rng) * std::sqrt(2.f / 27.f);
    for (auto &l : n.L) {
        int cin = l.cin1 + l.cin2;
        float sd = std::sqrt(2.f / (27.f * cin));
        for (int64_t e = 0; e < 27 * (int64_t)l.cout * cin; e++) h[l.woff + e] = N01(rng) * sd;
    }
}

static int unet_upload(UNet &n, const std::vector<float> &h, hipStream_t st) {
    AREC_HIP(hipMemcpyAsync(n.p, h.data(), sizeof(float) * n.nparam, hipMemcpyHostToDevice, st));
    to_bf16_kernel<<<ew(n.nparam), 256, 0, st>>>(n.p, n.pb, n.nparam);
    AREC_HIP(hipMemsetAsync(n.m, 0, sizeof(float) * n.nparam, st));
    AREC_HIP(hipMemsetAsync(n.v, 0, sizeof(float) * n.nparam, st));
    AREC_HIP(hipGetLastError());
    return 0;
}

// forward: x (FP32 cube [D][H][W]) -> y (FP32) = x + net(x * scale) / scale
static int unet_forward(UNet &n, const float *x, float *y, hipStream_t st) {
    int rc;
    Dims d0 = n.dims(0), d1 = n.dims(1), d2 = n.dims(2);
    conv_first_kernel<<<ew(n.vox(0)), 256, 0, st>>>(x, n.p + n.wfirst, n.p + n.bfirst, n.act[T1A], d0, n.scale);
    auto W = [&](int i) { return n.pb + n.L[i].woff; };
    auto B = [&](int i) { return n.p + n.L[i].boff; };
    if ((rc = launch_conv3(n.act[T1A], 16, nullptr, 0, W(0), B(0), n.act[S1], 16, d0, 1, st))) return rc;
    avgpool8_kernel<<<dim3((unsigned)((int64_t)d1.D * d1.H * ((d1.W * 2 + 255) / 256))), 256, 0, st>>>(
        n.act[S1], n.act[P1], d0, 16);
    if ((rc = launch_conv3(n.act[P1], 16, nullptr, 0, W(1), B(1), n.act[T2A], 32, d1, 1, st))) return rc;
    if ((rc = launch_conv3(n.act[T2A], 32, nullptr, 0, W(2), B(2), n.act[S2], 32, d1, 1, st))) return rc;
    avgpool8_kernel<<<dim3((unsigned)((int64_t)d2.D * d2.H * ((d2.W * 4 + 255) / 256))), 256, 0, st>>>(
        n.act[S2], n.act[P2], d1, 32);
    if ((rc = launch_conv3(n.act[P2], 32, nullptr, 0, W(3), B(3), n.act[T3A], 64, d2, 1, st))) return rc;
    if ((rc = launch_conv3(n.act[T3A], 64, nullptr, 0, W(4), B(4), n.act[B3], 64, d2, 1, st))) return rc;
    if ((rc = launch_conv3(n.act[S2], 32, n.act[B3], 64, W(5), B(5), n.act[T4A], 32, d1, 1, st))) return rc;
    if ((rc = launch_conv3(n.act[T4A], 32, nullptr, 0, W(6), B(6), n.act[T4B], 32, d1, 1, st))) return rc;
    if ((rc = launch_conv3(n.act[S1], 16, n.act[T4B], 32, W(7), B(7), n.act[T5A], 16, d0, 1, st))) return rc;
    if ((rc = launch_conv3(n.act[T5A], 16, nullptr, 0, W(8), B(8), n.act[T5B], 16, d0, 1, st))) return rc;
    float blast = 0.f;
    AREC_HIP(hipMemcpyAsync(&blast, n.p + n.blast, sizeof(float), hipMemcpyDeviceToHost, st));
    AREC_HIP(hipStreamSynchronize(st));
    conv_last_kernel<<<ew(n.vox(0)), 256, 0, st>>>(n.act[T5B], n.p + n.wlast, blast, x, y, n.vox(0), 1.f / n.scale);
    AREC_HIP(hipGetLastError());
    return 0;
}

// gradient of one MFMA convolution layer i given dOut (BF16, its output's
// gradient): dW, db into n.g; dIn (if din != nullptr) with Cin1+Cin2 channels.
static int conv_backward(UNet &n, int i, const bf16_t *in1, const bf16_t *in2, const bf16_t *out,
                         const bf16_t *dout, bf16_t *dz, bf16_t *din, Dims d, hipStream_t st) {
    const Layer &l = n.L[i];
    const int Ct = l.cin1 + l.cin2;
    const int64_t S = (int64_t)d.D * d.H * d.W;
    relu_mask_kernel<<<ew(S * l.cout), 256, 0, st>>>(dout, out, dz, S * l.cout);
    conv3_wgrad_kernel<<<27 * kWgChunks, kCvThreads, 0, st>>>(in1, l.cin1, in2, l.cin2, dz, l.cout, d, n.part);
    int nw = 27 * l.cout * Ct;
    reduce_chunks_kernel<<<ew(nw), 256, 0, st>>>(n.part, n.g + l.woff, nw, kWgChunks);
    bias_grad_kernel<<<l.cout, 256, 0, st>>>(dz, l.cout, S, n.g + l.boff);
    AREC_HIP(hipGetLastError());
    if (din) {
        weight_transpose_kernel<<<ew(nw), 256, 0, st>>>(n.pb + l.woff, n.wt, l.cout, Ct);
        int rc = launch_conv3(dz, l.cout, nullptr, 0, n.wt, nullptr, din, Ct, d, 0, st);
        if (rc) return rc;
    }
    return 0;
}

static int unet_train_step(UNet &n, const float *x, const float *target, float lr, double *loss,
                           hipStream_t st) {
    int rc;
    if ((rc = unet_forward(n, x, n.dy, st))) return rc;  // y in n.dy (overwritten by dY below)
    const int64_t S0 = n.vox(0);
    dim3 lg = ew(S0);
    mse_kernel<<<lg, 256, 0, st>>>(n.dy, target, n.dy, S0, n.lpart);
    Dims d0 = n.dims(0), d1 = n.dims(1), d2 = n.dims(2);
    // last layer: dT5B -> g0[0]
    conv_last_bwd_kernel<<<1024, 256, 0, st>>>(n.act[T5B], n.p + n.wlast, n.dy, S0, 1.f / n.scale, n.g0[0], n.part);
    AREC_HIP(hipGetLastError());
    {
        // the 17 sums (16 weights, bias) over 1024 partials, fixed order
        std::vector<float> hp(1024 * 17);
        AREC_HIP(hipMemcpyAsync(hp.data(), n.part, sizeof(float) * hp.size(), hipMemcpyDeviceToHost, st));
        AREC_HIP(hipStreamSynchronize(st));
        float gl[17] = {0};
        for (int b = 0; b < 1024; b++)
            for (int c = 0; c < 17; c++) gl[c] += hp[b * 17 + c];
        AREC_HIP(hipMemcpyAsync(n.g + n.wlast, gl, sizeof(float) * 17, hipMemcpyHostToDevice, st));
        AREC_HIP(hipStreamSynchronize(st));
    }
    bf16_t **g0 = n.g0, **g1 = n.g1, **g2 = n.g2;
    // decoder, full resolution: L8 (T5A -> T5B), L7 (S1 | up(T4B) -> T5A)
    if ((rc = conv_backward(n, 8, n.act[T5A], nullptr, n.act[T5B], g0[0], g0[1], g0[2], d0, st))) return rc;
    if ((rc = conv_backward(n, 7, n.act[S1], n.act[T4B], n.act[T5A], g0[2], g0[0], n.gin7, d0, st))) return rc;
    upsample_bwd_kernel<<<ew(n.vox(1) * 32), 256, 0, st>>>(n.gin7, 48, 16, 32, g1[0], d0);  // dT4B
    // decoder, half resolution: L6 (T4A -> T4B), L5 (S2 | up(B3) -> T4A)
    if ((rc = conv_backward(n, 6, n.act[T4A], nullptr, n.act[T4B], g1[0], g1[1], g1[2], d1, st))) return rc;
    if ((rc = conv_backward(n, 5, n.act[S2], n.act[B3], n.act[T4A], g1[2], g1[0], n.gin5, d1, st))) return rc;
    upsample_bwd_kernel<<<ew(n.vox(2) * 64), 256, 0, st>>>(n.gin5, 96, 32, 64, g2[0], d1);  // dB3
    // bottleneck: L4 (T3A -> B3), L3 (P2 -> T3A)
    if ((rc = conv_backward(n, 4, n.act[T3A], nullptr, n.act[B3], g2[0], g2[1], g2[2], d2, st))) return rc;
    if ((rc = conv_backward(n, 3, n.act[P2], nullptr, n.act[T3A], g2[2], g2[0], g2[1], d2, st))) return rc;
    // dS2 = skip part of the concatenated gradient + pool backward of dP2
    avgpool_bwd_add_kernel<<<ew(n.vox(1) * 32), 256, 0, st>>>(g2[1], n.gin5, 96, 0, g1[1], d1, 32);
    // encoder, half resolution: L2 (T2A -> S2), L1 (P1 -> T2A)
    if ((rc = conv_backward(n, 2, n.act[T2A], nullptr, n.act[S2], g1[1], g1[0], g1[2], d1, st))) return rc;
    if ((rc = conv_backward(n, 1, n.act[P1], nullptr, n.act[T2A], g1[2], g1[0], g1[1], d1, st))) return rc;
    // dS1 = skip part + pool backward of dP1
    avgpool_bwd_add_kernel<<<ew(n.vox(0) * 16), 256, 0, st>>>(g1[1], n.gin7, 48, 0, g0[1], d0, 16);
    // encoder, full resolution: L0 (T1A -> S1), then the first layer (x -> T1A)
    if ((rc = conv_backward(n, 0, n.act[T1A], nullptr, n.act[S1], g0[1], g0[0], g0[2], d0, st))) return rc;
    relu_mask_kernel<<<ew(S0 * 16), 256, 0, st>>>(g0[2], n.act[T1A], g0[0], S0 * 16);
    conv_first_wgrad_kernel<<<1024, 256, 0, st>>>(x, g0[0], d0, n.scale, n.part);
    reduce_chunks_kernel<<<ew(27 * 16 + 16), 256, 0, st>>>(n.part, n.g + n.wfirst, 27 * 16 + 16, 1024);
    AREC_HIP(hipGetLastError());
    // loss
    {
        std::vector<double> lp(lg.x);
        AREC_HIP(hipMemcpyAsync(lp.data(), n.lpart, sizeof(double) * lg.x, hipMemcpyDeviceToHost, st));
        AREC_HIP(hipStreamSynchronize(st));
        double s = 0;
        for (double q : lp) s += q;
        *loss = s / (double)S0;
    }
    // Adam
    n.step++;
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.f - std::pow(b1, (float)n.step), c2 = 1.f - std::pow(b2, (float)n.step);
    adam_kernel<<<ew(n.nparam), 256, 0, st>>>(n.p, n.g, n.m, n.v, n.pb, n.nparam, lr, b1, b2, c1, c2, 1e-8f);
    AREC_HIP(hipGetLastError());
    return 0;
}

}  // namespace arec

using namespace arec;

struct arec_unet {
    UNet net;
};

extern "C" {

int arec_unet_create(arec_ctx *ctx, arec_unet **out, int32_t D, int32_t H, int32_t W, int32_t training,
                     uint64_t seed) {
    if (!ctx || !out || D % 4 || H % 4 || W % 4 || D <= 0 || H <= 0 || W <= 0 || W > kWgMaxW * 16) {
        set_error("arec_unet_create: sizes must be positive multiples of 4");
        return AREC_EINVAL;
    }
    if (training && (W > kWgMaxW || (D * H) % kWgChunks || ((D / 2) * (H / 2)) % kWgChunks ||
                     ((D / 4) * (H / 4)) % kWgChunks)) {
        set_error("arec_unet_create: training needs W <= 128 and D*H/16 divisible by 128");
        return AREC_EINVAL;
    }
    arec_unet *u = new arec_unet;
    u->net.ctx = ctx;
    int rc = unet_alloc(u->net, Dims{D, H, W}, training != 0);
    if (rc) {
        delete u;
        return rc;
    }
    std::vector<float> h;
    unet_init_params(u->net, seed, h);
    if ((rc = unet_upload(u->net, h, ctx->stream))) {
        delete u;
        return rc;
    }
    *out = u;
    return 0;
}

int arec_unet_destroy(arec_unet *u) {
    delete u;
    return 0;
}

int64_t arec_unet_nparams(const arec_unet *u) { return u ? u->net.nparam : 0; }

int arec_unet_get_params(arec_unet *u, float *host) {
    if (!u || !host) return AREC_EINVAL;
    AREC_HIP(hipMemcpy(host, u->net.p, sizeof(float) * u->net.nparam, hipMemcpyDeviceToHost));
    return 0;
}

int arec_unet_set_params(arec_unet *u, const float *host) {
    if (!u || !host) return AREC_EINVAL;
    std::vector<float> h(host, host + u->net.nparam);
    int rc = unet_upload(u->net, h, u->net.ctx->stream);
    if (!rc) AREC_HIP(hipStreamSynchronize(u->net.ctx->stream));
    return rc;
}

int arec_unet_set_scale(arec_unet *u, float scale) {
    if (!u || !(scale > 0)) return AREC_EINVAL;
    u->net.scale = scale;
    return 0;
}

int arec_unet_forward(arec_unet *u, const float *x_cube, float *y_cube) {
    if (!u || !x_cube || !y_cube) return AREC_EINVAL;
    return unet_forward(u->net, x_cube, y_cube, u->net.ctx->stream);
}

int arec_unet_train_step(arec_unet *u, const float *x_cube, const float *t_cube, float lr, double *loss) {
    if (!u || !x_cube || !t_cube || !loss) return AREC_EINVAL;
    return unet_train_step(u->net, x_cube, t_cube, lr, loss, u->net.ctx->stream);
}

// Gradient of the loss with respect to every parameter at the current
// parameters (FP32 host array of nparams), without updating them.
int arec_unet_gradient(arec_unet *u, const float *x_cube, const float *t_cube, double *loss, float *grad_host) {
    if (!u || !x_cube || !t_cube || !loss || !grad_host) return AREC_EINVAL;
    UNet &n = u->net;
    std::vector<float> keep(n.nparam);
    AREC_HIP(hipMemcpy(keep.data(), n.p, sizeof(float) * n.nparam, hipMemcpyDeviceToHost));
    int64_t step = n.step;
    int rc = unet_train_step(n, x_cube, t_cube, 0.0f, loss, n.ctx->stream);
    if (rc) return rc;
    AREC_HIP(hipMemcpy(grad_host, n.g, sizeof(float) * n.nparam, hipMemcpyDeviceToHost));
    n.step = step;
    return unet_upload(n, keep, n.ctx->stream);
}

// Layer-level checks: one MFMA convolution (FP32 in/out, BF16 inside).
int arec_conv3_check(arec_ctx *ctx, const float *in1, int32_t C1, const float *in2, int32_t C2,
                     const float *w, const float *bias, float *out, int32_t Cout, int32_t D, int32_t H,
                     int32_t W, int32_t relu) {
    if (!ctx || !in1 || !w || !out) return AREC_EINVAL;
    hipStream_t st = ctx->stream;
    int64_t S = (int64_t)D * H * W, S2 = (int64_t)(D / 2) * (H / 2) * (W / 2);
    int Ct = C1 + C2;
    bf16_t *b1, *b2 = nullptr, *bw, *bo;
    AREC_HIP(hipMalloc((void **)&b1, 2 * S * C1));
    if (C2) AREC_HIP(hipMalloc((void **)&b2, 2 * S2 * C2));
    AREC_HIP(hipMalloc((void **)&bw, 2 * 27 * Cout * Ct));
    // output plus a guard of 4 W-rows filled with 0x7f7f, checked afterwards
    const int64_t guard = 4 * (int64_t)W * Cout;
    AREC_HIP(hipMalloc((void **)&bo, 2 * (S * Cout + guard)));
    AREC_HIP(hipMemsetAsync(bo + S * Cout, 0x7f, 2 * guard, st));
    to_bf16_kernel<<<ew(S * C1), 256, 0, st>>>(in1, b1, S * C1);
    if (C2) to_bf16_kernel<<<ew(S2 * C2), 256, 0, st>>>(in2, b2, S2 * C2);
    to_bf16_kernel<<<ew(27 * Cout * Ct), 256, 0, st>>>(w, bw, 27 * Cout * Ct);
    int rc = launch_conv3(b1, C1, b2, C2, bw, bias, bo, Cout, Dims{D, H, W}, relu, st);
    if (!rc) {
        // widen to FP32
        std::vector<uint16_t> h(S * Cout);
        hipMemcpyAsync(h.data(), bo, 2 * S * Cout, hipMemcpyDeviceToHost, st);
        hipStreamSynchronize(st);
        std::vector<float> f(S * Cout);
        for (int64_t i = 0; i < S * Cout; i++) {
            uint32_t u32 = (uint32_t)h[i] << 16;
            memcpy(&f[i], &u32, 4);
        }
        hipMemcpy(out, f.data(), 4 * S * Cout, hipMemcpyHostToDevice);
        std::vector<uint16_t> g(guard);
        hipMemcpy(g.data(), bo + S * Cout, 2 * guard, hipMemcpyDeviceToHost);
        for (int64_t i = 0; i < guard; i++)
            if (g[i] != 0x7f7f) {
                set_error("arec_conv3_check: the convolution wrote past its output");
                rc = AREC_EINVAL;
                break;
            }
    }
    hipFree(b1);
    hipFree(b2);
    hipFree(bw);
    hipFree(bo);
    return rc;
}

// Weight gradient of one MFMA convolution: dW[27][Cout][C1+C2] (FP32) from
// FP32 inputs rounded to BF16.
int arec_conv3_wgrad_check(arec_ctx *ctx, const float *in1, int32_t C1, const float *in2, int32_t C2,
                           const float *dz, int32_t Cout, int32_t D, int32_t H, int32_t W, float *dw) {
    if (!ctx || !in1 || !dz || !dw || W > kWgMaxW || (D * H) % kWgChunks) return AREC_EINVAL;
    hipStream_t st = ctx->stream;
    int64_t S = (int64_t)D * H * W, S2 = (int64_t)(D / 2) * (H / 2) * (W / 2);
    int Ct = C1 + C2;
    bf16_t *b1, *b2 = nullptr, *bz;
    float *part;
    AREC_HIP(hipMalloc((void **)&b1, 2 * S * C1));
    if (C2) AREC_HIP(hipMalloc((void **)&b2, 2 * S2 * C2));
    AREC_HIP(hipMalloc((void **)&bz, 2 * S * Cout));
    AREC_HIP(hipMalloc((void **)&part, 4 * (size_t)kWgChunks * 27 * Cout * Ct));
    to_bf16_kernel<<<ew(S * C1), 256, 0, st>>>(in1, b1, S * C1);
    if (C2) to_bf16_kernel<<<ew(S2 * C2), 256, 0, st>>>(in2, b2, S2 * C2);
    to_bf16_kernel<<<ew(S * Cout), 256, 0, st>>>(dz, bz, S * Cout);
    conv3_wgrad_kernel<<<27 * kWgChunks, kCvThreads, 0, st>>>(b1, C1, b2, C2, bz, Cout, Dims{D, H, W}, part);
    int nw = 27 * Cout * Ct;
    reduce_chunks_kernel<<<ew(nw), 256, 0, st>>>(part, dw, nw, kWgChunks);
    hipError_t e = hipGetLastError();
    hipStreamSynchronize(st);
    hipFree(b1);
    hipFree(b2);
    hipFree(bz);
    hipFree(part);
    return e == hipSuccess ? 0 : hip_fail(e, "arec_conv3_wgrad_check");
}

// cylinder <-> cube (radius as in arecAllocateCylImage, n = round(2r+1))
int arec_cyl_to_cube(arec_ctx *ctx, const float *x_dev, double radius, int64_t ny, float *cube_dev) {
    if (!ctx || !x_dev || !cube_dev) return AREC_EINVAL;
    std::vector<int> cord;
    int nrays = cyl_mask(radius, &cord);
    int n = (int)std::lround(2 * radius + 1);
    int *dc;
    AREC_HIP(hipMalloc((void **)&dc, sizeof(int) * cord.size()));
    AREC_HIP(hipMemcpyAsync(dc, cord.data(), sizeof(int) * cord.size(), hipMemcpyHostToDevice, ctx->stream));
    AREC_HIP(hipMemsetAsync(cube_dev, 0, sizeof(float) * (size_t)ny * n * n, ctx->stream));
    const int64_t ntile = (int64_t)((nrays + kCubeTile - 1) / kCubeTile) * ((ny + kCubeTile - 1) / kCubeTile);
    cyl_to_cube_kernel<<<(unsigned)std::max<int64_t>(1, std::min<int64_t>(ntile, 16384)), 256, 0, ctx->stream>>>(
        x_dev, dc, nrays, (int)ny, n, cube_dev);
    AREC_HIP(hipGetLastError());
    AREC_HIP(hipStreamSynchronize(ctx->stream));
    hipFree(dc);
    return 0;
}

int arec_cube_to_cyl(arec_ctx *ctx, const float *cube_dev, double radius, int64_t ny, float *x_dev) {
    if (!ctx || !x_dev || !cube_dev) return AREC_EINVAL;
    std::vector<int> cord;
    int nrays = cyl_mask(radius, &cord);
    int n = (int)std::lround(2 * radius + 1);
    int *dc;
    AREC_HIP(hipMalloc((void **)&dc, sizeof(int) * cord.size()));
    AREC_HIP(hipMemcpyAsync(dc, cord.data(), sizeof(int) * cord.size(), hipMemcpyHostToDevice, ctx->stream));
    const int64_t ntile = (int64_t)((nrays + kCubeTile - 1) / kCubeTile) * ((ny + kCubeTile - 1) / kCubeTile);
    cube_to_cyl_kernel<<<(unsigned)std::max<int64_t>(1, std::min<int64_t>(ntile, 16384)), 256, 0, ctx->stream>>>(
        cube_dev, dc, nrays, (int)ny, n, x_dev);
    AREC_HIP(hipGetLastError());
    AREC_HIP(hipStreamSynchronize(ctx->stream));
    hipFree(dc);
    return 0;
}

}  // extern "C"
