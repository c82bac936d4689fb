// Learned seed network kernels for gfx950 (no counterpart in the reference).
//
// A small correlation CNN predicts a coarse displacement field from the
// quarter-resolution reference and target; its forward and training passes run
// on these kernels.  Activations and weights are bfloat16 (NHWC, channel
// stride and offset so that concatenations are channel slices of one buffer);
// every product accumulates in float32 on the matrix cores
// (v_mfma_f32_16x16x32_bf16); gradients of activations and weights are float32.
//
// 3x3 convolutions (padding 1, stride 1 or 2) are implicit GEMMs:
//   forward   Y[m][co]  = sum_k  im2col(X)[m][k]  W[co][k]        k = tap*Cin + ci
//   data grad dX[m][ci] = sum_k' col(dY)[m][k']   WT[ci][k']      k' = tap*Cout + co
//   weight    dW[co][k] = sum_m  dY[m][co]        im2col(X)[m][k]
// with the operands gathered straight from HBM/L2 into the MFMA registers.
// MFMA 16x16x32 bf16 layout (wave64): lane l holds A[l%16][8(l/16)+i] and
// B[8(l/16)+i][l%16], i = 0..7, and D[4(l/16)+j][l%16], j = 0..3.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <mutex>

#include "ldic_device.h"
#include "livedic.h"

namespace ldic {
namespace {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

__device__ __forceinline__ s16x8 join8(s16x4 lo, s16x4 hi) {
    return s16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

constexpr float kLeaky = 0.1f;

// im2col value of input pixel (n, oy, ox) at GEMM index k (tap-major).
__device__ __forceinline__ uint16_t im2col_at(const ldic_nn_conv& g, const uint16_t* X, int n, int oy, int ox, int k) {
    if (k >= 9 * g.in_channels) return 0;
    const int tap = k / g.in_channels, ci = k - tap * g.in_channels;
    const int ky = tap / 3, kx = tap - 3 * ky;
    const int iy = oy * g.stride + ky - 1, ix = ox * g.stride + kx - 1;
    if (iy < 0 || iy >= g.height || ix < 0 || ix >= g.width) return 0;
    return X[(((size_t)n * g.height + iy) * g.width + ix) * g.in_stride + g.in_offset + ci];
}

// Four consecutive im2col values starting at k (k % 4 == 0).  VEC: the four
// share one tap (in_channels, in_stride and in_offset are multiples of 4).
template <bool VEC>
__device__ __forceinline__ s16x4 im2col4(const ldic_nn_conv& g, const uint16_t* X, int n, int oy, int ox, int k,
                                         bool valid) {
    s16x4 a = {0, 0, 0, 0};
    if (!valid) return a;
    if (VEC) {
        if (k >= 9 * g.in_channels) return a;
        const int tap = k / g.in_channels, ci = k - tap * g.in_channels;
        const int ky = tap / 3, kx = tap - 3 * ky;
        const int iy = oy * g.stride + ky - 1, ix = ox * g.stride + kx - 1;
        if (iy < 0 || iy >= g.height || ix < 0 || ix >= g.width) return a;
        const uint16_t* p = X + (((size_t)n * g.height + iy) * g.width + ix) * g.in_stride + g.in_offset + ci;
        return *reinterpret_cast<const s16x4*>(p);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = (short)im2col_at(g, X, n, oy, ox, k + i);
    return a;
}

__device__ __forceinline__ void pixel_of(int m, int Ho, int Wo, int& n, int& oy, int& ox) {
    n = m / (Ho * Wo);
    const int r = m - n * Ho * Wo;
    oy = r / Wo;
    ox = r - oy * Wo;
}

// Forward: a workgroup owns 64 output pixels (16 per wave) and up to 128
// output channels.  The weights are staged through LDS in chunks of kChunk
// GEMM columns, shared by the four waves; each wave gathers its own im2col
// rows straight into the MFMA A registers (both 32-column k-steps of a chunk
// are issued before the first MFMA that needs them).
constexpr int kChunk = 64;
constexpr int kChunkPitch = kChunk + 4;   // bf16 elements per LDS row (breaks bank aliasing)

// PARTIAL: blockIdx.z selects a range of kChunk-aligned GEMM columns (split-K
// for layers with few output pixels); range z stores its partial products in
// slice z of the float32 workspace Yf (slices of M x out_channels, dense), and
// k_conv_epilogue adds the slices in order, then bias and activation, so the
// result does not depend on which range finishes first.
template <bool VEC, int NT, bool PARTIAL>
__global__ __launch_bounds__(256) void k_conv_fwd(ldic_nn_conv g, const uint16_t* X, const uint16_t* Wt,
                                                  const float* bias, int leaky, uint16_t* Y, float* Yf,
                                                  int k_per_split) {
    __shared__ __align__(16) uint16_t Bs[16 * NT * kChunkPitch];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int M = g.batch * g.out_height * g.out_width;
    const int m0 = (blockIdx.x * 4 + wave) * 16;
    const int co_base = blockIdx.y * 16 * NT;
    const int rows = min(16 * NT, g.out_channels - co_base);
    const int ntiles = (rows + 15) / 16;
    const int row = lane & 15, quad = lane >> 4;
    const int m = m0 + row;
    int n = 0, oy = 0, ox = 0;
    const bool mvalid = m < M;
    if (mvalid) pixel_of(m, g.out_height, g.out_width, n, oy, ox);
    const int KP = (9 * g.in_channels + 15) & ~15;
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int kbeg = PARTIAL ? blockIdx.z * k_per_split : 0;
    const int kend = PARTIAL ? min(KP, kbeg + k_per_split) : KP;
    for (int k0 = kbeg; k0 < kend; k0 += kChunk) {
        const int kc = min(kChunk, kend - k0);
        // Stage W[co_base + r][k0 .. k0 + kc) in 8-byte pieces.
        __syncthreads();
        for (int i = threadIdx.x; i < rows * (kc / 4); i += 256) {
            const int r = i / (kc / 4), c4 = i - r * (kc / 4);
            *reinterpret_cast<s16x4*>(Bs + r * kChunkPitch + 4 * c4) =
                *reinterpret_cast<const s16x4*>(Wt + (size_t)(co_base + r) * KP + k0 + 4 * c4);
        }
        __syncthreads();
        if (m0 < M) {
            // kc is a multiple of 16 and a lane's 8 columns start at a multiple of 8,
            // so they lie either all inside the chunk or all outside it (zeros).
            s16x8 a[kChunk / 32];
#pragma unroll
            for (int s = 0; s < kChunk / 32; ++s) {
                const int kq = 32 * s + 8 * quad;
                a[s] = kq < kc ? join8(im2col4<VEC>(g, X, n, oy, ox, k0 + kq, mvalid),
                                       im2col4<VEC>(g, X, n, oy, ox, k0 + kq + 4, mvalid))
                               : s16x8{0, 0, 0, 0, 0, 0, 0, 0};
            }
#pragma unroll 1
            for (int s = 0; s < kChunk / 32; ++s) {
                if (s * 32 >= kc) break;
                const int kq = 32 * s + 8 * quad;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    if (t < ntiles) {
                        const int r = 16 * t + row;
                        s16x8 b = {0, 0, 0, 0, 0, 0, 0, 0};
                        if (r < rows && kq < kc) {
                            const uint16_t* src = Bs + r * kChunkPitch + kq;
                            b = join8(*reinterpret_cast<const s16x4*>(src), *reinterpret_cast<const s16x4*>(src + 4));
                        }
                        acc[t] = mfma32(a[s], b, acc[t]);
                    }
                }
            }
        }
    }
    if (m0 >= M) return;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        if (t >= ntiles) break;
        const int co = co_base + 16 * t + row;
        if (co >= g.out_channels) continue;
        if (PARTIAL) {
            float* slice = Yf + (size_t)blockIdx.z * M * g.out_channels;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int mm = m0 + 4 * quad + i;
                if (mm < M) slice[(size_t)mm * g.out_channels + co] = acc[t][i];
            }
            continue;
        }
        const float bco = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int mm = m0 + 4 * quad + i;
            if (mm >= M) continue;
            float v = acc[t][i] + bco;
            if (leaky) v = v > 0.0f ? v : kLeaky * v;
            const size_t o = (size_t)mm * g.out_stride + g.out_offset + co;
            if (Yf) Yf[o] = v;
            else Y[o] = f2bf(v);
        }
    }
}

// Sum of the split-K slices in slice order, then bias and activation.
__global__ void k_conv_epilogue(ldic_nn_conv g, const float* work, int slices, const float* bias, int leaky,
                                uint16_t* Y, float* Yf) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t M = (size_t)g.batch * g.out_height * g.out_width;
    const size_t count = M * g.out_channels;
    if (i >= count) return;
    const size_t m = i / g.out_channels;
    const int co = (int)(i - m * g.out_channels);
    float sum = work[i];
    for (int z = 1; z < slices; ++z) sum += work[(size_t)z * count + i];
    float v = sum + (bias ? bias[co] : 0.0f);
    if (leaky) v = v > 0.0f ? v : kLeaky * v;
    const size_t o = m * g.out_stride + g.out_offset + co;
    if (Yf) Yf[o] = v;
    else Y[o] = f2bf(v);
}

// Local correlation with one wave per pixel: lane j owns offset d0 + j of
// each 64-offset pass and reads its target feature vector with 16-byte loads;
// the reference vector (channels <= 64, one per lane) enters every product as
// a wave-uniform v_readlane operand, so no cross-lane reduction is needed and
// the 64 outputs of a pass are one coalesced store.  Requires channels, the
// strides and offsets to be multiples of 8.
__global__ __launch_bounds__(256) void k_corr_fwd_wave(ldic_nn_corr c, const uint16_t* F1, const uint16_t* F2,
                                                       uint16_t* out) {
    const int lane = threadIdx.x & 63;
    const size_t p = (size_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const size_t P = (size_t)c.batch * c.height * c.width;
    if (p >= P) return;
    const int D = 2 * c.radius + 1, D2 = D * D;
    const int n = (int)(p / ((size_t)c.height * c.width));
    const int r = (int)(p - (size_t)n * c.height * c.width);
    const int y = r / c.width, x = r - y * c.width;
    const float f1 = lane < c.channels ? bf2f(F1[p * c.f1_stride + c.f1_offset + lane]) : 0.0f;
    for (int d0 = 0; d0 < D2; d0 += kWave) {
        const int d = d0 + lane;
        const int y2 = y + d / D - c.radius, x2 = x + d % D - c.radius;
        const bool inb = d < D2 && y2 >= 0 && y2 < c.height && x2 >= 0 && x2 < c.width;
        const uint16_t* row = F2 + (((size_t)n * c.height + (inb ? y2 : 0)) * c.width + (inb ? x2 : 0)) * c.f2_stride +
                              c.f2_offset;
        float s = 0.0f;
        for (int k = 0; k < c.channels; k += 8) {
            uint4 v = inb ? *reinterpret_cast<const uint4*>(row + k) : make_uint4(0u, 0u, 0u, 0u);
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                s += readlane_f32(f1, k + 2 * j) * __uint_as_float(w[j] << 16);
                s += readlane_f32(f1, k + 2 * j + 1) * __uint_as_float(w[j] & 0xFFFF0000u);
            }
        }
        if (d < D2) out[p * c.out_stride + c.out_offset + d] = f2bf(s * c.scale);
    }
}

// Gradient gather for the data gradient: col(dY)[input pixel][k'] with
// k' = tap*Cout + co; the output pixel that tap (ky, kx) of input (iy, ix)
// feeds is ((iy + 1 - ky) / s, (ix + 1 - kx) / s) when both divide exactly.
__device__ __forceinline__ float dcol_at(const ldic_nn_conv& g, const float* dY, int n, int iy, int ix, int k) {
    if (k >= 9 * g.out_channels) return 0.0f;
    const int tap = k / g.out_channels, co = k - tap * g.out_channels;
    const int ky = tap / 3, kx = tap - 3 * ky;
    const int ty = iy + 1 - ky, tx = ix + 1 - kx;
    if (ty < 0 || tx < 0 || ty % g.stride || tx % g.stride) return 0.0f;
    const int oy = ty / g.stride, ox = tx / g.stride;
    if (oy >= g.out_height || ox >= g.out_width) return 0.0f;
    return dY[(((size_t)n * g.out_height + oy) * g.out_width + ox) * g.out_stride + g.out_offset + co];
}

template <bool VEC>
__device__ __forceinline__ s16x4 dcol4(const ldic_nn_conv& g, const float* dY, int n, int iy, int ix, int k,
                                       bool valid) {
    s16x4 a = {0, 0, 0, 0};
    if (!valid) return a;
    if (VEC) {
        if (k >= 9 * g.out_channels) return a;
        const int tap = k / g.out_channels, co = k - tap * g.out_channels;
        const int ky = tap / 3, kx = tap - 3 * ky;
        const int ty = iy + 1 - ky, tx = ix + 1 - kx;
        if (ty < 0 || tx < 0 || ty % g.stride || tx % g.stride) return a;
        const int oy = ty / g.stride, ox = tx / g.stride;
        if (oy >= g.out_height || ox >= g.out_width) return a;
        const float4 v = *reinterpret_cast<const float4*>(
            dY + (((size_t)n * g.out_height + oy) * g.out_width + ox) * g.out_stride + g.out_offset + co);
        a[0] = (short)f2bf(v.x);
        a[1] = (short)f2bf(v.y);
        a[2] = (short)f2bf(v.z);
        a[3] = (short)f2bf(v.w);
        return a;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = (short)f2bf(dcol_at(g, dY, n, iy, ix, k + i));
    return a;
}

// Data gradient: each wave owns 16 input pixels and up to 128 input channels;
// dX (float32, channel stride/offset of the forward input) is overwritten or
// accumulated.
template <bool VEC>
__global__ __launch_bounds__(256) void k_conv_dgrad(ldic_nn_conv g, const float* dY, const uint16_t* WT, float* dX,
                                                    int dx_stride, int dx_offset, int accumulate) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int M = g.batch * g.height * g.width;
    const int m0 = (blockIdx.x * 4 + wave) * 16;
    if (m0 >= M) return;
    const int ci_base = blockIdx.y * 128;
    const int ntiles = min(8, (g.in_channels - ci_base + 15) / 16);
    const int row = lane & 15, quad = lane >> 4;
    const int m = m0 + row;
    int n = 0, iy = 0, ix = 0;
    const bool mvalid = m < M;
    if (mvalid) pixel_of(m, g.height, g.width, n, iy, ix);
    const int KPT = (9 * g.out_channels + 15) & ~15;
    f32x4 acc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // 32 columns per step; KPT is a multiple of 16, so a lane's 8 columns lie
    // either all inside the row or all beyond it (zeros).
    for (int k0 = 0; k0 < KPT; k0 += 32) {
        const int k = k0 + 8 * quad;
        const bool kvalid = k < KPT;
        const s16x8 a = kvalid ? join8(dcol4<VEC>(g, dY, n, iy, ix, k, mvalid), dcol4<VEC>(g, dY, n, iy, ix, k + 4, mvalid))
                               : s16x8{0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            if (t < ntiles) {
                const int ci = ci_base + 16 * t + row;
                s16x8 b = {0, 0, 0, 0, 0, 0, 0, 0};
                if (ci < g.in_channels && kvalid) {
                    const uint16_t* src = WT + (size_t)ci * KPT + k;
                    b = join8(*reinterpret_cast<const s16x4*>(src), *reinterpret_cast<const s16x4*>(src + 4));
                }
                acc[t] = mfma32(a, b, acc[t]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        if (t >= ntiles) break;
        const int ci = ci_base + 16 * t + row;
        if (ci >= g.in_channels) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int mm = m0 + 4 * quad + i;
            if (mm >= M) continue;
            const size_t o = (size_t)mm * dx_stride + dx_offset + ci;
            dX[o] = accumulate ? dX[o] + acc[t][i] : acc[t][i];
        }
    }
}

// Weight gradient: grid (co tiles, k tiles, pixel chunks); each wave reduces
// kWgradPixels pixels and adds its 16x16 tile into dW with float atomics.
constexpr int kWgradPixels = 256;

template <bool VEC>
__global__ __launch_bounds__(256) void k_conv_wgrad(ldic_nn_conv g, const uint16_t* X, const float* dY, float* dW) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int M = g.batch * g.out_height * g.out_width;
    const int mstart = (blockIdx.z * 4 + wave) * kWgradPixels;
    if (mstart >= M) return;
    const int co0 = blockIdx.x * 16, k0 = blockIdx.y * 16;
    const int row = lane & 15, quad = lane >> 4;
    const int KP = (9 * g.in_channels + 15) & ~15;
    const int co = co0 + row;
    const bool covalid = co < g.out_channels;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int mend = min(M, mstart + kWgradPixels);
    // 32 pixels per step, 8 per lane.
    for (int mb = mstart; mb < mend; mb += 32) {
        s16x8 a = {0, 0, 0, 0, 0, 0, 0, 0}, b = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int mm = mb + 8 * quad + i;
            if (mm < mend) {
                if (covalid) a[i] = (short)f2bf(dY[(size_t)mm * g.out_stride + g.out_offset + co]);
                int n, oy, ox;
                pixel_of(mm, g.out_height, g.out_width, n, oy, ox);
                b[i] = (short)im2col_at(g, X, n, oy, ox, k0 + row);
            }
        }
        acc = mfma32(a, b, acc);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int c = co0 + 4 * quad + i;
        if (c < g.out_channels) atomicAdd(dW + (size_t)c * KP + k0 + row, acc[i]);
    }
}

// dB[co] += sum over pixels of dY[.][co]: one wave per 64-channel group and pixel chunk.
__global__ __launch_bounds__(64) void k_bias_grad(int pixels, int channels, int stride, int offset, const float* dY,
                                                  float* dB) {
    const int co = blockIdx.x * 64 + threadIdx.x;
    if (co >= channels) return;
    const int m0 = blockIdx.y * kWgradPixels, m1 = min(pixels, m0 + kWgradPixels);
    float s = 0.0f;
    for (int m = m0; m < m1; ++m) s += dY[(size_t)m * stride + offset + co];
    atomicAdd(dB + co, s);
}

// WT[ci][tap*Cout + co] = W[co][tap*Cin + ci] (zero padded to KPT columns).
__global__ void k_transpose_weights(int cin, int cout, const uint16_t* W, uint16_t* WT) {
    const int KP = (9 * cin + 15) & ~15, KPT = (9 * cout + 15) & ~15;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cin * KPT) return;
    const int ci = i / KPT, k = i - ci * KPT;
    uint16_t v = 0;
    if (k < 9 * cout) {
        const int tap = k / cout, co = k - tap * cout;
        v = W[(size_t)co * KP + tap * cin + ci];
    }
    WT[i] = v;
}

// dY *= leaky'(Y) in place (Y is the post-activation output; same layout).
__global__ void k_leaky_bwd(size_t count, const uint16_t* Y, float* dY) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    if (!(bf2f(Y[i]) > 0.0f)) dY[i] *= kLeaky;
}

// Local correlation: C[p][d] = scale * sum_c F1[p][c] F2[p + d][c], d over
// (2r+1)^2 offsets (dy outer), zero where p + d leaves the image.
__global__ __launch_bounds__(256) void k_corr_fwd(ldic_nn_corr c, const uint16_t* F1, const uint16_t* F2, uint16_t* out) {
    const int D = 2 * c.radius + 1, D2 = D * D;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)c.batch * c.height * c.width * D2;
    if (t >= total) return;
    const int d = (int)(t % D2);
    const size_t p = t / D2;
    const int n = (int)(p / ((size_t)c.height * c.width));
    const int r = (int)(p - (size_t)n * c.height * c.width);
    const int y = r / c.width, x = r - y * c.width;
    const int dy = d / D - c.radius, dx = d % D - c.radius;
    const int y2 = y + dy, x2 = x + dx;
    float s = 0.0f;
    if (y2 >= 0 && y2 < c.height && x2 >= 0 && x2 < c.width) {
        const uint16_t* a = F1 + p * c.f1_stride + c.f1_offset;
        const uint16_t* b = F2 + (((size_t)n * c.height + y2) * c.width + x2) * c.f2_stride + c.f2_offset;
        for (int k = 0; k < c.channels; ++k) s += bf2f(a[k]) * bf2f(b[k]);
    }
    out[p * c.out_stride + c.out_offset + d] = f2bf(s * c.scale);
}

// Correlation gradient, one thread per (pixel, channel):
//   dF1[p][k] += scale sum_d dC[p][d] F2[p+d][k]
//   dF2[q][k] += scale sum_d dC[q-d][d] F1[q-d][k]
__global__ __launch_bounds__(256) void k_corr_bwd(ldic_nn_corr c, const uint16_t* F1, const uint16_t* F2,
                                                  const float* dC, int dc_stride, int dc_offset, float* dF1,
                                                  int df1_stride, int df1_offset, float* dF2, int df2_stride,
                                                  int df2_offset) {
    const int D = 2 * c.radius + 1, D2 = D * D;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)c.batch * c.height * c.width * c.channels;
    if (t >= total) return;
    const int k = (int)(t % c.channels);
    const size_t p = t / c.channels;
    const int n = (int)(p / ((size_t)c.height * c.width));
    const int r = (int)(p - (size_t)n * c.height * c.width);
    const int y = r / c.width, x = r - y * c.width;
    float g1 = 0.0f, g2 = 0.0f;
    for (int d = 0; d < D2; ++d) {
        const int dy = d / D - c.radius, dx = d % D - c.radius;
        int y2 = y + dy, x2 = x + dx;
        if (y2 >= 0 && y2 < c.height && x2 >= 0 && x2 < c.width) {
            const size_t q = ((size_t)n * c.height + y2) * c.width + x2;
            g1 += dC[p * dc_stride + dc_offset + d] * bf2f(F2[q * c.f2_stride + c.f2_offset + k]);
        }
        y2 = y - dy;
        x2 = x - dx;
        if (y2 >= 0 && y2 < c.height && x2 >= 0 && x2 < c.width) {
            const size_t q = ((size_t)n * c.height + y2) * c.width + x2;
            g2 += dC[q * dc_stride + dc_offset + d] * bf2f(F1[q * c.f1_stride + c.f1_offset + k]);
        }
    }
    if (dF1) dF1[p * df1_stride + df1_offset + k] += c.scale * g1;
    if (dF2) dF2[p * df2_stride + df2_offset + k] += c.scale * g2;
}

// Bilinear warp of F (zero outside) by a float32 flow in pixels of F.
__global__ __launch_bounds__(256) void k_warp_fwd(int batch, int h, int w, int ch, const uint16_t* F, const float* flow,
                                                  uint16_t* out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)batch * h * w * ch) return;
    const int k = (int)(t % ch);
    const size_t p = t / ch;
    const int n = (int)(p / ((size_t)h * w));
    const int r = (int)(p - (size_t)n * h * w);
    const int y = r / w, x = r - y * w;
    const float sx = x + flow[2 * p], sy = y + flow[2 * p + 1];
    const float fx = floorf(sx), fy = floorf(sy);
    const int x0 = (int)fx, y0 = (int)fy;
    const float ax = sx - fx, ay = sy - fy;
    float v = 0.0f;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int xx = x0 + i, yy = y0 + j;
            if (xx < 0 || xx >= w || yy < 0 || yy >= h) continue;
            const float wgt = (i ? ax : 1.0f - ax) * (j ? ay : 1.0f - ay);
            v += wgt * bf2f(F[(((size_t)n * h + yy) * w + xx) * ch + k]);
        }
    out[t] = f2bf(v);
}

// Gradient of the warp with respect to F (the flow is treated as a constant).
__global__ __launch_bounds__(256) void k_warp_bwd(int batch, int h, int w, int ch, const float* flow, const float* dOut,
                                                  float* dF) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)batch * h * w * ch) return;
    const int k = (int)(t % ch);
    const size_t p = t / ch;
    const int n = (int)(p / ((size_t)h * w));
    const int r = (int)(p - (size_t)n * h * w);
    const int y = r / w, x = r - y * w;
    const float sx = x + flow[2 * p], sy = y + flow[2 * p + 1];
    const float fx = floorf(sx), fy = floorf(sy);
    const int x0 = (int)fx, y0 = (int)fy;
    const float ax = sx - fx, ay = sy - fy;
    const float g = dOut[t];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int xx = x0 + i, yy = y0 + j;
            if (xx < 0 || xx >= w || yy < 0 || yy >= h) continue;
            const float wgt = (i ? ax : 1.0f - ax) * (j ? ay : 1.0f - ay);
            atomicAdd(dF + (((size_t)n * h + yy) * w + xx) * ch + k, wgt * g);
        }
}

__device__ __forceinline__ float2 bilerp_flow(const float* f, int n, int h, int w, float x, float y) {
    x = fminf(fmaxf(x, 0.0f), (float)(w - 1));
    y = fminf(fmaxf(y, 0.0f), (float)(h - 1));
    const int x0 = min((int)x, w - 2 < 0 ? 0 : w - 2), y0 = min((int)y, h - 2 < 0 ? 0 : h - 2);
    const int x1 = min(x0 + 1, w - 1), y1 = min(y0 + 1, h - 1);
    const float ax = x - x0, ay = y - y0;
    const float* b = f + (size_t)n * h * w * 2;
    float2 r;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const float v00 = b[((size_t)y0 * w + x0) * 2 + c], v01 = b[((size_t)y0 * w + x1) * 2 + c];
        const float v10 = b[((size_t)y1 * w + x0) * 2 + c], v11 = b[((size_t)y1 * w + x1) * 2 + c];
        const float v = (1.0f - ay) * ((1.0f - ax) * v00 + ax * v01) + ay * ((1.0f - ax) * v10 + ax * v11);
        if (c == 0) r.x = v;
        else r.y = v;
    }
    return r;
}

// Fine flow = 2 * bilinear(coarse flow) at (x/2, y/2): fine pixel x sits at
// coarse coordinate x/2 (both pyramids are centred on even pixels).  Also
// writes the bfloat16 copy into a channel slice (nullable).
__global__ void k_upsample_flow(int batch, int hc, int wc, int hf, int wf, const float* coarse, float* fine,
                                uint16_t* fine_bf, int bf_stride, int bf_offset) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= batch * hf * wf) return;
    const int n = t / (hf * wf), r = t - n * hf * wf;
    const int y = r / wf, x = r - y * wf;
    const float2 v = bilerp_flow(coarse, n, hc, wc, 0.5f * x, 0.5f * y);
    fine[2 * (size_t)t] = 2.0f * v.x;
    fine[2 * (size_t)t + 1] = 2.0f * v.y;
    if (fine_bf) {
        fine_bf[(size_t)t * bf_stride + bf_offset] = f2bf(2.0f * v.x);
        fine_bf[(size_t)t * bf_stride + bf_offset + 1] = f2bf(2.0f * v.y);
    }
}

__global__ void k_copy_slice(size_t pixels, int channels, const uint16_t* src, int ss, int so, uint16_t* dst, int ds,
                             int dof) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= pixels * channels) return;
    const size_t p = t / channels;
    const int k = (int)(t - p * channels);
    dst[p * ds + dof + k] = src[p * ss + so + k];
}

__global__ void k_add_slice_f32(size_t pixels, int channels, const float* src, int ss, int so, float* dst, int ds,
                                int dof) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= pixels * channels) return;
    const size_t p = t / channels;
    const int k = (int)(t - p * channels);
    dst[p * ds + dof + k] += src[p * ss + so + k];
}

// uint8 images (batch, h, w) -> bfloat16 (batch, h, w, 1): (v - 128) / 64.
__global__ void k_u8_to_bf16(size_t count, const uint8_t* src, uint16_t* dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = f2bf(((float)src[i] - 128.0f) * (1.0f / 64.0f));
}

// Huber (delta 1) loss of a flow against its target, weight w / count; the
// gradient is written, the loss added into *loss.
__global__ __launch_bounds__(256) void k_flow_loss(int count, const float* flow, const float* truth, float weight,
                                                   float* grad, float* loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < 2 * count) {
        const float d = flow[i] - truth[i];
        const float ad = fabsf(d);
        l = ad < 1.0f ? 0.5f * d * d : ad - 0.5f;
        grad[i] = weight / count * fminf(fmaxf(d, -1.0f), 1.0f);
    }
    l = wave_sum(l * weight / count);
    if ((threadIdx.x & 63) == 0 && l != 0.0f) atomicAdd(loss, l);
}

// Adam with bias correction; the bfloat16 copy follows the float32 master.
__global__ void k_adam(size_t count, float* param, const float* grad, float* m, float* v, float lr, float b1, float b2,
                       float eps, float c1, float c2, uint16_t* param_bf) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = param[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    param[i] = p;
    if (param_bf) param_bf[i] = f2bf(p);
}

__global__ void k_to_bf16(size_t count, const float* src, uint16_t* dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = f2bf(src[i]);
}

// Affine seeds of a subset grid from a flow field with `scale` finest-level
// pixels per flow pixel (flow pixel j sits at finest coordinate j * scale):
// translation = scale * flow(centre / scale); gradients from central
// differences one flow pixel apart.
__global__ void k_seed_from_flow(ldic_nn_seed s, const float* flow, float* state, uint32_t* active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= s.grid_width * s.grid_height) return;
    const int gx = i % s.grid_width, gy = i / s.grid_width;
    const float cx = gx * s.subset_stride + 0.5f * (s.subset_size - 1);
    const float cy = gy * s.subset_stride + 0.5f * (s.subset_size - 1);
    const float fx = cx / s.scale, fy = cy / s.scale;
    const float2 c = bilerp_flow(flow, 0, s.flow_height, s.flow_width, fx, fy);
    const float2 l = bilerp_flow(flow, 0, s.flow_height, s.flow_width, fx - 1.0f, fy);
    const float2 r = bilerp_flow(flow, 0, s.flow_height, s.flow_width, fx + 1.0f, fy);
    const float2 u = bilerp_flow(flow, 0, s.flow_height, s.flow_width, fx, fy - 1.0f);
    const float2 d = bilerp_flow(flow, 0, s.flow_height, s.flow_width, fx, fy + 1.0f);
    float* o = state + (size_t)i * s.state_stride;
    for (int k = 0; k < s.state_stride; ++k) o[k] = 0.0f;
    o[0] = s.scale * c.x;
    o[3] = s.scale * c.y;
    if (s.gradients) {
        o[1] = 0.5f * (r.x - l.x);
        o[2] = 0.5f * (d.x - u.x);
        o[4] = 0.5f * (r.y - l.y);
        o[5] = 0.5f * (d.y - u.y);
    }
    active[i] = s.selected ? (uint32_t)(s.selected[i] != 0) : 1u;
}

inline unsigned blocks(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

inline int status(hipError_t e) { return e == hipSuccess ? LDIC_OK : LDIC_ERR_DEVICE; }

bool vec_input(const ldic_nn_conv& g) {
    return g.in_channels % 4 == 0 && g.in_stride % 4 == 0 && g.in_offset % 4 == 0;
}

bool vec_output(const ldic_nn_conv& g) {
    return g.out_channels % 4 == 0 && g.out_stride % 4 == 0 && g.out_offset % 4 == 0;
}

int check_conv(const ldic_nn_conv* g) {
    if (!g || g->batch < 1 || g->height < 1 || g->width < 1 || g->in_channels < 1 || g->out_channels < 1 ||
        (g->stride != 1 && g->stride != 2) || g->in_stride < g->in_offset + g->in_channels ||
        g->out_stride < g->out_offset + g->out_channels ||
        g->out_height != (g->height - 1) / g->stride + 1 || g->out_width != (g->width - 1) / g->stride + 1)
        return LDIC_ERR_ARGUMENT;
    return LDIC_OK;
}

template <bool PARTIAL>
void conv_fwd_launch(int tile, bool vec, dim3 grid, hipStream_t st, const ldic_nn_conv& g, const uint16_t* x,
                            const uint16_t* w, const float* bias, int leaky, uint16_t* y, float* yf, int kps) {
    switch (tile) {
        case 32:
            if (vec) k_conv_fwd<true, 2, PARTIAL><<<grid, 256, 0, st>>>(g, x, w, bias, leaky, y, yf, kps);
            else k_conv_fwd<false, 2, PARTIAL><<<grid, 256, 0, st>>>(g, x, w, bias, leaky, y, yf, kps);
            break;
        case 64:
            if (vec) k_conv_fwd<true, 4, PARTIAL><<<grid, 256, 0, st>>>(g, x, w, bias, leaky, y, yf, kps);
            else k_conv_fwd<false, 4, PARTIAL><<<grid, 256, 0, st>>>(g, x, w, bias, leaky, y, yf, kps);
            break;
        default:
            if (vec) k_conv_fwd<true, 8, PARTIAL><<<grid, 256, 0, st>>>(g, x, w, bias, leaky, y, yf, kps);
            else k_conv_fwd<false, 8, PARTIAL><<<grid, 256, 0, st>>>(g, x, w, bias, leaky, y, yf, kps);
            break;
    }
}

}  // namespace
}  // namespace ldic

namespace ldic {
int report_error(int code, const char* message);  // ldic_runtime.hip
}

namespace {

// Scratch of the allocation-free split-K forward (ldic_nn_conv_forward_split_static):
// the float32 partial sums of one call, at most kStaticSplitFloats of them.  Each
// GPU the library runs on holds its own copy until the process ends.  Host code
// never passes `scratch` itself to a kernel: on the host it names a stand-in, not
// a device address; static_scratch() looks up each device's address once.
constexpr size_t kStaticSplitFloats = size_t(1) << 20;  // 4 MiB
__device__ __align__(256) float scratch[kStaticSplitFloats];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch[kMaxScratchDevices] = {};

// The current device's address of `scratch`, from hipGetSymbolAddress on the first
// call on that device (made with the device current).
int static_scratch(float** out) {
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxScratchDevices)
        return ldic::report_error(LDIC_ERR_DEVICE, "no usable current HIP device for the static scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return ldic::report_error(LDIC_ERR_DEVICE, "hipGetSymbolAddress of the static scratch failed");
        g_scratch[device] = static_cast<float*>(p);
    }
    *out = g_scratch[device];
    return LDIC_OK;
}

}  // namespace

using namespace ldic;

extern "C" {

int ldic_nn_conv_forward(const ldic_nn_conv* g, const uint16_t* x, const uint16_t* w, const float* bias,
                         int32_t leaky, uint16_t* y, float* y_f32, void* stream) {
    if (check_conv(g) != LDIC_OK || !x || !w || (!y && !y_f32)) return LDIC_ERR_ARGUMENT;
    return ldic_nn_conv_forward_split(g, x, w, bias, leaky, y, y_f32, nullptr, 1, stream);
}

int ldic_nn_conv_forward_split(const ldic_nn_conv* g, const uint16_t* x, const uint16_t* w, const float* bias,
                               int32_t leaky, uint16_t* y, float* y_f32, float* workspace, int32_t splits,
                               void* stream) {
    if (check_conv(g) != LDIC_OK || !x || !w || (!y && !y_f32) || splits < 1 || (splits > 1 && !workspace))
        return LDIC_ERR_ARGUMENT;
    const int M = g->batch * g->out_height * g->out_width;
    int tile = g->out_channel_tile;
    if (tile == 0) {
        // Enough workgroups to fill the device: narrower channel tiles for small pixel counts.
        tile = 128;
        while (tile > 32 && (size_t)blocks(M, 64) * ((g->out_channels + tile - 1) / tile) < 2048) tile /= 2;
    }
    if (tile != 32 && tile != 64 && tile != 128) return LDIC_ERR_ARGUMENT;
    const int KP = (9 * g->in_channels + 15) & ~15;
    const int chunks = (KP + kChunk - 1) / kChunk;
    splits = std::min(splits, chunks);
    hipStream_t st = (hipStream_t)stream;
    const bool vec = vec_input(*g);
    if (splits == 1) {
        dim3 grid(blocks(M, 64), (g->out_channels + tile - 1) / tile);
        conv_fwd_launch<false>(tile, vec, grid, st, *g, x, w, bias, leaky, y, y_f32, KP);
        return status(hipGetLastError());
    }
    const size_t count = (size_t)M * g->out_channels;
    const int kps = (chunks + splits - 1) / splits * kChunk;
    const int slices = (KP + kps - 1) / kps;   // <= splits slices of the workspace are written
    dim3 grid(blocks(M, 64), (g->out_channels + tile - 1) / tile, slices);
    conv_fwd_launch<true>(tile, vec, grid, st, *g, x, w, nullptr, 0, nullptr, workspace, kps);
    k_conv_epilogue<<<blocks(count), 256, 0, st>>>(*g, workspace, slices, bias, leaky, y, y_f32);
    return status(hipGetLastError());
}

int ldic_nn_conv_forward_split_static(const ldic_nn_conv* g, const uint16_t* x, const uint16_t* w, const float* bias,
                                      int32_t leaky, uint16_t* y, float* y_f32, int32_t splits, void* stream) {
    if (check_conv(g) != LDIC_OK || !x || !w || (!y && !y_f32) || splits < 1)
        return report_error(LDIC_ERR_ARGUMENT, "invalid convolution arguments");
    const size_t count = (size_t)g->batch * g->out_height * g->out_width * g->out_channels;
    if (splits > 1 && count * (size_t)splits > kStaticSplitFloats)
        return report_error(LDIC_ERR_ARGUMENT,
                            "the split-K partial sums (splits * batch * out_height * out_width * out_channels floats) "
                            "exceed the static scratch of 1048576 floats; use ldic_nn_conv_forward_split with a "
                            "workspace");
    float* workspace = nullptr;
    if (splits > 1) {
        const int rc = static_scratch(&workspace);
        if (rc != LDIC_OK) return rc;
    }
    return ldic_nn_conv_forward_split(g, x, w, bias, leaky, y, y_f32, workspace, splits, stream);
}

int ldic_nn_conv_backward(const ldic_nn_conv* g, const uint16_t* x, const uint16_t* w_transposed, const float* dy,
                          float* dx, int32_t dx_stride, int32_t dx_offset, int32_t accumulate, float* dw, float* db,
                          void* stream) {
    if (check_conv(g) != LDIC_OK || !dy) return LDIC_ERR_ARGUMENT;
    hipStream_t st = (hipStream_t)stream;
    if (dx) {
        if (!w_transposed || dx_stride < dx_offset + g->in_channels) return LDIC_ERR_ARGUMENT;
        const int M = g->batch * g->height * g->width;
        dim3 grid(blocks(M, 64), (g->in_channels + 127) / 128);
        if (vec_output(*g))
            k_conv_dgrad<true><<<grid, 256, 0, st>>>(*g, dy, w_transposed, dx, dx_stride, dx_offset, accumulate);
        else
            k_conv_dgrad<false><<<grid, 256, 0, st>>>(*g, dy, w_transposed, dx, dx_stride, dx_offset, accumulate);
        if (hipGetLastError() != hipSuccess) return LDIC_ERR_DEVICE;
    }
    const int M = g->batch * g->out_height * g->out_width;
    if (dw) {
        if (!x) return LDIC_ERR_ARGUMENT;
        const int KP = (9 * g->in_channels + 15) & ~15;
        dim3 grid((g->out_channels + 15) / 16, KP / 16, blocks(M, 4 * kWgradPixels));
        if (vec_input(*g))
            k_conv_wgrad<true><<<grid, 256, 0, st>>>(*g, x, dy, dw);
        else
            k_conv_wgrad<false><<<grid, 256, 0, st>>>(*g, x, dy, dw);
        if (hipGetLastError() != hipSuccess) return LDIC_ERR_DEVICE;
    }
    if (db) {
        dim3 grid((g->out_channels + 63) / 64, blocks(M, kWgradPixels));
        k_bias_grad<<<grid, 64, 0, st>>>(M, g->out_channels, g->out_stride, g->out_offset, dy, db);
        if (hipGetLastError() != hipSuccess) return LDIC_ERR_DEVICE;
    }
    return LDIC_OK;
}

int ldic_nn_transpose_weights(int32_t in_channels, int32_t out_channels, const uint16_t* w, uint16_t* w_transposed,
                              void* stream) {
    if (in_channels < 1 || out_channels < 1 || !w || !w_transposed) return LDIC_ERR_ARGUMENT;
    const int KPT = (9 * out_channels + 15) & ~15;
    k_transpose_weights<<<blocks((size_t)in_channels * KPT), 256, 0, (hipStream_t)stream>>>(in_channels, out_channels,
                                                                                          w, w_transposed);
    return status(hipGetLastError());
}

int ldic_nn_leaky_backward(size_t count, const uint16_t* y, float* dy, void* stream) {
    if (!y || !dy) return LDIC_ERR_ARGUMENT;
    if (count) k_leaky_bwd<<<blocks(count), 256, 0, (hipStream_t)stream>>>(count, y, dy);
    return status(hipGetLastError());
}

int ldic_nn_corr_forward(const ldic_nn_corr* c, const uint16_t* f1, const uint16_t* f2, uint16_t* out, void* stream) {
    if (!c || !f1 || !f2 || !out || c->radius < 0 || c->channels < 1) return LDIC_ERR_ARGUMENT;
    const int D = 2 * c->radius + 1;
    const size_t pixels = (size_t)c->batch * c->height * c->width;
    if (!pixels) return LDIC_OK;
    if (c->channels <= kWave && c->channels % 8 == 0 && c->f1_stride % 8 == 0 && c->f1_offset % 8 == 0 &&
        c->f2_stride % 8 == 0 && c->f2_offset % 8 == 0)
        k_corr_fwd_wave<<<blocks(pixels, 4), 256, 0, (hipStream_t)stream>>>(*c, f1, f2, out);
    else
        k_corr_fwd<<<blocks(pixels * D * D), 256, 0, (hipStream_t)stream>>>(*c, f1, f2, out);
    return status(hipGetLastError());
}

int ldic_nn_corr_backward(const ldic_nn_corr* c, const uint16_t* f1, const uint16_t* f2, const float* dc,
                          int32_t dc_stride, int32_t dc_offset, float* df1, int32_t df1_stride, int32_t df1_offset,
                          float* df2, int32_t df2_stride, int32_t df2_offset, void* stream) {
    if (!c || !f1 || !f2 || !dc) return LDIC_ERR_ARGUMENT;
    const size_t total = (size_t)c->batch * c->height * c->width * c->channels;
    if (total)
        k_corr_bwd<<<blocks(total), 256, 0, (hipStream_t)stream>>>(*c, f1, f2, dc, dc_stride, dc_offset, df1, df1_stride,
                                                                   df1_offset, df2, df2_stride, df2_offset);
    return status(hipGetLastError());
}

int ldic_nn_warp_forward(int32_t batch, int32_t height, int32_t width, int32_t channels, const uint16_t* f,
                         const float* flow, uint16_t* out, void* stream) {
    if (!f || !flow || !out) return LDIC_ERR_ARGUMENT;
    const size_t total = (size_t)batch * height * width * channels;
    if (total) k_warp_fwd<<<blocks(total), 256, 0, (hipStream_t)stream>>>(batch, height, width, channels, f, flow, out);
    return status(hipGetLastError());
}

int ldic_nn_warp_backward(int32_t batch, int32_t height, int32_t width, int32_t channels, const float* flow,
                          const float* d_out, float* d_f, void* stream) {
    if (!flow || !d_out || !d_f) return LDIC_ERR_ARGUMENT;
    const size_t total = (size_t)batch * height * width * channels;
    if (total)
        k_warp_bwd<<<blocks(total), 256, 0, (hipStream_t)stream>>>(batch, height, width, channels, flow, d_out, d_f);
    return status(hipGetLastError());
}

int ldic_nn_upsample_flow(int32_t batch, int32_t coarse_height, int32_t coarse_width, int32_t fine_height,
                          int32_t fine_width, const float* coarse, float* fine, uint16_t* fine_bf16,
                          int32_t bf16_stride, int32_t bf16_offset, void* stream) {
    if (!coarse || !fine || coarse_height < 1 || coarse_width < 1) return LDIC_ERR_ARGUMENT;
    const size_t total = (size_t)batch * fine_height * fine_width;
    if (total)
        k_upsample_flow<<<blocks(total), 256, 0, (hipStream_t)stream>>>(batch, coarse_height, coarse_width, fine_height,
                                                                        fine_width, coarse, fine, fine_bf16,
                                                                        bf16_stride, bf16_offset);
    return status(hipGetLastError());
}

int ldic_nn_copy_slice(size_t pixels, int32_t channels, const uint16_t* src, int32_t src_stride, int32_t src_offset,
                       uint16_t* dst, int32_t dst_stride, int32_t dst_offset, void* stream) {
    if (!src || !dst) return LDIC_ERR_ARGUMENT;
    if (pixels * channels)
        k_copy_slice<<<blocks(pixels * channels), 256, 0, (hipStream_t)stream>>>(pixels, channels, src, src_stride,
                                                                                 src_offset, dst, dst_stride, dst_offset);
    return status(hipGetLastError());
}

int ldic_nn_add_slice(size_t pixels, int32_t channels, const float* src, int32_t src_stride, int32_t src_offset,
                      float* dst, int32_t dst_stride, int32_t dst_offset, void* stream) {
    if (!src || !dst) return LDIC_ERR_ARGUMENT;
    if (pixels * channels)
        k_add_slice_f32<<<blocks(pixels * channels), 256, 0, (hipStream_t)stream>>>(
            pixels, channels, src, src_stride, src_offset, dst, dst_stride, dst_offset);
    return status(hipGetLastError());
}

int ldic_nn_images_to_bf16(size_t count, const uint8_t* src, uint16_t* dst, void* stream) {
    if (!src || !dst) return LDIC_ERR_ARGUMENT;
    if (count) k_u8_to_bf16<<<blocks(count), 256, 0, (hipStream_t)stream>>>(count, src, dst);
    return status(hipGetLastError());
}

int ldic_nn_to_bf16(size_t count, const float* src, uint16_t* dst, void* stream) {
    if (!src || !dst) return LDIC_ERR_ARGUMENT;
    if (count) k_to_bf16<<<blocks(count), 256, 0, (hipStream_t)stream>>>(count, src, dst);
    return status(hipGetLastError());
}

int ldic_nn_flow_loss(int32_t count, const float* flow, const float* truth, float weight, float* grad, float* loss,
                      void* stream) {
    if (!flow || !truth || !grad || !loss || count < 1) return LDIC_ERR_ARGUMENT;
    k_flow_loss<<<blocks(2 * (size_t)count), 256, 0, (hipStream_t)stream>>>(count, flow, truth, weight, grad, loss);
    return status(hipGetLastError());
}

int ldic_nn_adam(size_t count, float* param, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
                 float eps, int32_t step, uint16_t* param_bf16, void* stream) {
    if (!param || !grad || !m || !v || step < 1) return LDIC_ERR_ARGUMENT;
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    if (count)
        k_adam<<<blocks(count), 256, 0, (hipStream_t)stream>>>(count, param, grad, m, v, lr, beta1, beta2, eps, c1, c2,
                                                               param_bf16);
    return status(hipGetLastError());
}

int ldic_nn_seed_from_flow(const ldic_nn_seed* s, const float* flow, float* state, uint32_t* active, void* stream) {
    if (!s || !flow || !state || !active || s->flow_height < 1 || s->flow_width < 1 || !(s->scale > 0.0f) ||
        s->state_stride < 8)
        return LDIC_ERR_ARGUMENT;
    const int n = s->grid_width * s->grid_height;
    if (n > 0) k_seed_from_flow<<<blocks(n), 256, 0, (hipStream_t)stream>>>(*s, flow, state, active);
    return status(hipGetLastError());
}

}  // extern "C"
