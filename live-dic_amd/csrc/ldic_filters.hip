// Image-plane kernels: pyramid reduction, target B-spline prefilter and the
// float64 reference preparation planes (Gaussian, recursive B-spline, gradients).
#include "ldic_device.h"
#include "ldic_kernels.h"

namespace ldic {

// ---------------------------------------------------------------------------
// Exact integer 5x5 [1,4,6,4,1]^2 / 256 reduction with truncation
// (shaders/full_pyramid.wgsl:215-249).  A 32x16 output tile shares its
// reflected 68x36 byte footprint in LDS.
// ---------------------------------------------------------------------------
constexpr int kRedTX = 32, kRedTY = 16;
constexpr int kRedW = 2 * kRedTX + 4, kRedH = 2 * kRedTY + 4;

__global__ __launch_bounds__(256) void k_reduce(const uint8_t* __restrict__ src, int w, int h,
                                                uint8_t* __restrict__ dst) {
    __shared__ uint32_t px[kRedH * kRedW];
    const int ox = blockIdx.x * kRedTX * 2 - 2;
    const int oy = blockIdx.y * kRedTY * 2 - 2;
    for (int i = threadIdx.x; i < kRedH * kRedW; i += 256) {
        int ty = i / kRedW, tx = i - ty * kRedW;
        int sx = mirror_clamp(ox + tx, w), sy = mirror_clamp(oy + ty, h);
        px[i] = src[(size_t)sy * w + sx];
    }
    __syncthreads();
    const int dw = w / 2, dh = h / 2;
    const uint32_t taps[5] = {1u, 4u, 6u, 4u, 1u};
    for (int o = threadIdx.x; o < kRedTX * kRedTY; o += 256) {
        int ly = o / kRedTX, lx = o - ly * kRedTX;
        int x = blockIdx.x * kRedTX + lx, y = blockIdx.y * kRedTY + ly;
        if (x >= dw || y >= dh) continue;
        uint32_t v = 0;
#pragma unroll
        for (int j = 0; j < 5; ++j)
#pragma unroll
            for (int i = 0; i < 5; ++i) v += taps[j] * taps[i] * px[(2 * ly + j) * kRedW + 2 * lx + i];
        dst[(size_t)y * dw + x] = (uint8_t)(v / 256u);
    }
}

hipError_t launch_reduce(const uint8_t* src, int w, int h, uint8_t* dst, hipStream_t s) {
    int dw = w / 2, dh = h / 2;
    if (dw < 1 || dh < 1) return hipErrorInvalidValue;
    dim3 grid((dw + kRedTX - 1) / kRedTX, (dh + kRedTY - 1) / kRedTY);
    k_reduce<<<grid, 256, 0, s>>>(src, w, h, dst);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Target prefilter: 25-tap FIR (Gaussian sigma=1 combined with the truncated
// inverse cubic B-spline filter) along x then y, mirror boundaries
// (shaders/icgn.wgsl:145-214).  Both passes are fused through LDS: the block
// stages its reflected 88x56 byte footprint as bytes (aligned dword loads in
// the interior, reflected byte loads at the image border), filters the rows
// into a float tile, then filters the columns.  Sums keep the reference tap
// order without FMA contraction, so each output equals the two-pass float32
// evaluation bit for bit.
// ---------------------------------------------------------------------------
constexpr int kPfR = 12;
constexpr int kPfTX = 64, kPfTY = 32;
constexpr int kPfInW = kPfTX + 2 * kPfR, kPfInH = kPfTY + 2 * kPfR;
constexpr int kPfInWords = kPfInW / 4;   // 22 dwords of bytes per staged row
constexpr int kPfRun = 4;                // outputs per thread and pass
static_assert(kPfInW % 4 == 0 && kPfR % 4 == 0 && kPfRun == 4 && kPfTX % kPfRun == 0 && kPfTY % kPfRun == 0, "tile");

__global__ __launch_bounds__(256) void k_prefilter(const uint8_t* __restrict__ src, int w, int h,
                                                   const float* __restrict__ fir31, float* __restrict__ coeff) {
#pragma clang fp contract(off)
    __shared__ uint32_t in[kPfInH * kPfInWords];
    __shared__ float4 mid[kPfInH * kPfTX / 4];
    __shared__ float taps[2 * kPfR + 1];
    const int bx = blockIdx.x * kPfTX, by = blockIdx.y * kPfTY;
    if (threadIdx.x < 2 * kPfR + 1) taps[threadIdx.x] = fir31[threadIdx.x + 3];
    const bool interior = bx >= kPfR && bx + kPfTX + kPfR <= w && by >= kPfR && by + kPfTY + kPfR <= h &&
                          (w & 3) == 0 && (reinterpret_cast<uintptr_t>(src) & 3) == 0;
    if (interior) {
        // bx - R and every row start are multiples of 4: whole dwords, no reflection.
        const uint8_t* base = src + (size_t)(by - kPfR) * w + (bx - kPfR);
        for (int i = threadIdx.x; i < kPfInH * kPfInWords; i += 256) {
            const int ty = i / kPfInWords, dx = i - ty * kPfInWords;
            in[i] = *reinterpret_cast<const uint32_t*>(base + (size_t)ty * w + 4 * dx);
        }
    } else {
        uint8_t* in8 = reinterpret_cast<uint8_t*>(in);
        for (int i = threadIdx.x; i < kPfInH * kPfInW; i += 256) {
            const int ty = i / kPfInW, tx = i - ty * kPfInW;
            const int sx = mirror_clamp(bx + tx - kPfR, w), sy = mirror_clamp(by + ty - kPfR, h);
            in8[i] = src[(size_t)sy * w + sx];
        }
    }
    __syncthreads();
    float tp[2 * kPfR + 1];
#pragma unroll
    for (int t = 0; t < 2 * kPfR + 1; ++t) tp[t] = taps[t];
    // Rows: each thread filters kPfRun consecutive outputs from one register
    // window of 2R + kPfRun inputs, read from LDS once per run.
    for (int i = threadIdx.x; i < kPfInH * (kPfTX / kPfRun); i += 256) {
        const int ty = i / (kPfTX / kPfRun), q = i - ty * (kPfTX / kPfRun);
        const uint32_t* row = in + ty * kPfInWords + q;
        float win[2 * kPfR + kPfRun];
#pragma unroll
        for (int j = 0; j < (2 * kPfR + kPfRun) / 4; ++j) {
            const uint32_t v = row[j];
            win[4 * j + 0] = (float)(v & 0xffu);
            win[4 * j + 1] = (float)((v >> 8) & 0xffu);
            win[4 * j + 2] = (float)((v >> 16) & 0xffu);
            win[4 * j + 3] = (float)(v >> 24);
        }
        float r[kPfRun];
#pragma unroll
        for (int o = 0; o < kPfRun; ++o) {
            r[o] = 0.0f;
#pragma unroll
            for (int t = 0; t < 2 * kPfR + 1; ++t) r[o] = r[o] + tp[t] * win[o + t];
        }
        mid[ty * (kPfTX / 4) + q] = make_float4(r[0], r[1], r[2], r[3]);
    }
    __syncthreads();
    // Columns: the same register window along y.
    const float* midf = reinterpret_cast<const float*>(mid);
    for (int i = threadIdx.x; i < (kPfTY / kPfRun) * kPfTX; i += 256) {
        const int ty = (i / kPfTX) * kPfRun, tx = i - (i / kPfTX) * kPfTX;
        const int x = bx + tx;
        if (x >= w) continue;
        const float* col = midf + ty * kPfTX + tx;
        float win[2 * kPfR + kPfRun];
#pragma unroll
        for (int t = 0; t < 2 * kPfR + kPfRun; ++t) win[t] = col[t * kPfTX];
#pragma unroll
        for (int o = 0; o < kPfRun; ++o) {
            const int y = by + ty + o;
            if (y >= h) break;
            float r = 0.0f;
#pragma unroll
            for (int t = 0; t < 2 * kPfR + 1; ++t) r = r + tp[t] * win[o + t];
            coeff[(size_t)y * w + x] = r;
        }
    }
}

hipError_t launch_prefilter(const uint8_t* src, int w, int h, const float* fir31, float* coeff, hipStream_t s) {
    dim3 grid((w + kPfTX - 1) / kPfTX, (h + kPfTY - 1) / kPfTY);
    k_prefilter<<<grid, 256, 0, s>>>(src, w, h, fir31, coeff);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Reference preparation planes in float64 (wgpu_backend.py:544-637).
// ---------------------------------------------------------------------------

// numpy.pad(mode="reflect"): one reflection without edge repeat.
__device__ __forceinline__ int reflect(int c, int n) {
    if (c < 0) c = -c;
    if (c >= n) c = 2 * n - 2 - c;
    return min(max(c, 0), n - 1);
}

struct Taps7 {
    double w[7];
};

__global__ __launch_bounds__(256) void k_gauss_rows(const uint8_t* __restrict__ src, int w, int h, Taps7 t,
                                                    double* __restrict__ out) {
#pragma clang fp contract(off)
    int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const uint8_t* row = src + (size_t)y * w;
    double acc = t.w[0] * (double)row[reflect(x - 3, w)];
#pragma unroll
    for (int i = 1; i < 7; ++i) acc = acc + t.w[i] * (double)row[reflect(x + i - 3, w)];
    out[(size_t)y * w + x] = acc;
}

// Each thread filters kGaussRows consecutive rows of its column from a register
// window of kGaussRows + 6 input rows (one load per input row instead of seven
// per output); the sums are formed in the same order as a one-output thread.
constexpr int kGaussRows = 8;

__global__ __launch_bounds__(256) void k_gauss_cols(const double* __restrict__ src, int w, int h, Taps7 t,
                                                    double* __restrict__ out) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * 256 + threadIdx.x, y0 = blockIdx.y * kGaussRows;
    if (x >= w) return;
    const int rows = min(kGaussRows, h - y0);
    double v[kGaussRows + 6];
#pragma unroll
    for (int j = 0; j < kGaussRows + 6; ++j)
        v[j] = j < rows + 6 ? src[(size_t)reflect(y0 + j - 3, h) * w + x] : 0.0;
#pragma unroll
    for (int r = 0; r < kGaussRows; ++r) {
        if (r >= rows) break;
        double acc = t.w[0] * v[r];
#pragma unroll
        for (int i = 1; i < 7; ++i) acc = acc + t.w[i] * v[r + i];
        out[(size_t)(y0 + r) * w + x] = acc;
    }
}

hipError_t launch_gaussian(const uint8_t* src, int w, int h, const double* w7, double* tmp, double* out,
                           hipStream_t s) {
    Taps7 t;
    for (int i = 0; i < 7; ++i) t.w[i] = w7[i];
    dim3 grid((w + 255) / 256, h), grid_cols((w + 255) / 256, (h + kGaussRows - 1) / kGaussRows);
    k_gauss_rows<<<grid, 256, 0, s>>>(src, w, h, t, tmp);
    k_gauss_cols<<<grid_cols, 256, 0, s>>>(tmp, w, h, t, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_to_double(const uint8_t* __restrict__ src, size_t n, double* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = (double)src[i];
}

hipError_t launch_to_double(const uint8_t* src, int w, int h, double* out, hipStream_t s) {
    size_t n = (size_t)w * h;
    k_to_double<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(src, n, out);
    return hipGetLastError();
}

// Cubic B-spline interpolation prefilter as segmented first-order recursions
// (wgpu_backend.py:564-586).  Every line is cut into kSeg-sample segments that
// run in parallel.  A segment starts its causal recursion kLook samples early
// from a zero state (the start error decays as |z|^kLook < 1e-18) and its
// anti-causal recursion kLook samples late; segments touching a line end use
// the reference's exact boundary conditions.  Lines of the column pass are
// adjacent in memory, so consecutive threads read consecutive addresses.  The
// recursion is serial along a line but its loads are not: they are issued
// kChunk at a time, so a lane has kChunk loads in flight instead of one (the
// row pass, whose lanes read 64 different cache lines per load, is
// latency-bound otherwise).
constexpr int kSeg = 64;
constexpr int kLook = 32;
constexpr int kChunk = 8;

struct Spline {
    double z, gain, end_scale;
    double powers[24];  // z^i, i < horizon (numpy: pole ** arange)
    int horizon;
};

__device__ __forceinline__ double at(const double* p, int line, int i, int w, bool columns) {
    return columns ? p[(size_t)i * w + line] : p[(size_t)line * w + i];
}

template <bool COLUMNS>
__global__ __launch_bounds__(256) void k_spline_causal(const double* __restrict__ in, double* __restrict__ out, int w,
                                                       int h, Spline sp) {
#pragma clang fp contract(off)
    const int lines = COLUMNS ? w : h;
    const int n = COLUMNS ? h : w;
    const int segs = (n + kSeg - 1) / kSeg;
    int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= lines * segs) return;
    int line = id % lines, seg = id / lines;
    int a = seg * kSeg, b = min(n, a + kSeg);
    int start = max(0, a - kLook);
    double c;
    if (start == 0) {
        // Causal initial value: sum_{i < horizon} z^i s_i.
        double acc = 0.0;
        for (int i = 0; i < sp.horizon; ++i) acc = acc + (sp.gain * at(in, line, i, w, COLUMNS)) * sp.powers[i];
        c = acc;
        if (a == 0) out[COLUMNS ? (size_t)line : (size_t)line * w] = c;
    } else {
        c = sp.gain * at(in, line, start, w, COLUMNS);
    }
    for (int i0 = start + 1; i0 < b; i0 += kChunk) {
        double x[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) x[k] = i0 + k < b ? at(in, line, i0 + k, w, COLUMNS) : 0.0;
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
            const int i = i0 + k;
            if (i >= b) break;
            c = sp.gain * x[k] + sp.z * c;
            if (i >= a) out[COLUMNS ? (size_t)i * w + line : (size_t)line * w + i] = c;
        }
    }
}

template <bool COLUMNS>
__global__ __launch_bounds__(256) void k_spline_anticausal(const double* __restrict__ cp, double* __restrict__ out,
                                                           int w, int h, Spline sp) {
#pragma clang fp contract(off)
    const int lines = COLUMNS ? w : h;
    const int n = COLUMNS ? h : w;
    const int segs = (n + kSeg - 1) / kSeg;
    int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= lines * segs) return;
    int line = id % lines, seg = id / lines;
    int a = seg * kSeg, b = min(n, a + kSeg);
    int stop = min(n - 1, b - 1 + kLook);
    double c;
    if (stop == n - 1) {
        c = n >= 2 ? sp.end_scale * (sp.z * at(cp, line, n - 2, w, COLUMNS) + at(cp, line, n - 1, w, COLUMNS))
                   : at(cp, line, 0, w, COLUMNS);
        if (b == n) out[COLUMNS ? (size_t)(n - 1) * w + line : (size_t)line * w + n - 1] = c;
    } else {
        c = sp.z * (0.0 - at(cp, line, stop, w, COLUMNS));
    }
    for (int i0 = stop - 1; i0 >= a; i0 -= kChunk) {
        double x[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; ++k) x[k] = i0 - k >= a ? at(cp, line, i0 - k, w, COLUMNS) : 0.0;
#pragma unroll
        for (int k = 0; k < kChunk; ++k) {
            const int i = i0 - k;
            if (i < a) break;
            c = sp.z * (c - x[k]);
            if (i < b) out[COLUMNS ? (size_t)i * w + line : (size_t)line * w + i] = c;
        }
    }
}

hipError_t launch_bspline(double* plane, double* tmp, int w, int h, hipStream_t s) {
    Spline sp;
    sp.z = sqrt(3.0) - 2.0;
    sp.gain = (1.0 - sp.z) * (1.0 - 1.0 / sp.z);
    sp.end_scale = sp.z / (sp.z * sp.z - 1.0);
    int horizon = (int)ceil(log(1e-12) / log(fabs(sp.z)));
    if (horizon > 24) return hipErrorInvalidValue;
    for (int i = 0; i < horizon; ++i) sp.powers[i] = pow(sp.z, (double)i);
    {
        int rows = ((w + kSeg - 1) / kSeg) * h;
        sp.horizon = min(w, horizon);
        k_spline_causal<false><<<(rows + 255) / 256, 256, 0, s>>>(plane, tmp, w, h, sp);
        k_spline_anticausal<false><<<(rows + 255) / 256, 256, 0, s>>>(tmp, plane, w, h, sp);
    }
    {
        int cols = ((h + kSeg - 1) / kSeg) * w;
        sp.horizon = min(h, horizon);
        k_spline_causal<true><<<(cols + 255) / 256, 256, 0, s>>>(plane, tmp, w, h, sp);
        k_spline_anticausal<true><<<(cols + 255) / 256, 256, 0, s>>>(tmp, plane, w, h, sp);
    }
    return hipGetLastError();
}

// Gradients of the B-spline interpolant at the pixel centres, or central
// differences for the direct interpolants (wgpu_backend.py:624-637); fills the
// float32 reference plane (intensity, gx, gy, 0).
__global__ __launch_bounds__(256) void k_gradients(const double* __restrict__ c, const double* __restrict__ intensity,
                                                   int w, int h, int bspline, double* __restrict__ gx,
                                                   double* __restrict__ gy, float4* __restrict__ ref) {
#pragma clang fp contract(off)
    int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    int xm = reflect(x - 1, w), xp = reflect(x + 1, w);
    int ym = reflect(y - 1, h), yp = reflect(y + 1, h);
    auto P = [&](int yy, int xx) { return c[(size_t)yy * w + xx]; };
    double dx = 0.5 * (P(y, xp) - P(y, xm));
    double dy = 0.5 * (P(yp, x) - P(ym, x));
    if (bspline) {
        dx = (0.5 * (P(ym, xp) - P(ym, xm)) + 4.0 * dx + 0.5 * (P(yp, xp) - P(yp, xm))) / 6.0;
        dy = (0.5 * (P(yp, xm) - P(ym, xm)) + 4.0 * dy + 0.5 * (P(yp, xp) - P(ym, xp))) / 6.0;
    }
    size_t i = (size_t)y * w + x;
    gx[i] = dx;
    gy[i] = dy;
    ref[i] = make_float4((float)intensity[i], (float)dx, (float)dy, 0.0f);
}

hipError_t launch_gradients(const double* coeff, const double* intensity, int w, int h, int bspline, double* gx,
                            double* gy, float4* reference, hipStream_t s) {
    dim3 grid((w + 255) / 256, h);
    k_gradients<<<grid, 256, 0, s>>>(coeff, intensity, w, h, bspline, gx, gy, reference);
    return hipGetLastError();
}

}  // namespace ldic
