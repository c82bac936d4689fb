// Analytic speckle renderer: the synthetic reference/target pairs used for
// measurement, accuracy checks and training data of the learned seed.
//
// A pixel x of the deformed image shows the reference texture at X with
// X + u(X) = x (u given in reference coordinates).  X is found by Newton
// iteration in float64; the texture is a background plus isotropic Gaussian
// blobs, evaluated from a cell-binned blob list; independent Gaussian noise
// (counter-based hash, Box-Muller) is added before rounding and clipping.
#include <hip/hip_runtime.h>

#include "ldic_device.h"
#include "livedic.h"

namespace ldic {

struct RenderArgs {
    int width, height;
    int cells_x, cells_y, cell;
    float cell_origin;
    const float4* blobs;        // x, y, amplitude, unused; sorted by cell
    const int* cell_start;      // cells_x*cells_y + 1 offsets
    float sigma, background, noise_sigma;
    unsigned long long seed;
    ldic_field field;
    const ldic_field* fields;   // optional per-image fields (device), batch index = blockIdx.z
    const int32_t* offsets;     // optional per-image texture offsets (device, x and y)
    uint8_t* out;
    float* out_float;           // optional noiseless float image
};

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ void field_eval(const ldic_field& f, double X, double Y, double& u, double& v, double& uxd, double& uyd,
                           double& vxd, double& vyd) {
    const double dx = X - f.origin_x, dy = Y - f.origin_y;
    const double* p = f.poly;
    u = p[0] + p[1] * dx + p[2] * dy + 0.5 * p[6] * dx * dx + p[7] * dx * dy + 0.5 * p[8] * dy * dy;
    v = p[3] + p[4] * dx + p[5] * dy + 0.5 * p[9] * dx * dx + p[10] * dx * dy + 0.5 * p[11] * dy * dy;
    uxd = p[1] + p[6] * dx + p[7] * dy;
    uyd = p[2] + p[7] * dx + p[8] * dy;
    vxd = p[4] + p[9] * dx + p[10] * dy;
    vyd = p[5] + p[10] * dx + p[11] * dy;
    for (int c = 0; c < 2; ++c) {
        const double* w = f.wave[c];  // amplitude, kx, ky, phase
        if (w[0] == 0.0) continue;
        double arg = w[1] * X + w[2] * Y + w[3];
        double s = sin(arg), co = cos(arg);
        if (c == 0) {
            u += w[0] * s;
            uxd += w[0] * w[1] * co;
            uyd += w[0] * w[2] * co;
        } else {
            v += w[0] * s;
            vxd += w[0] * w[1] * co;
            vyd += w[0] * w[2] * co;
        }
    }
}

__global__ __launch_bounds__(256) void k_render(RenderArgs a) {
    int x = blockIdx.x * 16 + (threadIdx.x & 15);
    int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= a.width || y >= a.height) return;
    const int b = blockIdx.z;
    const ldic_field& field = a.fields ? a.fields[b] : a.field;
    // Texture coordinates of this pixel.
    const double xt = x + (a.offsets ? a.offsets[2 * b] : 0), yt = y + (a.offsets ? a.offsets[2 * b + 1] : 0);
    double X = xt, Y = yt;
    if (field.enabled) {
        for (int it = 0; it < 20; ++it) {
            double u, v, ux, uy, vx, vy;
            field_eval(field, X, Y, u, v, ux, uy, vx, vy);
            double rx = X + u - xt, ry = Y + v - yt;
            double j11 = 1.0 + ux, j12 = uy, j21 = vx, j22 = 1.0 + vy;
            double det = j11 * j22 - j12 * j21;
            X -= (j22 * rx - j12 * ry) / det;
            Y -= (-j21 * rx + j11 * ry) / det;
        }
    }
    const double inv2s2 = 1.0 / (2.0 * (double)a.sigma * (double)a.sigma);
    const double reach2 = (double)a.cell * a.cell;  // the 3x3 cell block covers this radius
    double value = a.background;
    int cx = (int)floor((X - a.cell_origin) / a.cell), cy = (int)floor((Y - a.cell_origin) / a.cell);
    for (int j = cy - 1; j <= cy + 1; ++j) {
        if (j < 0 || j >= a.cells_y) continue;
        for (int i = cx - 1; i <= cx + 1; ++i) {
            if (i < 0 || i >= a.cells_x) continue;
            int c = j * a.cells_x + i;
            for (int b = a.cell_start[c]; b < a.cell_start[c + 1]; ++b) {
                float4 bl = a.blobs[b];
                double ddx = X - bl.x, ddy = Y - bl.y;
                double r2 = ddx * ddx + ddy * ddy;
                if (r2 <= reach2) value += bl.z * exp(-r2 * inv2s2);
            }
        }
    }
    size_t idx = (size_t)y * a.width + x;
    if (a.out_float) a.out_float[idx] = (float)value;
    if (a.noise_sigma > 0.0f) {
        const unsigned long long seed = a.seed + 0x9E3779B97F4A7C15ull * (unsigned long long)b;
        unsigned long long h1 = splitmix64(seed ^ (idx * 2ull + 1ull));
        unsigned long long h2 = splitmix64(h1 ^ 0xD1B54A32D192ED03ull);
        double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        double u2 = ((double)(h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        value += a.noise_sigma * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
    value = rint(value);
    a.out[(size_t)b * a.width * a.height + idx] = (uint8_t)fmin(fmax(value, 0.0), 255.0);
}

// Zero fill of `bytes` bytes: whole 32-bit words, then the trailing bytes.  A
// kernel rather than hipMemsetAsync, so the fill is an ordinary kernel node when
// a caller captures the stream into a graph.
__global__ __launch_bounds__(256) void k_zero(uint32_t* words, size_t count, uint8_t* tail, int tail_bytes) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) words[i] = 0u;
    if (blockIdx.x == 0 && (int)threadIdx.x < tail_bytes) tail[threadIdx.x] = 0;
}

}  // namespace ldic

extern "C" int ldic_render_speckle(const ldic_speckle* sp, const ldic_field* field, int32_t width, int32_t height,
                                   float noise_sigma, uint64_t noise_seed, uint8_t* device_out,
                                   float* device_float_out, void* stream) {
    if (!sp || !device_out || width < 1 || height < 1 || sp->cell < 1) return LDIC_ERR_ARGUMENT;
    ldic::RenderArgs a{};
    a.width = width;
    a.height = height;
    a.cells_x = sp->cells_x;
    a.cells_y = sp->cells_y;
    a.cell = sp->cell;
    a.cell_origin = sp->cell_origin;
    a.blobs = reinterpret_cast<const float4*>(sp->device_blobs);
    a.cell_start = sp->device_cell_start;
    a.sigma = sp->sigma;
    a.background = sp->background;
    a.noise_sigma = noise_sigma;
    a.seed = noise_seed;
    if (field) a.field = *field;
    else a.field.enabled = 0;
    a.out = device_out;
    a.out_float = device_float_out;
    dim3 grid((width + 15) / 16, (height + 15) / 16);
    ldic::k_render<<<grid, 256, 0, (hipStream_t)stream>>>(a);
    return hipGetLastError() == hipSuccess ? LDIC_OK : LDIC_ERR_DEVICE;
}

extern "C" int ldic_render_speckle_batch(const ldic_speckle* sp, const ldic_field* device_fields,
                                         const int32_t* device_offsets, int32_t count, int32_t width, int32_t height,
                                         float noise_sigma, uint64_t noise_seed, uint8_t* device_out, void* stream) {
    if (!sp || !device_out || width < 1 || height < 1 || count < 1 || count > 65535 || sp->cell < 1)
        return LDIC_ERR_ARGUMENT;
    ldic::RenderArgs a{};
    a.width = width;
    a.height = height;
    a.cells_x = sp->cells_x;
    a.cells_y = sp->cells_y;
    a.cell = sp->cell;
    a.cell_origin = sp->cell_origin;
    a.blobs = reinterpret_cast<const float4*>(sp->device_blobs);
    a.cell_start = sp->device_cell_start;
    a.sigma = sp->sigma;
    a.background = sp->background;
    a.noise_sigma = noise_sigma;
    a.seed = noise_seed;
    a.field.enabled = 0;
    a.fields = device_fields;
    a.offsets = device_offsets;
    a.out = device_out;
    dim3 grid((width + 15) / 16, (height + 15) / 16, count);
    ldic::k_render<<<grid, 256, 0, (hipStream_t)stream>>>(a);
    return hipGetLastError() == hipSuccess ? LDIC_OK : LDIC_ERR_DEVICE;
}

extern "C" int ldic_device_alloc(size_t bytes, void** out) {
    if (!out) return LDIC_ERR_ARGUMENT;
    return hipMalloc(out, bytes ? bytes : 1) == hipSuccess ? LDIC_OK : LDIC_ERR_MEMORY;
}

extern "C" int ldic_device_free(void* p) { return hipFree(p) == hipSuccess ? LDIC_OK : LDIC_ERR_DEVICE; }

extern "C" int ldic_copy_to_device(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? LDIC_OK : LDIC_ERR_DEVICE;
}

extern "C" int ldic_copy_to_host(void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? LDIC_OK : LDIC_ERR_DEVICE;
}

extern "C" int ldic_device_synchronize(void) {
    return hipDeviceSynchronize() == hipSuccess ? LDIC_OK : LDIC_ERR_DEVICE;
}

extern "C" int ldic_device_zero(void* p, size_t bytes, void* stream) {
    if (!p && bytes) return LDIC_ERR_ARGUMENT;
    if (!bytes) return LDIC_OK;
    uint8_t* b = static_cast<uint8_t*>(p);
    // Leading bytes up to the first 32-bit boundary go through the byte path.
    size_t head = (4u - ((uintptr_t)b & 3u)) & 3u;
    if (head > bytes) head = bytes;
    if (head) ldic::k_zero<<<1, 256, 0, (hipStream_t)stream>>>(nullptr, 0, b, (int)head);
    const size_t rest = bytes - head, words = rest / 4;
    if (rest) {
        size_t blocks = (words + 255) / 256;
        blocks = blocks < 1 ? 1 : (blocks > 4096 ? 4096 : blocks);
        ldic::k_zero<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(reinterpret_cast<uint32_t*>(b + head), words,
                                                                        b + head + words * 4, (int)(rest % 4));
    }
    return hipGetLastError() == hipSuccess ? LDIC_OK : LDIC_ERR_DEVICE;
}

extern "C" int ldic_device_copy_async(void* dst, const void* src, size_t bytes, void* stream) {
    if ((!dst || !src) && bytes) return LDIC_ERR_ARGUMENT;
    if (!bytes) return LDIC_OK;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess
               ? LDIC_OK
               : LDIC_ERR_DEVICE;
}

extern "C" int ldic_stream_synchronize(void* stream) {
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? LDIC_OK : LDIC_ERR_DEVICE;
}
