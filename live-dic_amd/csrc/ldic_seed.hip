// Exhaustive integer ZNCC seed search on gfx950.
//
// No reference counterpart: the reference seeds the coarsest pyramid level
// with zeros, a Fourier estimate or an external matcher
// (full_pyramid_backend.py:1693-1716); its nearest formulation is the
// candidate ZNCC of shaders/spatial_seed.wgsl:65-190.  Here every integer
// shift in [-R, R]^2 around a prior is scored for every subset, on the raw
// uint8 planes and with exact integer moments, so the peak does not depend on
// summation order:
//   n = S^2, num = n Sfg - Sf Sg, vf = n Sff - Sf^2, vg = n Sgg - Sg^2,
//   zncc = num / sqrt(vf vg) in float64, rounded to float32.
//
// One workgroup per subset.  The target window ((2R+S) rows) and the
// reference subset live in LDS as packed bytes; each lane owns one row shift
// dy and four consecutive column shifts, and forms the four products of a
// reference word with v_alignbyte-shifted target words by v_dot4_u32_u8, so
// one LDS word feeds four shifts.  Row shifts are processed in bands of
// kBand: for each band the window sums Sg and Sgg of every shift come from
// vertical, then in-place horizontal running sums in LDS, which bounds LDS to
// the window plus 2 kBand (2R+S-1) integers for any radius.
#include "ldic_device.h"
#include "ldic_kernels.h"

#include <utility>

namespace ldic {
namespace {

constexpr int kBand = 16;

struct SearchGeom {
    int K;    // shifts per axis, 2R + 1
    int G;    // groups of four column shifts, ceil(K / 4)
    int NW;   // reference words per row, ceil(S / 4)
    int TWW;  // window words per row, G + NW
    int TH;   // window rows, K + S - 1
    int VW;   // running-sum row length, K + S - 1
};

__device__ __forceinline__ uint32_t udot4(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_udot4(a, b, c, false);
}

__device__ __forceinline__ void better(float s, int i, float& best, int& besti) {
    if (s > best || (s == best && i < besti)) {
        best = s;
        besti = i;
    }
}

// The score of one shift from its exact moments; -2 when undefined.
__device__ __forceinline__ float zncc_score(bool inside, int64_t n, int64_t sf, int64_t vf, int64_t sg, int64_t sgg,
                                            uint32_t sfg) {
    if (!inside || vf <= 0) return -2.0f;
    const int64_t vg = n * sgg - sg * sg;
    if (vg <= 0) return -2.0f;
    const int64_t num = n * (int64_t)sfg - sf * sg;
    return (float)((double)num / sqrt((double)vf * (double)vg));
}

struct Window {
    const uint32_t* win;
    const uint32_t* refw;
    int TWW, NW, S, K, wx0, wy0, width, height;
    int64_t n, sf, vf;
    __device__ bool inside(int dx, int dy) const {
        const int x0 = wx0 + dx, y0 = wy0 + dy;
        return x0 >= 0 && y0 >= 0 && x0 + S <= width && y0 + S <= height;
    }
    // Direct evaluation of one shift (used for the peak's neighbours).
    __device__ float score(int dx, int dy) const {
        const uint8_t* wb = reinterpret_cast<const uint8_t*>(win);
        const int pitch = 4 * TWW;
        int64_t sg = 0, sgg = 0;
        uint32_t sfg = 0u;
        for (int r = 0; r < S; ++r) {
            const uint8_t* row = wb + (dy + r) * pitch + dx;
            for (int c = 0; c < S; ++c) {
                const int g = row[c];
                sg += g;
                sgg += g * g;
            }
            const uint32_t* frow = refw + r * NW;
            for (int c = 0; c < NW; ++c) {
                uint32_t word = 0u;
#pragma unroll
                for (int b = 0; b < 4; ++b) word |= (uint32_t)row[4 * c + b] << (8 * b);
                sfg = udot4(frow[c], word, sfg);
            }
        }
        return zncc_score(inside(dx, dy), n, sf, vf, sg, sgg, sfg);
    }
};

__global__ __launch_bounds__(256) void k_zncc_search(SearchArgs a, SearchGeom g) {
    extern __shared__ uint32_t smem[];
    __shared__ uint32_t moments[2];
    __shared__ float wave_best[4];
    __shared__ int wave_besti[4];
    __shared__ float neighbour[4];
    const int subset = xcd_block(blockIdx.x, gridDim.x);
    const int tid = threadIdx.x, nthreads = blockDim.x;
    const int S = a.S, R = a.R, K = g.K, G = g.G, NW = g.NW, TWW = g.TWW, TH = g.TH, VW = g.VW;

    if (a.mask && !a.mask[subset]) {
        if (tid < a.seed_stride && a.seed) a.seed[(size_t)subset * a.seed_stride + tid] = 0.0f;
        if (tid == 0) {
            a.shift[2 * subset] = 0.0f;
            a.shift[2 * subset + 1] = 0.0f;
            a.score[subset] = __builtin_nanf("");
            if (a.active) a.active[subset] = 0u;
        }
        return;
    }

    uint32_t* win = smem;              // TH x TWW words of target bytes
    uint32_t* refw = win + TH * TWW;   // S x NW words of reference bytes
    int* V1 = reinterpret_cast<int*>(refw + S * NW);  // kBand x VW
    int* V2 = V1 + kBand * VW;

    const int gx = subset % a.grid_w, gy = subset / a.grid_w;
    const int ox = gx * a.D, oy = gy * a.D;
    int px = 0, py = 0;
    if (a.prior) {
        // A prior that is not finite or lies beyond the image counts as 0.
        const float fx = a.prior[(size_t)subset * a.prior_stride];
        const float fy = a.prior[(size_t)subset * a.prior_stride + 3];
        px = fabsf(fx) <= (float)a.width ? (int)rintf(fx) : 0;
        py = fabsf(fy) <= (float)a.height ? (int)rintf(fy) : 0;
    }
    const int wx0 = ox + px - R, wy0 = oy + py - R;
    if (tid < 2) moments[tid] = 0u;

    for (int i = tid; i < TH * TWW; i += nthreads) {
        const int row = i / TWW, wc = i - row * TWW;
        const int y = wy0 + row, x = wx0 + 4 * wc;
        uint32_t word = 0u;
        if (y >= 0 && y < a.height) {
            const uint8_t* src = a.target + (size_t)y * a.width;
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (x + b >= 0 && x + b < a.width) word |= (uint32_t)src[x + b] << (8 * b);
        }
        win[i] = word;
    }
    uint32_t pf = 0u, pff = 0u;
    for (int i = tid; i < S * NW; i += nthreads) {
        const int r = i / NW, c = i - r * NW;
        const uint8_t* src = a.reference + (size_t)(oy + r) * a.width + ox;
        uint32_t word = 0u;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (4 * c + b < S) word |= (uint32_t)src[4 * c + b] << (8 * b);
        refw[i] = word;
        pf = udot4(word, 0x01010101u, pf);
        pff = udot4(word, word, pff);
    }
    __syncthreads();
    if (pf) atomicAdd(&moments[0], pf);
    if (pff) atomicAdd(&moments[1], pff);

    Window W{win, refw, TWW, NW, S, K, wx0, wy0, a.width, a.height, (int64_t)S * S, 0, 0};
    const uint8_t* wb = reinterpret_cast<const uint8_t*>(win);
    const int pitch = 4 * TWW;
    float best = -3.0f;
    int besti = 0x7fffffff;
    for (int band = 0; band < K; band += kBand) {
        const int bh = min(kBand, K - band);
        // Column sums of S rows for the band's row shifts, then row sums of S
        // columns in place: V[y][dx] becomes the window sum of (dx, band + y).
        for (int x = tid; x < VW; x += nthreads) {
            int s1 = 0, s2 = 0;
            for (int r = 0; r < S; ++r) {
                const int v = wb[(band + r) * pitch + x];
                s1 += v;
                s2 += v * v;
            }
            V1[x] = s1;
            V2[x] = s2;
            for (int y = 1; y < bh; ++y) {
                const int vo = wb[(band + y - 1) * pitch + x], vn = wb[(band + y + S - 1) * pitch + x];
                s1 += vn - vo;
                s2 += vn * vn - vo * vo;
                V1[y * VW + x] = s1;
                V2[y * VW + x] = s2;
            }
        }
        __syncthreads();
        for (int t = tid; t < 2 * bh; t += nthreads) {
            int* row = (t < bh ? V1 : V2) + (t < bh ? t : t - bh) * VW;
            int s = 0;
            for (int c = 0; c < S; ++c) s += row[c];
            for (int x = 0; x < K; ++x) {
                const int old = row[x];
                const int next = x + 1 < K ? s - old + row[x + S] : 0;
                row[x] = s;
                s = next;
            }
        }
        __syncthreads();
        W.sf = moments[0];
        W.vf = W.n * (int64_t)moments[1] - W.sf * W.sf;
        for (int t = tid; t < bh * G; t += nthreads) {
            const int y = t / G, gi = t - y * G, dy = band + y;
            uint32_t acc0 = 0u, acc1 = 0u, acc2 = 0u, acc3 = 0u;
            for (int r = 0; r < S; ++r) {
                const uint32_t* wrow = win + (dy + r) * TWW + gi;
                const uint32_t* frow = refw + r * NW;
                uint32_t lo = wrow[0];
                for (int c = 0; c < NW; ++c) {
                    const uint32_t hi = wrow[c + 1];
                    const uint32_t f = frow[c];
                    acc0 = udot4(f, lo, acc0);
                    acc1 = udot4(f, __builtin_amdgcn_alignbyte(hi, lo, 1), acc1);
                    acc2 = udot4(f, __builtin_amdgcn_alignbyte(hi, lo, 2), acc2);
                    acc3 = udot4(f, __builtin_amdgcn_alignbyte(hi, lo, 3), acc3);
                    lo = hi;
                }
            }
            const uint32_t acc[4] = {acc0, acc1, acc2, acc3};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int dx = 4 * gi + j;
                if (dx < K) {
                    const float score = zncc_score(W.inside(dx, dy), W.n, W.sf, W.vf, V1[y * VW + dx],
                                                   V2[y * VW + dx], acc[j]);
                    better(score, dy * K + dx, best, besti);
                }
            }
        }
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) {
        const float os = __shfl_xor(best, off);
        const int oi = __shfl_xor(besti, off);
        better(os, oi, best, besti);
    }
    if ((tid & (kWave - 1)) == 0) {
        wave_best[tid >> 6] = best;
        wave_besti[tid >> 6] = besti;
    }
    __syncthreads();
    for (int w = 0; w < nthreads / kWave; ++w) better(wave_best[w], wave_besti[w], best, besti);
    const int bx = besti % K, by = besti / K;
    // Scores of the four neighbours of the peak (left, right, up, down).
    if (tid < 4) {
        const int nx = bx + (tid == 0 ? -1 : tid == 1 ? 1 : 0), ny = by + (tid == 2 ? -1 : tid == 3 ? 1 : 0);
        neighbour[tid] = (nx >= 0 && nx < K && ny >= 0 && ny < K) ? W.score(nx, ny) : -2.0f;
    }
    __syncthreads();
    if (tid != 0) return;
    const float c0 = best;
    float subx = 0.0f, suby = 0.0f;
    if (c0 > -1.5f) {
        if (bx > 0 && bx < K - 1) {
            const float l = neighbour[0], r = neighbour[1];
            const float den = (l + r) - 2.0f * c0;
            if (l > -1.5f && r > -1.5f && den < 0.0f) subx = (0.5f * (l - r)) / den;
        }
        if (by > 0 && by < K - 1) {
            const float u = neighbour[2], d = neighbour[3];
            const float den = (u + d) - 2.0f * c0;
            if (u > -1.5f && d > -1.5f && den < 0.0f) suby = (0.5f * (u - d)) / den;
        }
    }
    const float sx = (float)(bx - R + px) + subx, sy = (float)(by - R + py) + suby;
    a.shift[2 * subset] = sx;
    a.shift[2 * subset + 1] = sy;
    a.score[subset] = c0;
    if (a.seed) {
        // Translation from the search (the prior's where no shift was valid);
        // the deformation terms of the prior (gradients, curvatures) are kept,
        // zero without a prior.
        float* sd = a.seed + (size_t)subset * a.seed_stride;
        const float* pr = a.prior ? a.prior + (size_t)subset * a.prior_stride : nullptr;
        for (int k = 0; k < a.seed_stride; ++k) {
            const float v = pr && k < a.prior_stride ? pr[k] : 0.0f;
            sd[k] = isfinite(v) ? v : 0.0f;
        }
        if (c0 > -1.5f) {
            sd[0] = sx;
            sd[3] = sy;
        }
    }
    if (a.active) a.active[subset] = 1u;
}

SearchGeom geometry(int S, int R) {
    SearchGeom g;
    g.K = 2 * R + 1;
    g.G = (g.K + 3) / 4;
    g.NW = (S + 3) / 4;
    g.TWW = g.G + g.NW;
    g.TH = g.K + S - 1;
    g.VW = g.K + S - 1;
    return g;
}

__global__ void k_retry_mask(int count, int P, int stride, const float* state, const float* zncc, float zncc_min,
                             const uint8_t* selected, uint8_t* mask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float status = state[(size_t)i * stride + P];
    // NaN ZNCC (no correlation defined) fails the comparison and is retried.
    const bool good = status == 1.0f && (zncc == nullptr || zncc[i] >= zncc_min);
    mask[i] = (selected == nullptr || selected[i]) && !good ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Neighbour re-seeding: every selected subset that is still unaccepted after
// the seeded solve (and its ZNCC fallback) takes the warp of its nearest
// accepted subset, extrapolated to its own centre, as a new seed.  Sources
// are accepted subsets with at least two accepted 8-neighbours, so a lone
// false acceptance does not spread.  The nearest source is found by jump
// flooding over the subset grid (Rong & Tan 2006, "Jump flooding in GPU with
// applications to Voronoi diagram and distance transform"): steps N/2, ..., 1
// and one more step 1; ties go to the smaller subset index.
// ---------------------------------------------------------------------------
constexpr uint32_t kNoSource = 0xFFFFFFFFu;

__device__ __forceinline__ bool accepted_at(const float* state, int stride, int P, int i) {
    return state[(size_t)i * stride + P] >= 0.5f;
}

__global__ void k_reseed_sources(int gw, int gh, int P, int stride, const float* state, uint32_t* nearest) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= gw * gh) return;
    uint32_t out = kNoSource;
    if (accepted_at(state, stride, P, i)) {
        const int x = i % gw, y = i / gw;
        int n = 0;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = x + dx, ny = y + dy;
                if ((dx || dy) && nx >= 0 && ny >= 0 && nx < gw && ny < gh && accepted_at(state, stride, P, ny * gw + nx))
                    ++n;
            }
        if (n >= 2) out = (uint32_t)i;
    }
    nearest[i] = out;
}

__global__ void k_jump_flood(int gw, int gh, int step, const uint32_t* in, uint32_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= gw * gh) return;
    const int x = i % gw, y = i / gw;
    uint32_t best = in[i];
    int bd = 0x7FFFFFFF;
    if (best != kNoSource) {
        const int bx = (int)(best % (uint32_t)gw) - x, by = (int)(best / (uint32_t)gw) - y;
        bd = bx * bx + by * by;
    }
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int nx = x + dx * step, ny = y + dy * step;
            if ((dx == 0 && dy == 0) || nx < 0 || ny < 0 || nx >= gw || ny >= gh) continue;
            const uint32_t c = in[ny * gw + nx];
            if (c == kNoSource) continue;
            const int cx = (int)(c % (uint32_t)gw) - x, cy = (int)(c / (uint32_t)gw) - y;
            const int d = cx * cx + cy * cy;
            if (d < bd || (d == bd && c < best)) {
                best = c;
                bd = d;
            }
        }
    out[i] = best;
}

// The extrapolation is written with explicit roundings so that it does not
// depend on floating-point contraction.
__global__ void k_reseed_apply(int gw, int count, int P, int stride, int M, float D, const float* state,
                               const float* metadata, const uint8_t* selected, const uint32_t* nearest, float* seed,
                               uint32_t* active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float meta1 = metadata[(size_t)i * M + 1];
    const bool usable = (meta1 >= 0.5f || meta1 == -4.0f) && (selected == nullptr || selected[i]);
    const uint32_t src = nearest[i];
    if (!usable || accepted_at(state, stride, P, i) || src == kNoSource) {
        active[i] = 0u;
        return;
    }
    const float* q = state + (size_t)src * stride;
    const float dx = D * (float)(i % gw - (int)(src % (uint32_t)gw));
    const float dy = D * (float)(i / gw - (int)(src / (uint32_t)gw));
    float* o = seed + (size_t)i * stride;
    o[0] = __fadd_rn(__fadd_rn(q[0], __fmul_rn(q[1], dx)), __fmul_rn(q[2], dy));
    o[1] = q[1];
    o[2] = q[2];
    o[3] = __fadd_rn(__fadd_rn(q[3], __fmul_rn(q[4], dx)), __fmul_rn(q[5], dy));
    o[4] = q[4];
    o[5] = q[5];
    for (int k = 6; k < stride; ++k) o[k] = 0.0f;
    active[i] = 1u;
}

}  // namespace

hipError_t launch_reseed(int grid_w, int grid_h, int model, float D, const float* state, const float* metadata,
                         const uint8_t* selected, uint32_t* near_a, uint32_t* near_b, float* seed, uint32_t* active,
                         hipStream_t s) {
    const int count = grid_w * grid_h;
    if (count <= 0) return hipSuccess;
    const int P = param_count(model), stride = state_stride(model), M = metadata_stride(model);
    const dim3 grid((count + 255) / 256), block(256);
    hipLaunchKernelGGL(k_reseed_sources, grid, block, 0, s, grid_w, grid_h, P, stride, state, near_a);
    int step = 1;
    while (step * 2 < max(grid_w, grid_h)) step *= 2;
    uint32_t *in = near_a, *out = near_b;
    for (bool extra = false;; ) {
        hipLaunchKernelGGL(k_jump_flood, grid, block, 0, s, grid_w, grid_h, step, in, out);
        std::swap(in, out);
        if (step > 1) {
            step /= 2;
        } else if (!extra) {
            extra = true;
        } else {
            break;
        }
    }
    hipLaunchKernelGGL(k_reseed_apply, grid, block, 0, s, grid_w, count, P, stride, M, D, state, metadata, selected,
                       in, seed, active);
    return hipGetLastError();
}

hipError_t launch_retry_mask(int count, int model, const float* state, const float* zncc, float zncc_min,
                             const uint8_t* selected, uint8_t* mask, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_retry_mask, dim3((count + 255) / 256), dim3(256), 0, s, count, param_count(model),
                       state_stride(model), state, zncc, zncc_min, selected, mask);
    return hipGetLastError();
}

size_t zncc_search_lds_bytes(int S, int R) {
    const SearchGeom g = geometry(S, R);
    return sizeof(uint32_t) * ((size_t)g.TH * g.TWW + (size_t)S * g.NW + 2 * (size_t)kBand * g.VW);
}

hipError_t launch_zncc_search(const SearchArgs& a, hipStream_t s) {
    const int count = a.grid_w * a.grid_h;
    if (count <= 0) return hipSuccess;
    const SearchGeom g = geometry(a.S, a.R);
    const int tasks = min(g.K, kBand) * g.G;
    const int waves = tasks >= 192 ? 4 : (tasks + kWave - 1) / kWave;
    hipLaunchKernelGGL(k_zncc_search, dim3(count), dim3(kWave * waves), zncc_search_lds_bytes(a.S, a.R), s, a, g);
    return hipGetLastError();
}

}  // namespace ldic
