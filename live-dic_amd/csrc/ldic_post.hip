// Per-subset grid kernels: coarse-to-fine seed transfer, UOD / median outlier
// pass and the final affine ZNCC.  One thread (or one wave for ZNCC) per
// subset; the state grid is a few MB, so these are latency-bound side passes.
#include "ldic_device.h"
#include <algorithm>

#include "ldic_kernels.h"

namespace ldic {

// ---------------------------------------------------------------------------
// Seed transfer (shaders/full_pyramid.wgsl:163-433): bilinear blend of the
// four coarse warps re-centred on the fine subset centre, translation scaled
// by source/destination scale and curvature by its inverse.
// ---------------------------------------------------------------------------
__device__ void accumulate_seed(const SeedArgs& a, int si, float weight, float spx, float spy, float dpx, float dpy,
                                float ts, float acc[12], float& wsum) {
    const bool q = a.model == kQuadratic;
    const float* s = a.src_state + (size_t)si * (q ? 16 : 8);
    float status = q ? s[12] : s[6];
    if (status < -0.5f) return;
    float dx = dpx - spx, dy = dpy - spy;
    float tu = s[0] + s[1] * dx + s[2] * dy;
    float tv = s[3] + s[4] * dx + s[5] * dy;
    float ux = s[1], uy = s[2], vx = s[4], vy = s[5];
    if (q) {
        tu += 0.5f * s[6] * dx * dx + s[7] * dx * dy + 0.5f * s[8] * dy * dy;
        tv += 0.5f * s[9] * dx * dx + s[10] * dx * dy + 0.5f * s[11] * dy * dy;
        ux += s[6] * dx + s[7] * dy;
        uy += s[7] * dx + s[8] * dy;
        vx += s[9] * dx + s[10] * dy;
        vy += s[10] * dx + s[11] * dy;
    }
    acc[0] += weight * tu * ts;
    acc[1] += weight * ux;
    acc[2] += weight * uy;
    acc[3] += weight * tv * ts;
    acc[4] += weight * vx;
    acc[5] += weight * vy;
    if (q) {
        float cs = 1.0f / ts;
        for (int k = 6; k < 12; ++k) acc[k] += weight * s[k] * cs;
    }
    wsum += weight;
}

// Weighted blend of the permitted coarse corners (and the one-based fallback
// neighbour) at destination subset di (interpolate_subset_seed,
// full_pyramid.wgsl:313-400); acc holds the normalised estimate when wsum > 0.
__device__ void interpolate_seed(const SeedArgs& a, int di, uint32_t perm, float acc[12], float& wsum) {
    const bool q = a.model == kQuadratic;
    const int stride = q ? 16 : 8;
    int dxg = di % a.dst_grid_w, dyg = di / a.dst_grid_w;
    float so = (1.0f - a.src_scale) * 0.5f, dof = (1.0f - a.dst_scale) * 0.5f;
    float dcx = 0.5f * a.S + a.D * (float)dxg, dcy = 0.5f * a.S + a.D * (float)dyg;
    float scx = (a.dst_scale * dcx + (dof - so)) / a.src_scale;
    float scy = (a.dst_scale * dcy + (dof - so)) / a.src_scale;
    float sgx = (scx - 0.5f * a.S) / a.D, sgy = (scy - 0.5f * a.S) / a.D;
    int lx = clampi((int)floorf(sgx), 0, max(0, a.src_grid_w - 2));
    int ly = clampi((int)floorf(sgy), 0, max(0, a.src_grid_h - 2));
    float fx = fminf(fmaxf(sgx - (float)lx, 0.0f), 1.0f);
    float fy = fminf(fmaxf(sgy - (float)ly, 0.0f), 1.0f);
    float ts = a.src_scale / a.dst_scale;
    for (int k = 0; k < 12; ++k) acc[k] = 0.0f;
    wsum = 0.0f;
    for (int corner = 0; corner < 4; ++corner) {
        if ((perm & (1u << corner)) == 0u) continue;
        int ox = corner & 1, oy = corner >> 1;
        int sx = lx + ox, sy = ly + oy;
        if (sx >= a.src_grid_w || sy >= a.src_grid_h) continue;
        float wx = ox ? fx : 1.0f - fx, wy = oy ? fy : 1.0f - fy;
        accumulate_seed(a, sy * a.src_grid_w + sx, wx * wy, 0.5f * a.S + a.D * (float)sx,
                        0.5f * a.S + a.D * (float)sy, scx, scy, ts, acc, wsum);
    }
    uint32_t fallback = perm >> 5u;
    if (wsum == 0.0f && fallback != 0u) {
        int si = (int)fallback - 1;
        float st = a.src_state[(size_t)si * stride + (q ? 12 : 6)];
        if (st >= 0.5f) {
            int sx = si % a.src_grid_w, sy = si / a.src_grid_w;
            accumulate_seed(a, si, 1.0f, 0.5f * a.S + a.D * (float)sx, 0.5f * a.S + a.D * (float)sy, scx, scy, ts,
                            acc, wsum);
        }
    }
    if (wsum > 0.0f)
        for (int k = 0; k < 12; ++k) acc[k] /= wsum;
}

__global__ __launch_bounds__(256) void k_transfer_seed(SeedArgs a) {
    int di = blockIdx.x * 256 + threadIdx.x;
    if (di >= a.dst_count) return;
    const bool q = a.model == kQuadratic;
    const int stride = q ? 16 : 8;
    uint32_t perm = a.permissions ? a.permissions[di] : 31u;
    float acc[12], wsum;
    interpolate_seed(a, di, perm, acc, wsum);
    float4* o = reinterpret_cast<float4*>(a.dst_seed + (size_t)di * stride);
    // Seed availability (TRACK_SEED_AVAILABILITY): -1 without a complete
    // stencil; the temporal projection also marks seeds it does not count
    // (project_temporal_seed, full_pyramid.wgsl:437-449).
    const bool counted = wsum > 0.0f && (perm & 16u);
    float marker = (wsum >= 0.99999f && (!a.projection || counted)) ? 0.0f : -1.0f;
    if (q) {
        o[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
        o[1] = make_float4(acc[4], acc[5], acc[6], acc[7]);
        o[2] = make_float4(acc[8], acc[9], acc[10], acc[11]);
        o[3] = make_float4(a.track_availability ? marker : 0.0f, 0.0f, 0.0f, 0.0f);
    } else {
        o[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
        o[1] = make_float4(acc[4], acc[5], a.track_availability ? marker : 0.0f, 0.0f);
    }
    a.dst_active[di] = (perm >> 4u) & 1u;
    if (a.seeded_count && counted) atomicAdd(a.seeded_count, 1u);
}

// refine_temporal_seed (full_pyramid.wgsl:451-515): where the fine subset's
// history and every permitted coarse corner's history share one age, replace
// the absolute pyramid seed by the fine history plus the coarse correction
// (absolute seed minus the coarse history interpolated like a seed).
__global__ __launch_bounds__(256) void k_temporal_refine(RefineArgs r) {
    const SeedArgs& a = r.edge;
    const int di = blockIdx.x * 256 + threadIdx.x;
    if (di >= a.dst_count) return;
    if (a.dst_active[di] == 0u) return;
    const uint32_t age = r.fine_ages[di];
    if (age == 0xFFFFFFFFu || age > r.max_age) return;
    const bool q = a.model == kQuadratic;
    const int stride = q ? 16 : 8;
    float* seed = a.dst_seed + (size_t)di * stride;
    if ((q ? seed[12] : seed[6]) < -0.5f) return;
    const uint32_t perm = a.permissions ? a.permissions[di] : 31u;
    const int dxg = di % a.dst_grid_w, dyg = di / a.dst_grid_w;
    const float cx = (a.dst_scale * (0.5f * a.S + (float)dxg * a.D) + (a.src_scale - a.dst_scale) * 0.5f) / a.src_scale;
    const float cy = (a.dst_scale * (0.5f * a.S + (float)dyg * a.D) + (a.src_scale - a.dst_scale) * 0.5f) / a.src_scale;
    const int lx = clampi((int)floorf((cx - 0.5f * a.S) / a.D), 0, max(0, a.src_grid_w - 2));
    const int ly = clampi((int)floorf((cy - 0.5f * a.S) / a.D), 0, max(0, a.src_grid_h - 2));
    for (int corner = 0; corner < 4; ++corner) {
        if ((perm & (1u << corner)) == 0u) continue;
        const int sx = lx + (corner & 1), sy = ly + (corner >> 1);
        if (sx >= a.src_grid_w || sy >= a.src_grid_h) continue;
        if (r.coarse_ages[sy * a.src_grid_w + sx] != age) return;
    }
    float pred[12], wsum;
    interpolate_seed(a, di, perm, pred, wsum);
    if (wsum < 0.99999f) return;
    const float* prev = r.fine_history + (size_t)di * stride;
    // Parameters in state order: tx ux uy ty vx vy (+ 6 curvature terms at 6..11).
    for (int k = 0; k < 6; ++k) seed[k] = prev[k] + (seed[k] - pred[k]);
    if (q)
        for (int k = 6; k < 12; ++k) seed[k] = prev[k] + (seed[k] - pred[k]);
}

hipError_t launch_temporal_refine(const RefineArgs& a, hipStream_t s) {
    if (a.edge.dst_count <= 0) return hipSuccess;
    k_temporal_refine<<<(a.edge.dst_count + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

// predict_temporal_subset (full_pyramid.wgsl:83-113): a subset whose current
// and previous measurements are both fresh (age 0) and accepted gets
// 2 * current - previous for its parameters; the status and iteration words
// are copied.  Every other subset's seed is its history as is.
__global__ void k_temporal_predict(int count, int stride, int P, const float* history, const float* previous,
                                   const uint32_t* ages, const uint32_t* previous_ages, float* prediction,
                                   uint32_t* predicted_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float* h = history + (size_t)i * stride;
    const float* p = previous + (size_t)i * stride;
    const bool predict = ages[i] == 0u && previous_ages[i] == 0u && h[P] >= 0.5f && p[P] >= 0.5f;
    float* o = prediction + (size_t)i * stride;
    for (int k = 0; k < stride; ++k) o[k] = (predict && k < P) ? 2.0f * h[k] - p[k] : h[k];
    if (predict && predicted_count) atomicAdd(predicted_count, 1u);
}

// save_previous_temporal_state (full_pyramid.wgsl:115-137): before the merge
// replaces the history with an accepted measurement, keep the history as the
// previous measurement; otherwise the previous age grows (saturating below the
// never-set sentinel).
__global__ void k_temporal_save_previous(int count, int stride, int P, const float* raw, const float* output,
                                         const float* history, const uint32_t* ages, float* previous,
                                         uint32_t* previous_ages) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float rs = raw[(size_t)i * stride + P], fs = output[(size_t)i * stride + P];
    if ((rs == 1.0f || rs == 2.0f) && (fs == 1.0f || fs == 2.0f)) {
        for (int k = 0; k < stride; ++k) previous[(size_t)i * stride + k] = history[(size_t)i * stride + k];
        previous_ages[i] = ages[i];
    } else {
        const uint32_t age = previous_ages[i];
        if (age != 0xFFFFFFFFu) previous_ages[i] = min(age, 0xFFFFFFFDu) + 1u;
    }
}

hipError_t launch_temporal_predict(int count, int model, const float* history, const float* previous,
                                   const uint32_t* ages, const uint32_t* previous_ages, float* prediction,
                                   uint32_t* predicted_count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    k_temporal_predict<<<(count + 255) / 256, 256, 0, s>>>(count, state_stride(model), param_count(model), history,
                                                          previous, ages, previous_ages, prediction, predicted_count);
    return hipGetLastError();
}

hipError_t launch_temporal_save_previous(int count, int model, const float* raw, const float* output,
                                         const float* history, const uint32_t* ages, float* previous,
                                         uint32_t* previous_ages, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    k_temporal_save_previous<<<(count + 255) / 256, 256, 0, s>>>(count, state_stride(model), param_count(model), raw,
                                                                output, history, ages, previous, previous_ages);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Temporal history of the finest field (full_pyramid.wgsl:519-575 and the reset
// of full_pyramid_backend.py:660-677): ages 0xFFFFFFFF mean "never set"; a
// subset whose final status is accepted (1 or 2) -- and, unless the filtered
// output is used, whose raw status is accepted too -- is copied into the
// history with age 0; every other subset ages, saturating below the sentinel,
// and its history status becomes -1 once the age exceeds max_age.
// ---------------------------------------------------------------------------
__global__ void k_temporal_reset(int count, int stride, int P, float* history, uint32_t* ages) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    for (int k = 0; k < stride; ++k) history[(size_t)i * stride + k] = k == P ? -1.0f : 0.0f;
    ages[i] = 0xFFFFFFFFu;
}

// raw == nullptr: no measurement this frame (a level the adaptive pyramid
// skipped), so every entry ages.  expired_count (optional) counts entries whose
// age has just passed max_age (REPORT_HISTORY_EXPIRY, full_pyramid.wgsl:548-556).
__global__ void k_temporal_merge(int count, int stride, int P, const float* raw, const float* output,
                                 int use_filtered, uint32_t max_age, float* history, uint32_t* ages,
                                 uint32_t* accepted_count, uint32_t* expired_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float rs = raw ? raw[(size_t)i * stride + P] : 0.0f, fs = raw ? output[(size_t)i * stride + P] : 0.0f;
    const bool raw_ok = rs == 1.0f || rs == 2.0f, final_ok = fs == 1.0f || fs == 2.0f;
    if (final_ok && (use_filtered || raw_ok)) {
        const float* src = (use_filtered ? output : raw) + (size_t)i * stride;
        for (int k = 0; k < stride; ++k) history[(size_t)i * stride + k] = src[k];
        ages[i] = 0u;
        if (accepted_count) atomicAdd(accepted_count, 1u);
    } else {
        const uint32_t age = ages[i];
        if (age != 0xFFFFFFFFu) {
            const uint32_t next = min(age, 0xFFFFFFFDu) + 1u;
            ages[i] = next;
            if (next > max_age) {
                history[(size_t)i * stride + P] = -1.0f;
                if (expired_count && next == max_age + 1u) atomicAdd(expired_count, 1u);
            }
        }
    }
}

// Links of a projection onto the finest grid: all four corners where the
// subset is selected, none elsewhere.
__global__ void k_links_from_selected(int count, const uint8_t* selected, uint32_t* links) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) links[i] = selected[i] ? 31u : 0u;
}

hipError_t launch_links_from_selected(int count, const uint8_t* selected, uint32_t* links, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    k_links_from_selected<<<(count + 255) / 256, 256, 0, s>>>(count, selected, links);
    return hipGetLastError();
}

__global__ void k_fill_u32(int count, uint32_t value, uint32_t* p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

hipError_t launch_temporal_reset(int count, int model, float* history, uint32_t* ages, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    k_temporal_reset<<<(count + 255) / 256, 256, 0, s>>>(count, state_stride(model), param_count(model), history, ages);
    return hipGetLastError();
}

hipError_t launch_temporal_merge(int count, int model, const float* raw, const float* output, int use_filtered,
                                 uint32_t max_age, float* history, uint32_t* ages, uint32_t* accepted_count,
                                 hipStream_t s, uint32_t* expired_count) {
    if (count <= 0) return hipSuccess;
    k_temporal_merge<<<(count + 255) / 256, 256, 0, s>>>(count, state_stride(model), param_count(model), raw, output,
                                                         use_filtered, max_age, history, ages, accepted_count,
                                                         expired_count);
    return hipGetLastError();
}

hipError_t launch_fill_u32(int count, uint32_t value, uint32_t* p, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    k_fill_u32<<<(count + 255) / 256, 256, 0, s>>>(count, value, p);
    return hipGetLastError();
}

// Summary of a packed state field for the host result: accepted (status > 0.5) and
// replaced (status > 1.5) counts, and float64 sums of tx, ty and iterations over the
// accepted subsets (the statistics RegistrationResult reports, engine.py:105-193).
// Stage 1: each workgroup sums one contiguous run of subsets (threads strided, then a
// fixed-order LDS tree) into its partial; stage 2: one workgroup sums the partials in
// the same fixed order.  The result does not depend on scheduling.
constexpr int kSummaryThreads = 256;

__device__ void summary_tree(double (*sums)[kSummaryThreads], double* out) {
    const int t = threadIdx.x;
    __syncthreads();
    for (int half = kSummaryThreads / 2; half > 0; half >>= 1) {
        if (t < half)
            for (int k = 0; k < 5; ++k) sums[k][t] += sums[k][t + half];
        __syncthreads();
    }
    if (t < 5) out[t] = sums[t][0];
}

__global__ void __launch_bounds__(kSummaryThreads) k_field_summary_partial(const float* state, int count, int stride,
                                                                           int P, int per_block, double* partial) {
    __shared__ double sums[5][kSummaryThreads];
    double acc = 0.0, rep = 0.0, tx = 0.0, ty = 0.0, it = 0.0;
    const int begin = blockIdx.x * per_block, end = min(count, begin + per_block);
    for (int i = begin + threadIdx.x; i < end; i += kSummaryThreads) {
        const float* q = state + (size_t)i * stride;
        const float status = q[P];
        if (status > 0.5f) {
            acc += 1.0;
            tx += (double)q[0];
            ty += (double)q[3];
            it += (double)q[P + 1];
        }
        if (status > 1.5f) rep += 1.0;
    }
    const int t = threadIdx.x;
    sums[0][t] = acc; sums[1][t] = rep; sums[2][t] = tx; sums[3][t] = ty; sums[4][t] = it;
    summary_tree(sums, partial + 5 * (size_t)blockIdx.x);
}

__global__ void __launch_bounds__(kSummaryThreads) k_field_summary_final(const double* partial, int blocks,
                                                                         double* out) {
    __shared__ double sums[5][kSummaryThreads];
    const int t = threadIdx.x;
    for (int k = 0; k < 5; ++k) {
        double v = 0.0;
        for (int b = t; b < blocks; b += kSummaryThreads) v += partial[5 * (size_t)b + k];
        sums[k][t] = v;
    }
    summary_tree(sums, out);
}

hipError_t launch_field_summary(const float* state, int count, int model, double* out, hipStream_t s) {
    // out[0..4] receives the summary; out[8 .. 8 + 5 * kSummaryMaxBlocks) holds the partials.
    const int per_block = std::max(kSummaryThreads, (count + kSummaryMaxBlocks - 1) / kSummaryMaxBlocks);
    const int blocks = std::max(1, (count + per_block - 1) / per_block);
    k_field_summary_partial<<<blocks, kSummaryThreads, 0, s>>>(state, count, state_stride(model), param_count(model),
                                                               per_block, out + 8);
    k_field_summary_final<<<1, kSummaryThreads, 0, s>>>(out + 8, blocks, out);
    return hipGetLastError();
}

__global__ void k_scatter_rows(int count, int row, const uint32_t* index, const float* src, float* dst) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * row) return;
    const int i = t / row, k = t - i * row;
    dst[(size_t)index[i] * row + k] = src[t];
}

hipError_t launch_scatter_rows(int count, int row, const uint32_t* index, const float* src, float* dst,
                               hipStream_t s) {
    if (count <= 0) return hipSuccess;
    k_scatter_rows<<<(count * row + 255) / 256, 256, 0, s>>>(count, row, index, src, dst);
    return hipGetLastError();
}

hipError_t launch_transfer_seed(const SeedArgs& a, hipStream_t s) {
    if (a.dst_count <= 0) return hipSuccess;
    k_transfer_seed<<<(a.dst_count + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Outlier pass (shaders/icgn.wgsl:1087-1605).  Median values are exact order
// statistics, so an insertion sort reproduces both the bubble-sorted 3x3 UOD
// median and the quickselect used by the wider median policy.
// ---------------------------------------------------------------------------
template <int CAP>
__device__ __forceinline__ float sorted_median(float* v, int n) {
    for (int i = 1; i < n; ++i) {
        float x = v[i];
        int j = i - 1;
        while (j >= 0 && v[j] > x) {
            v[j + 1] = v[j];
            --j;
        }
        v[j + 1] = x;
    }
    return (n & 1) ? v[n / 2] : 0.5f * (v[n / 2 - 1] + v[n / 2]);
}

__device__ __forceinline__ float recentred(const float* s, int model, int comp, float ox, float oy) {
    if (model == kQuadratic) {
        switch (comp) {
            case 0: return s[0] + s[1] * ox + s[2] * oy + 0.5f * s[6] * ox * ox + s[7] * ox * oy + 0.5f * s[8] * oy * oy;
            case 1: return s[1] + s[6] * ox + s[7] * oy;
            case 2: return s[2] + s[7] * ox + s[8] * oy;
            case 3: return s[3] + s[4] * ox + s[5] * oy + 0.5f * s[9] * ox * ox + s[10] * ox * oy + 0.5f * s[11] * oy * oy;
            case 4: return s[4] + s[9] * ox + s[10] * oy;
            case 5: return s[5] + s[10] * ox + s[11] * oy;
            default: return s[comp];
        }
    }
    switch (comp) {
        case 0: return s[0] + s[1] * ox + s[2] * oy;
        case 3: return s[3] + s[4] * ox + s[5] * oy;
        default: return s[comp];
    }
}

__constant__ int kComponentOrder[12] = {0, 3, 1, 2, 4, 5, 6, 7, 8, 9, 10, 11};

template <int CAP>
__device__ void median_policy(const MedianArgs& a, int i) {
    const int model = a.model;
    const int stride = model == kQuadratic ? 16 : 8;
    const float* own = a.state + (size_t)i * stride;
    float* out = a.filtered + (size_t)i * stride;
    float f[16];
    for (int k = 0; k < stride; ++k) f[k] = own[k];
    const float own_status = model == kQuadratic ? own[12] : own[6];
    const float ref_status = a.metadata[(size_t)i * (model == kQuadratic ? 28 : 16) + 1];
    bool skip = (own_status >= 0.5f && !a.scope) || (ref_status < 0.5f && ref_status != -4.0f);
    if (!skip) {
        const int px = i % a.grid_w, py = i / a.grid_w, gh = a.grid_count / a.grid_w;
        const int radius = a.size / 2;
        const int active = model == kTranslation ? 2 : (model == kQuadratic ? 12 : 6);
        const float* perm = a.roi_words ? a.roi_words + (size_t)i * 4 : nullptr;
        float values[CAP];
        float filtered[12];
        for (int k = 0; k < 12; ++k) filtered[k] = k < stride ? own[k] : 0.0f;
        bool complete = true;
        for (int ai = 0; ai < active && complete; ++ai) {
            int comp = kComponentOrder[ai];
            int count = 0;
            if (own_status >= 0.5f) values[count++] = recentred(own, model, comp, 0.0f, 0.0f);
            for (int dy = -radius; dy <= radius; ++dy)
                for (int dx = -radius; dx <= radius; ++dx) {
                    if (dx == 0 && dy == 0) continue;
                    int nx = px + dx, ny = py + dy;
                    if (nx < 0 || ny < 0 || nx >= a.grid_w || ny >= gh) continue;
                    if (a.roi_enabled) {
                        uint32_t bit = (uint32_t)((dy + 1) * 3 + dx + 1);
                        uint32_t word = (uint32_t)perm[0];
                        if (a.size > 3) {
                            bit = (uint32_t)((dy + 4) * 9 + dx + 4);
                            word = (uint32_t)perm[bit / 24u];
                        }
                        if ((word & (1u << (bit % 24u))) == 0u) continue;
                    }
                    const float* ns = a.state + (size_t)(ny * a.grid_w + nx) * stride;
                    if ((model == kQuadratic ? ns[12] : ns[6]) < 0.5f) continue;
                    values[count++] = recentred(ns, model, comp, -(float)a.D * dx, -(float)a.D * dy);
                }
            int minimum = (a.roi_enabled && a.size == 3 && own_status < 0.5f) ? 3 : 4;
            if (count < minimum) {
                complete = false;
                break;
            }
            filtered[comp] = sorted_median<CAP>(values, count);
        }
        if (complete) {
            for (int k = 0; k < 8; ++k) f[k] = filtered[k];
            if (model == kQuadratic) {
                for (int k = 8; k < 12; ++k) f[k] = filtered[k];
                f[12] = (a.seed_available && own_status < 0.5f) ? 0.0f : own[12];
            } else {
                f[6] = (a.seed_available && own_status < 0.5f) ? 0.0f : own[6];
                f[7] = own[7];
            }
        }
    }
    for (int k = 0; k < stride; ++k) out[k] = f[k];
}

// ---------------------------------------------------------------------------
// Register-only forms of the 3x3 passes.  The neighbour values of one
// component go into a fixed array padded with +inf, an odd-even transposition
// network sorts it, and the median of the first `count` entries is picked by
// selects, so no array is indexed at run time (the generic form above, used
// for wider windows, keeps its list in scratch memory).  Order statistics are
// exact, so the results equal the insertion-sorted medians bit for bit.
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void sort_network(float (&v)[N]) {
#pragma unroll
    for (int pass = 0; pass < N; ++pass)
#pragma unroll
        for (int j = pass & 1; j + 1 < N; j += 2) {
            const float lo = fminf(v[j], v[j + 1]), hi = fmaxf(v[j], v[j + 1]);
            v[j] = lo;
            v[j + 1] = hi;
        }
}

// Median of the `count` smallest entries of a sorted array (count >= 1).
template <int N>
__device__ __forceinline__ float pick_median(const float (&v)[N], int count) {
    float lo = v[0], hi = v[0];
#pragma unroll
    for (int c = 1; c <= N; ++c)
        if (c == count) {
            lo = v[(c - 1) / 2];
            hi = v[c / 2];
        }
    return (count & 1) ? hi : 0.5f * (lo + hi);
}

__device__ constexpr int component_order(int k) { return k == 1 ? 3 : k == 2 ? 1 : k == 3 ? 2 : k; }

struct Neighbours3 {
    const float* state[8];
    bool ok[8];
    float ox[8], oy[8];
    int count;
};

// The eight neighbours in the reference's order (dy outer, dx inner), their
// validity (in the grid, linked under an ROI, accepted) and recentring offsets.
__device__ __forceinline__ void neighbours3(const MedianArgs& a, int i, bool links, Neighbours3& n) {
    const int stride = a.model == kQuadratic ? 16 : 8;
    const int px = i % a.grid_w, py = i / a.grid_w, gh = a.grid_count / a.grid_w;
    const uint32_t perm = links ? (uint32_t)a.roi_words[(size_t)i * 4] : 0x1FFu;
    n.count = 0;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        const int e = d < 4 ? d : d + 1;
        const int dx = e % 3 - 1, dy = e / 3 - 1;
        const int nx = px + dx, ny = py + dy;
        const bool inside = nx >= 0 && ny >= 0 && nx < a.grid_w && ny < gh && ((perm >> e) & 1u);
        const float* ns = a.state + (size_t)(inside ? ny * a.grid_w + nx : i) * stride;
        n.state[d] = ns;
        n.ok[d] = inside && (a.model == kQuadratic ? ns[12] : ns[6]) >= 0.5f;
        n.ox[d] = -(float)a.D * dx;
        n.oy[d] = -(float)a.D * dy;
        n.count += n.ok[d] ? 1 : 0;
    }
}

__device__ void uod_policy3(const MedianArgs& a, int i) {
    const int model = a.model;
    const int stride = model == kQuadratic ? 16 : 8;
    const float* own = a.state + (size_t)i * stride;
    float* out = a.filtered + (size_t)i * stride;
    float f[16];
    for (int k = 0; k < stride; ++k) f[k] = own[k];
    const float own_status = model == kQuadratic ? own[12] : own[6];
    if (own_status >= 0.5f) {
        Neighbours3 n;
        // ROI links are honoured for every model.  The reference's UOD reads them for
        // translation and affine only (icgn.wgsl:1356-1400 has no ROI test) because it
        // rejects quadratic warps with an ROI (roi.py:45-46); this build supports that
        // combination, and a neighbour across a hole is not evidence about a subset.
        neighbours3(a, i, a.roi_enabled, n);
        if (n.count >= 4) {
            const int model_count = model == kTranslation ? 2 : (model == kQuadratic ? 12 : 6);
            const int median_count = a.policy == kReplace ? model_count : 2;
            float med[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) med[k] = k < stride ? own[k] : 0.0f;
            bool flagged = false;
#pragma unroll
            for (int ai = 0; ai < 2; ++ai) {
                const int comp = component_order(ai);
                float v[8];
#pragma unroll
                for (int d = 0; d < 8; ++d)
                    v[d] = n.ok[d] ? recentred(n.state[d], model, comp, n.ox[d], n.oy[d]) : __builtin_inff();
                sort_network(v);
                med[comp] = pick_median(v, n.count);
#pragma unroll
                for (int d = 0; d < 8; ++d)
                    v[d] = n.ok[d] ? fabsf(recentred(n.state[d], model, comp, n.ox[d], n.oy[d]) - med[comp]) : __builtin_inff();
                sort_network(v);
                const float spread = pick_median(v, n.count);
                if (fabsf(own[comp] - med[comp]) / (spread + 0.1f) > 2.0f) flagged = true;
            }
            if (flagged) {
                if (a.policy == kReplace) {
#pragma unroll
                    for (int ai = 2; ai < 12; ++ai) {
                        if (ai >= median_count) break;
                        const int comp = component_order(ai);
                        float v[8];
#pragma unroll
                        for (int d = 0; d < 8; ++d)
                            v[d] = n.ok[d] ? recentred(n.state[d], model, comp, n.ox[d], n.oy[d]) : __builtin_inff();
                        sort_network(v);
                        med[comp] = pick_median(v, n.count);
                    }
#pragma unroll
                    for (int ai = 0; ai < 12; ++ai)
                        if (ai < model_count) f[component_order(ai)] = med[component_order(ai)];
                }
                if (model == kQuadratic) {
                    f[12] = 2.0f;
                    f[14] = 0.0f;
                    f[15] = 0.0f;
                } else {
                    f[6] = 2.0f;
                }
            }
        }
    }
    for (int k = 0; k < stride; ++k) out[k] = f[k];
}

// median_policy for a 3x3 window: own value (if accepted) and the accepted,
// linked neighbours; at least 3 (ROI, rejected subset) or 4 values.
__device__ void median_policy3(const MedianArgs& a, int i) {
    const int model = a.model;
    const int stride = model == kQuadratic ? 16 : 8;
    const float* own = a.state + (size_t)i * stride;
    float* out = a.filtered + (size_t)i * stride;
    float f[16];
    for (int k = 0; k < stride; ++k) f[k] = own[k];
    const float own_status = model == kQuadratic ? own[12] : own[6];
    const float ref_status = a.metadata[(size_t)i * (model == kQuadratic ? 28 : 16) + 1];
    const bool skip = (own_status >= 0.5f && !a.scope) || (ref_status < 0.5f && ref_status != -4.0f);
    if (!skip) {
        Neighbours3 n;
        neighbours3(a, i, a.roi_enabled, n);
        const bool own_ok = own_status >= 0.5f;
        const int count = n.count + (own_ok ? 1 : 0);
        const int minimum = (a.roi_enabled && own_status < 0.5f) ? 3 : 4;
        if (count >= minimum) {
            const int active = model == kTranslation ? 2 : (model == kQuadratic ? 12 : 6);
            float filtered[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) filtered[k] = k < stride ? own[k] : 0.0f;
#pragma unroll
            for (int ai = 0; ai < 12; ++ai) {
                if (ai >= active) break;
                const int comp = component_order(ai);
                float v[9];
                v[0] = own_ok ? recentred(own, model, comp, 0.0f, 0.0f) : __builtin_inff();
#pragma unroll
                for (int d = 0; d < 8; ++d)
                    v[d + 1] = n.ok[d] ? recentred(n.state[d], model, comp, n.ox[d], n.oy[d]) : __builtin_inff();
                sort_network(v);
                filtered[comp] = pick_median(v, count);
            }
            for (int k = 0; k < 8; ++k) f[k] = filtered[k];
            if (model == kQuadratic) {
                for (int k = 8; k < 12; ++k) f[k] = filtered[k];
                f[12] = (a.seed_available && own_status < 0.5f) ? 0.0f : own[12];
            } else {
                f[6] = (a.seed_available && own_status < 0.5f) ? 0.0f : own[6];
                f[7] = own[7];
            }
        }
    }
    for (int k = 0; k < stride; ++k) out[k] = f[k];
}

// 3x3 passes: registers only.
__global__ __launch_bounds__(256) void k_median(MedianArgs a) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.grid_count) return;
    if (a.policy == kMedian)
        median_policy3(a, i);
    else
        uod_policy3(a, i);
}

// Median windows of 5x5 to 9x9 (up to 81 values in scratch memory).
__global__ __launch_bounds__(256) void k_median_wide(MedianArgs a) {
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= a.grid_count) return;
    median_policy<81>(a, i);
}

hipError_t launch_median(const MedianArgs& a, hipStream_t s) {
    if (a.grid_count <= 0) return hipSuccess;
    if (a.policy == kMedian && a.size > 3)
        k_median_wide<<<(a.grid_count + 255) / 256, 256, 0, s>>>(a);
    else
        k_median<<<(a.grid_count + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Final affine ZNCC at the solved warp (shaders/affine_zncc.wgsl:57-99).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_affine_zncc(ZnccArgs a) {
    const int lane = threadIdx.x & 63;
    const int index = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int N = a.grid_w * a.grid_h;
    if (index >= N) return;
    const float* p = a.state + (size_t)index * 8;
    if (p[6] != 1.0f) {
        if (lane == 0) a.out[index] = -2.0f;
        return;
    }
    const int gx = index % a.grid_w, gy = index / a.grid_w;
    const float half = ((float)a.S - 1.0f) * 0.5f;
    float sf = 0, sg = 0, sff = 0, sgg = 0, sfg = 0, n = 0;
    for (int k = lane; k < a.S * a.S; k += 64) {
        int cx = k % a.S, cy = k / a.S;
        int px = gx * a.D + cx, py = gy * a.D + cy;
        float dx = (float)cx - half, dy = (float)cy - half;
        float tx = (float)px + (p[0] + (p[1] * dx + p[2] * dy));
        float ty = (float)py + (p[3] + (p[4] * dx + p[5] * dy));
        if (tx >= 2.0f && ty >= 2.0f && tx <= (float)(a.width - 3) && ty <= (float)(a.height - 3)) {
            float fx = floorf(tx), fy = floorf(ty);
            int ix = (int)fx, iy = (int)fy;
            float ax = tx - fx, ay = ty - fy;
            auto raw = [&](int x, int y) { return (float)a.target[(size_t)y * a.width + x]; };
            float g;
            if (a.interpolation == kLinear) {
                float v00 = raw(ix, iy) / 255.0f, v10 = raw(ix + 1, iy) / 255.0f;
                float v01 = raw(ix, iy + 1) / 255.0f, v11 = raw(ix + 1, iy + 1) / 255.0f;
                float top = v00 * (1.0f - ax) + v10 * ax, bottom = v01 * (1.0f - ax) + v11 * ax;
                g = 255.0f * (top * (1.0f - ay) + bottom * ay);
            } else {
                float wx[4], wy[4];
                float t2x = ax * ax, t3x = t2x * ax, t2y = ay * ay, t3y = t2y * ay;
                if (a.interpolation == kKeys) {
                    wx[0] = -0.5f * ax + t2x - 0.5f * t3x; wx[1] = 1.0f - 2.5f * t2x + 1.5f * t3x;
                    wx[2] = 0.5f * ax + 2.0f * t2x - 1.5f * t3x; wx[3] = -0.5f * t2x + 0.5f * t3x;
                    wy[0] = -0.5f * ay + t2y - 0.5f * t3y; wy[1] = 1.0f - 2.5f * t2y + 1.5f * t3y;
                    wy[2] = 0.5f * ay + 2.0f * t2y - 1.5f * t3y; wy[3] = -0.5f * t2y + 0.5f * t3y;
                } else {
                    float ix1 = 1.0f - ax, iy1 = 1.0f - ay;
                    wx[0] = ix1 * ix1 * ix1 / 6.0f; wx[1] = (3.0f * t3x - 6.0f * t2x + 4.0f) / 6.0f;
                    wx[2] = (-3.0f * t3x + 3.0f * t2x + 3.0f * ax + 1.0f) / 6.0f; wx[3] = t3x / 6.0f;
                    wy[0] = iy1 * iy1 * iy1 / 6.0f; wy[1] = (3.0f * t3y - 6.0f * t2y + 4.0f) / 6.0f;
                    wy[2] = (-3.0f * t3y + 3.0f * t2y + 3.0f * ay + 1.0f) / 6.0f; wy[3] = t3y / 6.0f;
                }
                g = 0.0f;
                for (int j = 0; j < 4; ++j)
                    for (int i = 0; i < 4; ++i) {
                        float pix = a.interpolation == kBspline ? a.coeff[(size_t)(iy + j - 1) * a.width + ix + i - 1]
                                                                : raw(ix + i - 1, iy + j - 1);
                        g += wx[i] * wy[j] * pix;
                    }
            }
            float f = a.reference[(size_t)py * a.width + px].x;
            sf += f; sg += g; sff += f * f; sgg += g * g; sfg += f * g; n += 1.0f;
        }
    }
    sf = wave_sum(sf); sg = wave_sum(sg); sff = wave_sum(sff); sgg = wave_sum(sgg); sfg = wave_sum(sfg); n = wave_sum(n);
    if (lane == 0) {
        float vf = sff - sf * sf / fmaxf(n, 1.0f), vg = sgg - sg * sg / fmaxf(n, 1.0f);
        a.out[index] = (n == (float)(a.S * a.S) && vf > 1e-9f && vg > 1e-9f)
                           ? fminf(fmaxf((sfg - sf * sg / n) / sqrtf(vf * vg), -1.0f), 1.0f)
                           : -2.0f;
    }
}

hipError_t launch_affine_zncc(const ZnccArgs& a, hipStream_t s) {
    int N = a.grid_w * a.grid_h;
    if (N <= 0) return hipSuccess;
    k_affine_zncc<<<(N + 3) / 4, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace ldic
