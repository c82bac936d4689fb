// ROI planning on the device: the static per-level tables that the solver, the
// outlier pass and the seed transfer read (the reference builds them on the host,
// live_dic/roi.py:63-97, 178-263, 266-355, 358-437).
//
// - Window selection: a subset is selected when its S x S window holds no masked
//   pixel, read from a summed-area table of masked pixels.
// - Neighbour links: two subsets are linked when both are selected and the closed
//   pixel box around the segment joining their centres holds no masked pixel; 3x3
//   link bits (bit (dy+1)*3 + dx+1), and for wider medians 81 bits in four 24-bit
//   float words (bit (dy+4)*9 + dx+4).
// - Partial supports (roi_policy="partial"): for a window that is not entirely
//   allowed, the 4-connected component of allowed pixels grown from its centre (an
//   even width needs all four central pixels), optionally cut to the pixels whose
//   segment to the centre crosses only support pixels (visibility).  One wave per
//   window, one 64-bit row per lane: the component grows by OR-ing each row with
//   its shifts and its neighbours' rows until no row changes.
// - Seed links between pyramid levels: corner bits 0-3 (each coarse interpolation
//   corner usable), bit 4 (destination selected), and from bit 5 the one-based
//   index of the nearest usable coarse subset when no corner is.
// - Reduced masks: a coarse pixel is allowed when all 25 reflected source pixels
//   of its 5x5 window at (2x, 2y) are.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ldic_kernels.h"
#include "livedic.h"

namespace ldic {
namespace {

constexpr int kMaxPartialSubset = 64;  // one 64-bit row per lane

__device__ __forceinline__ int reflect(int i, int n) {
    if (i < 0) return -i;
    if (i > n - 1) return 2 * (n - 1) - i;
    return i;
}

__global__ void k_reduce_mask(const uint8_t* src, int w, int h, uint8_t* dst) {
    const int cw = w / 2, ch = h / 2;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cw || y >= ch) return;
    uint8_t all = 1;
    for (int dy = 0; dy < 5; ++dy) {
        const uint8_t* row = src + (size_t)reflect(2 * y + dy - 2, h) * w;
        for (int dx = 0; dx < 5; ++dx) all &= row[reflect(2 * x + dx - 2, w)] ? 1 : 0;
    }
    dst[(size_t)y * cw + x] = all;
}

// Summed-area table of masked pixels, (h + 1) x (w + 1) int32, first row and column zero.
// Rows: one workgroup per row, a running prefix over 256-wide chunks.
__global__ void __launch_bounds__(256) k_sat_rows(const uint8_t* mask, int w, int h, int* sat) {
    __shared__ int part[256];
    const int y = blockIdx.x, t = threadIdx.x;
    const size_t pitch = (size_t)w + 1;
    int* out = sat + (size_t)(y + 1) * pitch;
    if (t == 0) out[0] = 0;
    int carry = 0;
    for (int x0 = 0; x0 < w; x0 += 256) {
        const int x = x0 + t;
        int v = (x < w && !mask[(size_t)y * w + x]) ? 1 : 0;
        part[t] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            const int add = t >= off ? part[t - off] : 0;
            __syncthreads();
            part[t] += add;
            __syncthreads();
        }
        if (x < w) out[x + 1] = carry + part[t];
        carry += part[255];
        __syncthreads();
    }
}

// Columns: one thread per column, a running sum down the rows (coalesced across threads).
__global__ void k_sat_cols(int* sat, int w, int h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t pitch = (size_t)w + 1;
    if (x > w) return;
    sat[x] = 0;
    int acc = 0;
    for (int y = 1; y <= h; ++y) {
        acc += sat[(size_t)y * pitch + x];
        sat[(size_t)y * pitch + x] = acc;
    }
}

struct Sat {
    const int* t;
    int w, h;  // mask size; t is (h + 1) x (w + 1)
    __device__ int at(int x, int y) const { return t[(size_t)y * (w + 1) + x]; }
    // The half-open box [x0, x1) x [y0, y1) holds no masked pixel.
    __device__ bool clear(int x0, int y0, int x1, int y1) const {
        return at(x1, y1) - at(x1, y0) - at(x0, y1) + at(x0, y0) == 0;
    }
    // The closed pixel box around the segment (ax, ay)-(bx, by), clipped to the image.
    __device__ bool segment_clear(double ax, double ay, double bx, double by) const {
        const int x0 = max(0, (int)ceil(fmin(ax, bx) - 0.5));
        const int y0 = max(0, (int)ceil(fmin(ay, by) - 0.5));
        const int x1 = min(w, (int)floor(fmax(ax, bx) + 0.5) + 1);
        const int y1 = min(h, (int)floor(fmax(ay, by) + 0.5) + 1);
        return clear(x0, y0, x1, y1);
    }
};

__global__ void k_select_windows(Sat sat, int gw, int gh, int S, int D, uint8_t* selected) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= gw * gh) return;
    const int x = (i % gw) * D, y = (i / gw) * D;
    selected[i] = sat.clear(x, y, x + S, y + S) ? 1 : 0;
}

__global__ void k_links(Sat sat, const uint8_t* selected, int gw, int gh, int S, int D, int radius, uint32_t* bits,
                        float* words) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= gw * gh) return;
    const int gx = i % gw, gy = i / gw;
    const double half = 0.5 * (S - 1);
    const double cx = gx * (double)D + half, cy = gy * (double)D + half;
    const bool self = selected[i] != 0;
    uint32_t b = 0;
    float wd[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int dy = -radius; dy <= radius; ++dy)
        for (int dx = -radius; dx <= radius; ++dx) {
            const int nx = gx + dx, ny = gy + dy;
            bool ok = self && nx >= 0 && ny >= 0 && nx < gw && ny < gh && selected[ny * gw + nx] &&
                      sat.segment_clear(cx, cy, cx + dx * (double)D, cy + dy * (double)D);
            if (!ok) continue;
            if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) b |= 1u << ((dy + 1) * 3 + dx + 1);
            if (words) {
                const int bit = (dy + 4) * 9 + dx + 4;
                wd[bit / 24] += (float)(1u << (bit % 24));
            }
        }
    bits[i] = b;
    if (words)
        for (int k = 0; k < 4; ++k) words[(size_t)i * 4 + k] = wd[k];
}

// Cells touched by the segment from the window centre to each pixel centre (slab test
// per axis; touching an edge or corner counts): row p holds ceil(S*S/32) words,
// cell c in bit c % 32 of word c / 32.
__global__ void k_visibility_paths(int S, int nwords, uint32_t* paths) {
    const int n = S * S;
    const int p = blockIdx.x;
    const double centre = 0.5 * (S - 1);
    const double px = p % S, py = p / S;
    const double dxp = px - centre, dyp = py - centre;
    for (int w = threadIdx.x; w < nwords; w += blockDim.x) {
        uint32_t word = 0;
        for (int b = 0; b < 32; ++b) {
            const int c = w * 32 + b;
            if (c >= n) break;
            const double coord[2] = {(double)(c % S), (double)(c / S)};
            const double d[2] = {dxp, dyp};
            double lo = 0.0, hi = 1.0;
            for (int axis = 0; axis < 2; ++axis) {
                if (d[axis] == 0.0) {
                    if (fabs(coord[axis] - centre) > 0.5) hi = -1.0;
                    continue;
                }
                const double a = (coord[axis] - 0.5 - centre) / d[axis];
                const double e = (coord[axis] + 0.5 - centre) / d[axis];
                lo = fmax(lo, fmin(a, e));
                hi = fmin(hi, fmax(a, e));
            }
            if (lo <= hi) word |= 1u << b;
        }
        paths[(size_t)p * nwords + w] = word;
    }
}

__device__ __forceinline__ uint64_t lane_shift(uint64_t v, int from) {
    // v of lane `from`, or 0 outside the wave.
    const uint32_t lo = __shfl((uint32_t)v, from & 63), hi = __shfl((uint32_t)(v >> 32), from & 63);
    return (from < 0 || from > 63) ? 0ull : ((uint64_t)hi << 32) | lo;
}

// The 4-connected component of `allowed` grown from the centre; empty unless the
// central pixel (all four for an even width) is allowed.  Lane y holds row y.
__device__ uint64_t central_component(uint64_t allowed, int S, int lane) {
    const int lo = (S - 1) / 2, hi = S / 2;
    const uint64_t centre_bits = ((2ull << hi) - 1) & ~((1ull << lo) - 1);
    const bool ok = lane < lo || lane > hi || (allowed & centre_bits) == centre_bits;
    if (!__all(ok)) return 0ull;
    uint64_t comp = lane == lo ? (1ull << lo) : 0ull;
    for (;;) {
        const uint64_t up = lane_shift(comp, lane - 1), down = lane_shift(comp, lane + 1);
        const uint64_t grown = (comp | (comp << 1) | (comp >> 1) | up | down) & allowed;
        const bool changed = grown != comp;
        comp = grown;
        if (!__any(changed)) break;
    }
    return comp;
}

// Row-major packing of the rows (lane y, S bits each) into words in LDS.
__device__ void pack_rows(uint64_t row, int S, int lane, int nwords, uint32_t* words) {
    for (int k = lane; k < nwords; k += 64) words[k] = 0u;
    __syncthreads();
    if (lane < S && row) {
        for (int x = 0; x < S; ++x)
            if ((row >> x) & 1ull) {
                const int p = lane * S + x;
                atomicOr(&words[p >> 5], 1u << (p & 31));
            }
    }
    __syncthreads();
}

struct PartialArgs {
    const uint8_t* mask;
    int w, S, D, gw, nwords, visibility;
    const uint32_t* paths;      // visibility paths (visibility only)
    const uint32_t* candidates; // flat subset indices
    int count;
    uint32_t* words;            // count * nwords
    float* boxes;               // count * 4
    int* pixels;                // count: support pixels (0: no support)
};

// One workgroup (one wave) per candidate window.
__global__ void __launch_bounds__(64) k_partial_supports(PartialArgs a) {
    __shared__ uint32_t supp[(kMaxPartialSubset * kMaxPartialSubset + 31) / 32];
    __shared__ uint32_t vis[(kMaxPartialSubset * kMaxPartialSubset + 31) / 32];
    const int j = blockIdx.x, lane = threadIdx.x, S = a.S;
    if (j >= a.count) return;
    const uint32_t idx = a.candidates[j];
    const int x0 = (int)(idx % a.gw) * a.D, y0 = (int)(idx / a.gw) * a.D;
    uint64_t allowed = 0;
    if (lane < S) {
        const uint8_t* row = a.mask + (size_t)(y0 + lane) * a.w + x0;
        for (int x = 0; x < S; ++x)
            if (row[x]) allowed |= 1ull << x;
    }
    uint64_t comp = central_component(allowed, S, lane);
    if (a.visibility && __any(comp != 0ull)) {
        pack_rows(comp, S, lane, a.nwords, supp);
        for (int k = lane; k < a.nwords; k += 64) vis[k] = 0u;
        __syncthreads();
        for (int p = lane; p < S * S; p += 64) {
            const uint32_t* path = a.paths + (size_t)p * a.nwords;
            bool visible = true;
            for (int k = 0; k < a.nwords && visible; ++k) visible = (path[k] & ~supp[k]) == 0u;
            if (visible) atomicOr(&vis[p >> 5], 1u << (p & 31));
        }
        __syncthreads();
        uint64_t seen = 0;
        if (lane < S)
            for (int x = 0; x < S; ++x) {
                const int p = lane * S + x;
                if ((vis[p >> 5] >> (p & 31)) & 1u) seen |= 1ull << x;
            }
        comp = central_component(comp & seen, S, lane);
    }
    pack_rows(comp, S, lane, a.nwords, supp);
    for (int k = lane; k < a.nwords; k += 64) a.words[(size_t)j * a.nwords + k] = supp[k];
    // Box and pixel count by wave reductions.
    int n = __popcll(comp);
    int ymin = comp ? lane : 1 << 20, ymax = comp ? lane : -1;
    int xmin = comp ? __ffsll((unsigned long long)comp) - 1 : 1 << 20;
    int xmax = comp ? 63 - __clzll((long long)comp) : -1;
    for (int off = 32; off > 0; off >>= 1) {
        n += __shfl_xor(n, off);
        ymin = min(ymin, __shfl_xor(ymin, off));
        ymax = max(ymax, __shfl_xor(ymax, off));
        xmin = min(xmin, __shfl_xor(xmin, off));
        xmax = max(xmax, __shfl_xor(xmax, off));
    }
    if (lane == 0) {
        a.pixels[j] = n;
        float* box = a.boxes + (size_t)j * 4;
        box[0] = n ? (float)xmin : 0.0f;
        box[1] = n ? (float)ymin : 0.0f;
        box[2] = n ? (float)xmax : 0.0f;
        box[3] = n ? (float)ymax : 0.0f;
    }
}

struct SeedLinkArgs {
    Sat sat;                    // of the finest (reference) mask
    const uint8_t* src_sel;     // source level selection
    int sgw, sgh;
    const uint8_t* dst_sel;     // destination level selection
    int dgw, dgh;
    int src_scale, dst_scale, S, D;
    uint32_t* links;
};

__global__ void k_seed_links(SeedLinkArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.dgw * a.dgh) return;
    const int x = i % a.dgw, y = i / a.dgw;
    const double c = a.S / 2.0, half = 0.5 * (a.S - 1), D = a.D;
    const double ss = a.src_scale, ds = a.dst_scale;
    // Destination centre in source-level coordinates, and the interpolation cell.
    const double sx = (ds * (c + x * D) + (ss - ds) / 2.0) / ss;
    const double sy = (ds * (c + y * D) + (ss - ds) / 2.0) / ss;
    const int lx = (int)fmin(fmax(floor((sx - c) / D), 0.0), (double)max(0, a.sgw - 2));
    const int ly = (int)fmin(fmax(floor((sy - c) / D), 0.0), (double)max(0, a.sgh - 2));
    const bool dsel = a.dst_sel[i] != 0;
    const double fx = ds * (half + x * D), fy = ds * (half + y * D);
    uint32_t bits = dsel ? 16u : 0u;
    for (int corner = 0; corner < 4; ++corner) {
        const int nx = lx + (corner & 1), ny = ly + (corner >> 1);
        if (nx >= a.sgw || ny >= a.sgh || !dsel) continue;
        if (!a.src_sel[ny * a.sgw + nx]) continue;
        if (a.sat.segment_clear(fx, fy, ss * (half + nx * D), ss * (half + ny * D))) bits |= 1u << corner;
    }
    if (a.src_scale > a.dst_scale && (long long)a.sgw * a.sgh < (1ll << 27) - 1 && dsel && (bits & 15u) == 0u) {
        double best = INFINITY;
        uint32_t chosen = 0;
        for (int dy = -2; dy < 4; ++dy)
            for (int dx = -2; dx < 4; ++dx) {
                const int nx = lx + dx, ny = ly + dy;
                if (nx < 0 || ny < 0 || nx >= a.sgw || ny >= a.sgh) continue;
                const double ex = sx - (c + nx * D), ey = sy - (c + ny * D);
                const double dist = ex * ex + ey * ey;
                if (!a.src_sel[ny * a.sgw + nx] || !(dist < best)) continue;
                if (!a.sat.segment_clear(fx, fy, ss * (half + nx * D), ss * (half + ny * D))) continue;
                best = dist;
                chosen = (uint32_t)(ny * a.sgw + nx + 1);
            }
        bits |= chosen << 5;
    }
    a.links[i] = bits;
}

inline unsigned blocks_of(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace
}  // namespace ldic

using namespace ldic;

struct ldic_roi_planner {
    int device = 0;
    int levels = 0;
    struct Level {
        int w = 0, h = 0;
        uint8_t* mask = nullptr;   // device
        int* sat = nullptr;        // device, (h + 1) x (w + 1)
        bool all_allowed = false;
        uint8_t* selected = nullptr;  // device, after plan()
        int gw = 0, gh = 0, S = 0, D = 0;
        // Partial supports of the last plan(): device arrays and their count.
        uint32_t* partial_indices = nullptr;
        uint32_t* support_words = nullptr;
        float* support_boxes = nullptr;
        int partial_count = 0, nwords = 0;
        std::vector<int> support_pixels;  // host, per partial subset
    } lv[LDIC_MAX_LEVELS];
    std::vector<void*> owned;

    template <class T>
    hipError_t alloc(T** p, size_t count) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count * sizeof(T), 16));
        if (e == hipSuccess) owned.push_back(q);
        *p = static_cast<T*>(q);
        return e;
    }
    void release(void* p) {
        if (!p) return;
        auto it = std::find(owned.begin(), owned.end(), p);
        if (it != owned.end()) owned.erase(it);
        (void)hipFree(p);
    }
};

namespace {

int roi_error(int code, const char* message) { return report_error(code, message); }

#define ROI_HIP(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) return roi_error(LDIC_ERR_DEVICE, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

extern "C" {

int ldic_roi_planner_create(const uint8_t* mask, int32_t width, int32_t height, int32_t levels,
                            int32_t device_ordinal, ldic_roi_planner** out) {
    if (!out || !mask || width < 1 || height < 1 || levels < 1 || levels > LDIC_MAX_LEVELS)
        return roi_error(LDIC_ERR_ARGUMENT, "invalid ROI planner arguments");
    *out = nullptr;
    int count = 0;
    ROI_HIP(hipGetDeviceCount(&count));
    if (device_ordinal < 0 || device_ordinal >= count) return roi_error(LDIC_ERR_ARGUMENT, "invalid device ordinal");
    ROI_HIP(hipSetDevice(device_ordinal));
    auto* p = new ldic_roi_planner();
    p->device = device_ordinal;
    p->levels = levels;
    auto fail = [&](int rc) {
        ldic_roi_planner_destroy(p);
        return rc;
    };
    int w = width, h = height;
    for (int k = 0; k < levels; ++k) {
        if (k) {
            w /= 2;
            h /= 2;
        }
        if (w < 1 || h < 1) return fail(roi_error(LDIC_ERR_ARGUMENT, "the mask is too small for the pyramid"));
        auto& L = p->lv[k];
        L.w = w;
        L.h = h;
        hipError_t e = p->alloc(&L.mask, (size_t)w * h);
        if (e == hipSuccess && k == 0) e = hipMemcpy(L.mask, mask, (size_t)w * h, hipMemcpyHostToDevice);
        if (e == hipSuccess && k > 0) {
            const auto& U = p->lv[k - 1];
            k_reduce_mask<<<dim3(blocks_of((size_t)w), h), 256>>>(U.mask, U.w, U.h, L.mask);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = p->alloc(&L.sat, (size_t)(w + 1) * (h + 1));
        if (e == hipSuccess) {
            k_sat_rows<<<h, 256>>>(L.mask, w, h, L.sat);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            k_sat_cols<<<blocks_of((size_t)w + 1), 256>>>(L.sat, w, h);
            e = hipGetLastError();
        }
        int masked = 0;
        if (e == hipSuccess) e = hipMemcpy(&masked, L.sat + (size_t)h * (w + 1) + w, sizeof(int), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail(roi_error(LDIC_ERR_DEVICE, hipGetErrorString(e)));
        L.all_allowed = masked == 0;
    }
    *out = p;
    return LDIC_OK;
}

int ldic_roi_planner_mask(ldic_roi_planner* p, int32_t level, int32_t* width, int32_t* height, uint8_t* out_mask,
                          int32_t* all_allowed) {
    if (!p || level < 0 || level >= p->levels) return roi_error(LDIC_ERR_ARGUMENT, "level out of range");
    auto& L = p->lv[level];
    if (width) *width = L.w;
    if (height) *height = L.h;
    if (all_allowed) *all_allowed = L.all_allowed ? 1 : 0;
    if (out_mask) {
        ROI_HIP(hipSetDevice(p->device));
        ROI_HIP(hipMemcpy(out_mask, L.mask, (size_t)L.w * L.h, hipMemcpyDeviceToHost));
    }
    return LDIC_OK;
}

int ldic_roi_planner_plan(ldic_roi_planner* p, int32_t level, const ldic_roi_plan_options* o, uint8_t* out_selected,
                          uint32_t* out_neighbour_bits, float* out_median_words, float* out_coverage,
                          int32_t* out_partial_count) {
    if (!p || !o || level < 0 || level >= p->levels || !out_selected || !out_neighbour_bits)
        return roi_error(LDIC_ERR_ARGUMENT, "invalid ROI plan arguments");
    auto& L = p->lv[level];
    const int S = o->subset_size, D = o->subset_stride;
    if (S < 4 || D < 1 || D > S || L.w < S || L.h < S)
        return roi_error(LDIC_ERR_ARGUMENT, "the subset does not fit the level's mask");
    if (o->median_filter_size < 3 || o->median_filter_size > 9 || !(o->median_filter_size & 1))
        return roi_error(LDIC_ERR_ARGUMENT, "median_filter_size must be 3, 5, 7 or 9");
    if (o->partial && S > kMaxPartialSubset)
        return roi_error(LDIC_ERR_ARGUMENT, "partial ROI supports are planned for subsets up to 64 pixels");
    ROI_HIP(hipSetDevice(p->device));
    const int gw = 1 + (L.w - S) / D, gh = 1 + (L.h - S) / D;
    const size_t n = (size_t)gw * gh;
    p->release(L.selected);
    p->release(L.partial_indices);
    p->release(L.support_words);
    p->release(L.support_boxes);
    L.selected = nullptr;
    L.partial_indices = L.support_words = nullptr;
    L.support_boxes = nullptr;
    L.partial_count = 0;
    L.support_pixels.clear();
    L.gw = gw;
    L.gh = gh;
    L.S = S;
    L.D = D;
    ROI_HIP(p->alloc(&L.selected, n));
    const Sat sat{L.sat, L.w, L.h};
    k_select_windows<<<blocks_of(n), 256>>>(sat, gw, gh, S, D, L.selected);
    ROI_HIP(hipGetLastError());
    std::vector<uint8_t> sel(n);
    ROI_HIP(hipMemcpy(sel.data(), L.selected, n, hipMemcpyDeviceToHost));
    std::vector<float> coverage(n);
    for (size_t i = 0; i < n; ++i) coverage[i] = sel[i] ? 1.0f : 0.0f;
    if (o->partial) {
        // Candidates: every window that is not entirely allowed, in index order.
        std::vector<uint32_t> cand;
        for (size_t i = 0; i < n; ++i)
            if (!sel[i]) cand.push_back((uint32_t)i);
        const int nwords = (S * S + 31) / 32;
        L.nwords = nwords;
        if (!cand.empty()) {
            uint32_t *dcand = nullptr, *dwords = nullptr, *paths = nullptr;
            float* dboxes = nullptr;
            int* dpix = nullptr;
            ROI_HIP(p->alloc(&dcand, cand.size()));
            ROI_HIP(p->alloc(&dwords, cand.size() * nwords));
            ROI_HIP(p->alloc(&dboxes, cand.size() * 4));
            ROI_HIP(p->alloc(&dpix, cand.size()));
            ROI_HIP(hipMemcpy(dcand, cand.data(), cand.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
            if (o->visibility) {
                ROI_HIP(p->alloc(&paths, (size_t)S * S * nwords));
                k_visibility_paths<<<S * S, 64>>>(S, nwords, paths);
                ROI_HIP(hipGetLastError());
            }
            PartialArgs a{L.mask, L.w, S, D, gw, nwords, o->visibility ? 1 : 0, paths, dcand, (int)cand.size(),
                          dwords, dboxes, dpix};
            k_partial_supports<<<(unsigned)cand.size(), 64>>>(a);
            ROI_HIP(hipGetLastError());
            std::vector<int> pix(cand.size());
            ROI_HIP(hipMemcpy(pix.data(), dpix, pix.size() * sizeof(int), hipMemcpyDeviceToHost));
            std::vector<uint32_t> keep;
            for (size_t j = 0; j < cand.size(); ++j)
                if (pix[j] > 0) keep.push_back((uint32_t)j);
            const int m = (int)keep.size();
            L.partial_count = m;
            if (m) {
                // Compact the supported candidates (their order is the subset order).
                std::vector<uint32_t> words((size_t)cand.size() * nwords), idx(m), cw((size_t)m * nwords);
                std::vector<float> boxes(cand.size() * 4), cb((size_t)m * 4);
                ROI_HIP(hipMemcpy(words.data(), dwords, words.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
                ROI_HIP(hipMemcpy(boxes.data(), dboxes, boxes.size() * sizeof(float), hipMemcpyDeviceToHost));
                for (int k = 0; k < m; ++k) {
                    const uint32_t j = keep[k];
                    idx[k] = cand[j];
                    std::memcpy(&cw[(size_t)k * nwords], &words[(size_t)j * nwords], nwords * sizeof(uint32_t));
                    std::memcpy(&cb[(size_t)k * 4], &boxes[(size_t)j * 4], 4 * sizeof(float));
                    sel[cand[j]] = 1;
                    // float64 quotient rounded once to float32, as the reference stores it.
                    coverage[cand[j]] = (float)((double)pix[j] / (double)(S * S));
                    L.support_pixels.push_back(pix[j]);
                }
                ROI_HIP(p->alloc(&L.partial_indices, (size_t)m));
                ROI_HIP(p->alloc(&L.support_words, (size_t)m * nwords));
                ROI_HIP(p->alloc(&L.support_boxes, (size_t)m * 4));
                ROI_HIP(hipMemcpy(L.partial_indices, idx.data(), idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
                ROI_HIP(hipMemcpy(L.support_words, cw.data(), cw.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
                ROI_HIP(hipMemcpy(L.support_boxes, cb.data(), cb.size() * sizeof(float), hipMemcpyHostToDevice));
                ROI_HIP(hipMemcpy(L.selected, sel.data(), n, hipMemcpyHostToDevice));
            }
            p->release(dcand);
            p->release(dwords);
            p->release(dboxes);
            p->release(dpix);
            p->release(paths);
        }
    }
    uint32_t* dbits = nullptr;
    float* dwords = nullptr;
    const bool wide = o->median_filter_size > 3;
    ROI_HIP(p->alloc(&dbits, n));
    if (wide) ROI_HIP(p->alloc(&dwords, n * 4));
    k_links<<<blocks_of(n), 256>>>(sat, L.selected, gw, gh, S, D, wide ? o->median_filter_size / 2 : 1, dbits,
                                   dwords);
    ROI_HIP(hipGetLastError());
    ROI_HIP(hipMemcpy(out_neighbour_bits, dbits, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (wide && out_median_words) ROI_HIP(hipMemcpy(out_median_words, dwords, n * 4 * sizeof(float), hipMemcpyDeviceToHost));
    p->release(dbits);
    p->release(dwords);
    std::memcpy(out_selected, sel.data(), n);
    if (out_coverage) std::memcpy(out_coverage, coverage.data(), n * sizeof(float));
    if (out_partial_count) *out_partial_count = L.partial_count;
    return LDIC_OK;
}

int ldic_roi_planner_partial(ldic_roi_planner* p, int32_t level, uint32_t* out_indices, uint32_t* out_support_words,
                             float* out_support_boxes) {
    if (!p || level < 0 || level >= p->levels) return roi_error(LDIC_ERR_ARGUMENT, "level out of range");
    auto& L = p->lv[level];
    if (!L.partial_count) return LDIC_OK;
    if (!out_indices || !out_support_words || !out_support_boxes)
        return roi_error(LDIC_ERR_ARGUMENT, "output buffers are required");
    ROI_HIP(hipSetDevice(p->device));
    const size_t m = (size_t)L.partial_count;
    ROI_HIP(hipMemcpy(out_indices, L.partial_indices, m * sizeof(uint32_t), hipMemcpyDeviceToHost));
    ROI_HIP(hipMemcpy(out_support_words, L.support_words, m * L.nwords * sizeof(uint32_t), hipMemcpyDeviceToHost));
    ROI_HIP(hipMemcpy(out_support_boxes, L.support_boxes, m * 4 * sizeof(float), hipMemcpyDeviceToHost));
    return LDIC_OK;
}

int ldic_roi_planner_seed_links(ldic_roi_planner* p, int32_t source_level, int32_t destination_level,
                                uint32_t* out_links) {
    if (!p || !out_links || source_level < 0 || source_level >= p->levels || destination_level < 0 ||
        destination_level >= p->levels)
        return roi_error(LDIC_ERR_ARGUMENT, "invalid seed-link arguments");
    auto& src = p->lv[source_level];
    auto& dst = p->lv[destination_level];
    if (!src.selected || !dst.selected || src.S != dst.S || src.D != dst.D)
        return roi_error(LDIC_ERR_STATE, "plan both levels with the same subset geometry first");
    ROI_HIP(hipSetDevice(p->device));
    const size_t n = (size_t)dst.gw * dst.gh;
    uint32_t* dl = nullptr;
    ROI_HIP(p->alloc(&dl, n));
    const auto& R = p->lv[0];
    SeedLinkArgs a{Sat{R.sat, R.w, R.h}, src.selected, src.gw, src.gh, dst.selected, dst.gw, dst.gh,
                   1 << source_level, 1 << destination_level, src.S, src.D, dl};
    k_seed_links<<<blocks_of(n), 256>>>(a);
    ROI_HIP(hipGetLastError());
    ROI_HIP(hipMemcpy(out_links, dl, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    p->release(dl);
    return LDIC_OK;
}

void ldic_roi_planner_destroy(ldic_roi_planner* p) {
    if (!p) return;
    (void)hipSetDevice(p->device);
    (void)hipDeviceSynchronize();
    for (void* q : p->owned) (void)hipFree(q);
    delete p;
}

}  // extern "C"
