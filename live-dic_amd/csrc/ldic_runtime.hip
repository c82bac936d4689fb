// Host runtime of liblivedic.so: engine, reference preparation, per-frame
// launch sequence, streaming slots and the C ABI declared in livedic.h.
//
// Memory plan (MI355X, 288 GB HBM): everything a frame touches is allocated
// once per reference -- reference planes, metadata and inverse Hessians per
// level, and per slot the target pyramid, coefficient planes, state, seed and
// outlier buffers plus pinned host staging.  A frame performs no allocation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "ldic_kernels.h"
#include "livedic.h"

using namespace ldic;

namespace {

thread_local std::string g_error;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

}  // namespace

int ldic::report_error(int code, const char* message) {
    g_error = message;
    return code;
}

namespace {

#define HIP_CHECK(expr)                                                                                \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return set_error(LDIC_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define CHECK(expr)            \
    do {                       \
        int r_ = (expr);       \
        if (r_ != LDIC_OK) return r_; \
    } while (0)

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// Device allocations owned by one reference generation or one slot.
struct Arena {
    std::vector<void*> device;
    std::vector<void*> pinned;
    template <class T>
    int alloc(T** out, size_t count) {
        void* p = nullptr;
        size_t bytes = count * sizeof(T);
        if (bytes == 0) bytes = sizeof(T);
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return set_error(LDIC_ERR_MEMORY, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        device.push_back(p);
        *out = static_cast<T*>(p);
        return LDIC_OK;
    }
    template <class T>
    int alloc_pinned(T** out, size_t count) {
        void* p = nullptr;
        hipError_t e = hipHostMalloc(&p, count * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) return set_error(LDIC_ERR_MEMORY, "hipHostMalloc failed: %s", hipGetErrorString(e));
        pinned.push_back(p);
        *out = static_cast<T*>(p);
        return LDIC_OK;
    }
    void release() {
        for (void* p : device) (void)hipFree(p);
        for (void* p : pinned) (void)hipHostFree(p);
        device.clear();
        pinned.clear();
    }
};

struct Level {
    int width = 0, height = 0, gw = 0, gh = 0;
    uint8_t* image = nullptr;  // reference image of this level
    float4* reference = nullptr;
    float* metadata = nullptr;
    float* inverse = nullptr;
    uint8_t* valid = nullptr;
    uint8_t* selected = nullptr;
    float* roi_words = nullptr;
    uint32_t* seed_links = nullptr;
    bool roi = false;
    // Partial-support subsets (roi_policy="partial"): list, support bits, boxes.
    uint32_t* partial_list = nullptr;
    int partial_count = 0;
    uint32_t* support = nullptr;
    float4* support_boxes = nullptr;
    ldic_level_config config{30, 1e-3f, 0.8f};
    int count() const { return gw * gh; }
};

struct SlotLevel {
    uint8_t* target = nullptr;
    float* coeff = nullptr;
    float* state = nullptr;
    float* filtered = nullptr;
    float* seed = nullptr;
    uint32_t* active = nullptr;
    float* shift = nullptr;         // ZNCC search: best shift per subset
    float* score = nullptr;         // ZNCC search: its score
    uint8_t* mask = nullptr;        // subsets to re-solve from a ZNCC seed
    float* retry_seed = nullptr;
    uint32_t* retry_active = nullptr;
    uint32_t* near_a = nullptr;     // neighbour re-seeding: nearest accepted subset (ping-pong)
    uint32_t* near_b = nullptr;
};

constexpr int kEventRing = 256;

struct Slot {
    Arena arena;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    // Per-frame solver events; frames queued since the last wait own
    // ring entries [ring_first, ring_first + ring_count).
    hipEvent_t ev_solve0[kEventRing] = {}, ev_solve1[kEventRing] = {};
    int ring_first = 0, ring_count = 0;
    double drained_solve_ms = 0.0;
    int drained_frames = 0;
    SlotLevel lv[LDIC_MAX_LEVELS];
    uint8_t* pinned_in = nullptr;
    uint8_t* pinned_ref = nullptr;
    uint8_t* pinned_valid = nullptr;
    float* pinned_out = nullptr;
    float* zncc = nullptr;
    float* pinned_zncc = nullptr;
    uint32_t* counts = nullptr;         // temporal: seeded coarsest subsets, merged finest subsets
    uint32_t* pinned_counts = nullptr;
    double* summary = nullptr;          // launch_field_summary of the finest field read back
    double* pinned_summary = nullptr;
    bool summary_ready = false;         // the pending frame's readback carries a summary
    int levels_used = 0;                // pyramid levels of the slot's last frame
    bool adaptive_frame = false;        // it ran under the adaptive pyramid
    bool feedback_pending = false;      // its counts have not yet set the next frame's levels
    bool pending = false;
    bool resident = false;
    double upload_ms = 0.0;
    double submitted_at = 0.0;
};

}  // namespace

struct ldic_engine {
    ldic_options opt{};
    // Batches (ldic_batch_create) that hold this engine's device addresses: while any
    // does, calls that would free, re-prepare, reconfigure or use slot 0 are refused.
    int batch_refs = 0;
    int device = 0;
    hipDeviceProp_t prop{};
    int levels = 0;
    Level lv[LDIC_MAX_LEVELS];
    Arena ref_arena;
    std::vector<Slot*> slots;
    float* fir = nullptr;       // 31 FIR taps on the device
    Arena const_arena;
    hipStream_t prep_stream = nullptr;
    int tile_cap = 0;
    size_t solve_lds = 0;
    bool ready = false;
    double gauss7[7];
    double weight_sum = 0.0;
    // Reference-preparation scratch (float64 planes of the finest level) and the
    // integration weights, kept for consecutive-pair registration.
    double *prep_g = nullptr, *prep_a = nullptr, *prep_b = nullptr, *prep_gx = nullptr, *prep_gy = nullptr;
    double* weights = nullptr;
    ldic_initialization init{LDIC_INIT_ZERO, 0, -2.0f, 0, 0xFFFFFFFEu};
    // Temporal history of the finest field and the projection links onto the coarsest grid.
    float* hist_state = nullptr;
    uint32_t* hist_ages = nullptr;
    // Velocity prediction: previous accepted measurement, its ages, the seed.
    float* prev_state = nullptr;
    uint32_t* prev_ages = nullptr;
    float* prediction = nullptr;
    // Fine correction: history of the coarsest level.
    float* coarse_state = nullptr;
    uint32_t* coarse_ages = nullptr;
    // Adaptive pyramid: levels of the next frame, reference-valid finest
    // subsets, and the links of a projection onto the finest grid.
    int adaptive_next = 0;
    int64_t adaptive_population = 0;
    std::vector<uint8_t> valid_host;    // finest reference_valid, partial admissions included
    uint32_t* fine_links = nullptr;
    uint32_t* temporal_links = nullptr;
    // Recorded after each frame's history merge; the next frame's projection,
    // on whichever slot, waits for it, so streams keep the chronological chain.
    hipEvent_t temporal_done = nullptr;
    bool temporal_pending = false;
};

namespace {

constexpr int kMaxSearchRadius = 64;
constexpr int kMaxSearchSubset = 255;   // keeps the uint32 dot-product sums exact
constexpr size_t kMaxSearchLds = 65536;

int output_floats(const ldic_engine* e) { return e->lv[0].count() * state_stride(e->opt.motion_model); }

void free_slots(ldic_engine* e) {
    for (Slot* s : e->slots) {
        if (s->stream) (void)hipStreamSynchronize(s->stream);
        s->arena.release();
        if (s->ev_begin) (void)hipEventDestroy(s->ev_begin);
        for (int i = 0; i < kEventRing; ++i) {
            if (s->ev_solve0[i]) (void)hipEventDestroy(s->ev_solve0[i]);
            if (s->ev_solve1[i]) (void)hipEventDestroy(s->ev_solve1[i]);
        }
        if (s->ev_end) (void)hipEventDestroy(s->ev_end);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
    }
    e->slots.clear();
}

int create_slot(ldic_engine* e, Slot** out) {
    Slot* s = new Slot();
    e->slots.push_back(s);
    HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreate(&s->ev_begin));
    for (int i = 0; i < kEventRing; ++i) {
        HIP_CHECK(hipEventCreate(&s->ev_solve0[i]));
        HIP_CHECK(hipEventCreate(&s->ev_solve1[i]));
    }
    HIP_CHECK(hipEventCreate(&s->ev_end));
    const int stride = state_stride(e->opt.motion_model);
    for (int k = 0; k < e->levels; ++k) {
        const Level& L = e->lv[k];
        SlotLevel& sl = s->lv[k];
        size_t px = (size_t)L.width * L.height;
        size_t n = (size_t)L.count();
        CHECK(s->arena.alloc(&sl.target, px));
        if (e->opt.interpolation == kBspline) CHECK(s->arena.alloc(&sl.coeff, px));
        CHECK(s->arena.alloc(&sl.state, n * stride));
        CHECK(s->arena.alloc(&sl.filtered, n * stride));
        CHECK(s->arena.alloc(&sl.seed, n * stride));
        CHECK(s->arena.alloc(&sl.active, n));
        CHECK(s->arena.alloc(&sl.shift, n * 2));
        CHECK(s->arena.alloc(&sl.score, n));
        CHECK(s->arena.alloc(&sl.mask, n));
        CHECK(s->arena.alloc(&sl.retry_seed, n * stride));
        CHECK(s->arena.alloc(&sl.retry_active, n));
        CHECK(s->arena.alloc(&sl.near_a, n));
        CHECK(s->arena.alloc(&sl.near_b, n));
    }
    CHECK(s->arena.alloc(&s->zncc, (size_t)e->lv[0].count()));
    CHECK(s->arena.alloc(&s->counts, 4));
    CHECK(s->arena.alloc_pinned(&s->pinned_counts, 4));
    std::memset(s->pinned_counts, 0, 4 * sizeof(uint32_t));
    CHECK(s->arena.alloc_pinned(&s->pinned_in, (size_t)e->lv[0].width * e->lv[0].height));
    CHECK(s->arena.alloc_pinned(&s->pinned_ref, (size_t)e->lv[0].width * e->lv[0].height));
    CHECK(s->arena.alloc_pinned(&s->pinned_valid, (size_t)e->lv[0].count()));
    CHECK(s->arena.alloc_pinned(&s->pinned_out, (size_t)output_floats(e)));
    CHECK(s->arena.alloc_pinned(&s->pinned_zncc, (size_t)e->lv[0].count()));
    CHECK(s->arena.alloc(&s->summary, (size_t)kSummaryDoubles));
    CHECK(s->arena.alloc_pinned(&s->pinned_summary, 8));
    *out = s;
    return LDIC_OK;
}

int ensure_slots(ldic_engine* e, int count) {
    HIP_CHECK(hipSetDevice(e->device));
    while ((int)e->slots.size() < count) {
        Slot* s = nullptr;
        CHECK(create_slot(e, &s));
    }
    return LDIC_OK;
}

const float* level_output(const ldic_engine* e, const SlotLevel& sl) {
    return e->opt.outlier_policy == kDisabled ? sl.state : sl.filtered;
}

SolveArgs solve_args(const ldic_engine* e, int k, const SlotLevel& sl, const float* seed, const uint32_t* active,
                     int preserve) {
    const Level& L = e->lv[k];
    SolveArgs a{};
    a.width = L.width;
    a.height = L.height;
    a.grid_w = L.gw;
    a.grid_h = L.gh;
    a.S = e->opt.subset_size;
    a.D = e->opt.subset_stride;
    a.iterations = L.config.iterations;
    a.tolerance = L.config.tolerance;
    a.zncc_min = L.config.zncc_min;
    a.inv_var = 1.0f / (e->opt.gaussian_sigma * e->opt.gaussian_sigma);
    a.roi_enabled = L.roi ? 1 : 0;
    a.preserve_disabled = preserve;
    a.tile_cap = e->tile_cap;
    a.reference = L.reference;
    a.metadata = L.metadata;
    a.inverse = L.inverse;
    a.coeff = sl.coeff;
    a.target = sl.target;
    a.seed = seed;
    a.active = active;
    a.state = sl.state;
    return a;
}

SearchArgs search_args(const ldic_engine* e, int k, const SlotLevel& sl, int radius) {
    const Level& L = e->lv[k];
    SearchArgs q{};
    q.width = L.width;
    q.height = L.height;
    q.grid_w = L.gw;
    q.grid_h = L.gh;
    q.S = e->opt.subset_size;
    q.D = e->opt.subset_stride;
    q.R = radius;
    q.reference = L.image;
    q.target = sl.target;
    q.shift = sl.shift;
    q.score = sl.score;
    q.seed_stride = state_stride(e->opt.motion_model);
    return q;
}

// The partial-support pass of a level (solve_partial / solve_partial_seeded /
// retry_partial_seeded, icgn.wgsl:1040-1085): the weighted solver variant over
// the level's partial list, uniform weights expressed as an infinite sigma.
int launch_partial(ldic_engine* e, int k, SlotLevel& sl, const float* seed, const uint32_t* active, int preserve,
                   hipStream_t st, float* zncc_out = nullptr, int patience = 0, int min_iters = 0) {
    const Level& L = e->lv[k];
    if (L.partial_count <= 0) return LDIC_OK;
    SolveArgs a = solve_args(e, k, sl, seed, active, preserve);
    a.partial_list = L.partial_list;
    a.partial_count = L.partial_count;
    a.support = L.support;
    a.support_boxes = L.support_boxes;
    a.zncc_out = zncc_out;
    a.patience = patience;
    a.min_iters = min_iters;
    if (e->opt.spatial_weighting == 0) a.inv_var = 0.0f;
    HIP_CHECK(launch_solve(e->opt.motion_model, e->opt.interpolation, 1, a, st));
    return LDIC_OK;
}

// Neighbour re-seeding passes of the finest level: unaccepted subsets restart
// from the warp of their nearest accepted subset and are re-solved; the rest
// keep their state.
int reseed_passes(ldic_engine* e, SlotLevel& sl, hipStream_t st) {
    const Level& L = e->lv[0];
    const int model = e->opt.motion_model;
    for (int pass = 0; pass < e->init.reseed_passes; ++pass) {
        HIP_CHECK(launch_reseed(L.gw, L.gh, model, (float)e->opt.subset_stride, sl.state, L.metadata,
                                L.roi ? L.selected : nullptr, sl.near_a, sl.near_b, sl.retry_seed, sl.retry_active,
                                st));
        SolveArgs a = solve_args(e, 0, sl, sl.retry_seed, sl.retry_active, 1);
        a.patience = e->init.divergence_patience;
        a.min_iters = e->init.divergence_min_iterations;
        HIP_CHECK(launch_solve(model, e->opt.interpolation, e->opt.spatial_weighting, a, st));
        CHECK(launch_partial(e, 0, sl, sl.retry_seed, sl.retry_active, 1, st, nullptr, a.patience, a.min_iters));
    }
    return LDIC_OK;
}

int has_partial(const ldic_engine* e) {
    for (int k = 0; k < e->levels; ++k)
        if (e->lv[k].partial_count > 0) return 1;
    return 0;
}

int launch_median_level(ldic_engine* e, int k, SlotLevel& sl, bool seed_available, hipStream_t st) {
    if (e->opt.outlier_policy == kDisabled) return LDIC_OK;
    const Level& L = e->lv[k];
    MedianArgs m{};
    m.grid_w = L.gw;
    m.grid_count = L.count();
    m.D = e->opt.subset_stride;
    m.model = e->opt.motion_model;
    m.policy = e->opt.outlier_policy;
    m.size = e->opt.median_filter_size;
    m.scope = e->opt.median_filter_scope;
    m.roi_enabled = L.roi ? 1 : 0;
    m.seed_available = seed_available ? 1 : 0;
    m.state = sl.state;
    m.metadata = L.metadata;
    m.roi_words = L.roi_words;
    m.filtered = sl.filtered;
    HIP_CHECK(launch_median(m, st));
    return LDIC_OK;
}

// The per-frame command sequence of full_pyramid_backend.py:1436-1569: target
// reductions fine-to-coarse, then per level (coarse-to-fine) seed transfer,
// prefilter, solve and outlier pass.  The target of level 0 is in place.
// Sum the solver times of all queued frames (waits for them).
int drain_ring(Slot* s) {
    for (int i = 0; i < s->ring_count; ++i) {
        int r = (s->ring_first + i) % kEventRing;
        HIP_CHECK(hipEventSynchronize(s->ev_solve1[r]));
        float ms = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&ms, s->ev_solve0[r], s->ev_solve1[r]));
        s->drained_solve_ms += ms;
        s->drained_frames += 1;
    }
    s->ring_first = (s->ring_first + s->ring_count) % kEventRing;
    s->ring_count = 0;
    return LDIC_OK;
}

// Forget every temporal history (the reset of full_pyramid_backend.py:660-677
// covers the finest, previous and coarse histories).
hipError_t reset_histories(ldic_engine* e, hipStream_t st) {
    const int model = e->opt.motion_model;
    e->adaptive_next = e->levels;
    hipError_t r = launch_temporal_reset(e->lv[0].count(), model, e->hist_state, e->hist_ages, st);
    if (r == hipSuccess) r = launch_temporal_reset(e->lv[0].count(), model, e->prev_state, e->prev_ages, st);
    if (r == hipSuccess)
        r = launch_temporal_reset(e->lv[e->levels - 1].count(), model, e->coarse_state, e->coarse_ages, st);
    return r;
}

SeedArgs edge_args(const ldic_engine* e, int k) {
    const Level& src = e->lv[k + 1];
    const Level& lev = e->lv[k];
    SeedArgs sa{};
    sa.src_grid_w = src.gw;
    sa.src_grid_h = src.gh;
    sa.dst_grid_w = lev.gw;
    sa.dst_count = lev.count();
    sa.src_scale = (float)(1 << (k + 1));
    sa.dst_scale = (float)(1 << k);
    sa.S = (float)e->opt.subset_size;
    sa.D = (float)e->opt.subset_stride;
    sa.model = e->opt.motion_model;
    sa.permissions = lev.seed_links;
    return sa;
}

int enqueue_frame(ldic_engine* e, Slot* s) {
    hipStream_t st = s->stream;
    const bool adaptive = e->init.kind == LDIC_INIT_TEMPORAL && e->init.adaptive_pyramid && e->levels == 2;
    if (adaptive)
        for (Slot* o : e->slots)
            if (o->pending)
                return set_error(LDIC_ERR_STATE, "the adaptive pyramid takes one frame at a time; retire the last one");
    const int L = adaptive ? e->adaptive_next : e->levels;
    s->levels_used = L;
    s->adaptive_frame = adaptive;
    s->feedback_pending = adaptive;
    if (s->ring_count == kEventRing) CHECK(drain_ring(s));
    const int ring = (s->ring_first + s->ring_count) % kEventRing;
    s->ring_count += 1;
    for (int k = 1; k < L; ++k)
        HIP_CHECK(launch_reduce(s->lv[k - 1].target, e->lv[k - 1].width, e->lv[k - 1].height, s->lv[k].target, st));
    const bool temporal = e->init.kind == LDIC_INIT_TEMPORAL;
    const bool velocity = temporal && e->init.temporal_prediction == 1;
    const bool correct = temporal && e->init.temporal_fine_correction != 0 && e->levels == 2;
    for (int k = L - 1; k >= 0; --k) {
        SlotLevel& sl = s->lv[k];
        const Level& lev = e->lv[k];
        const float* seed = nullptr;
        const uint32_t* active = nullptr;
        if (k < L - 1) {
            SeedArgs sa = edge_args(e, k);
            sa.track_availability = correct ? 1 : 0;
            sa.src_state = level_output(e, s->lv[k + 1]);
            sa.dst_seed = sl.seed;
            sa.dst_active = sl.active;
            HIP_CHECK(launch_transfer_seed(sa, st));
            if (k == 0 && correct) {
                // refine_temporal_seed (full_pyramid.wgsl:451-515).
                RefineArgs r{};
                r.edge = sa;
                r.edge.src_state = e->coarse_state;
                r.fine_history = e->hist_state;
                r.fine_ages = e->hist_ages;
                r.coarse_ages = e->coarse_ages;
                r.max_age = e->init.temporal_max_age;
                HIP_CHECK(launch_temporal_refine(r, st));
            }
            seed = sl.seed;
            active = sl.active;
        }
        if (k == L - 1 && e->init.kind == LDIC_INIT_TEMPORAL) {
            // Project the last accepted finest field onto this (coarsest) grid
            // (project_temporal_seed, full_pyramid.wgsl:437-449).
            const Level& fine = e->lv[0];
            if (e->temporal_pending) HIP_CHECK(hipStreamWaitEvent(st, e->temporal_done, 0));
            HIP_CHECK(launch_fill_u32(4, 0u, s->counts, st));
            if (velocity)
                HIP_CHECK(launch_temporal_predict(fine.count(), e->opt.motion_model, e->hist_state, e->prev_state,
                                                  e->hist_ages, e->prev_ages, e->prediction, s->counts + 2, st));
            SeedArgs sa{};
            sa.src_grid_w = fine.gw;
            sa.src_grid_h = fine.gh;
            sa.dst_grid_w = lev.gw;
            sa.dst_count = lev.count();
            sa.src_scale = 1.0f;
            sa.dst_scale = (float)(1 << k);
            sa.S = (float)e->opt.subset_size;
            sa.D = (float)e->opt.subset_stride;
            sa.model = e->opt.motion_model;
            sa.track_availability = correct ? 1 : 0;
            sa.projection = 1;
            sa.src_state = velocity ? e->prediction : e->hist_state;
            // The adaptive pyramid's finest-only frames project onto the finest
            // grid (fine_projection, full_pyramid_backend.py:838-868).
            sa.permissions = k == e->levels - 1 ? e->temporal_links : e->fine_links;
            sa.dst_seed = sl.seed;
            sa.dst_active = sl.active;
            sa.seeded_count = s->counts;
            HIP_CHECK(launch_transfer_seed(sa, st));
            seed = sl.seed;
            active = sl.active;
        }
        if (k == L - 1 && e->init.kind == LDIC_INIT_ZNCC) {
            // Translation seeds of the coarsest level from the exhaustive
            // ZNCC search; ROI-excluded subsets stay disabled.
            SearchArgs q = search_args(e, k, sl, e->init.search_radius);
            q.mask = lev.roi ? lev.selected : nullptr;
            q.seed = sl.seed;
            q.active = sl.active;
            HIP_CHECK(launch_zncc_search(q, st));
            seed = sl.seed;
            active = sl.active;
        }
        if (e->opt.interpolation == kBspline)
            HIP_CHECK(launch_prefilter(sl.target, lev.width, lev.height, e->fir, sl.coeff, st));
        if (k == 0) HIP_CHECK(hipEventRecord(s->ev_solve0[ring], st));
        SolveArgs a = solve_args(e, k, sl, seed, active, 0);
        if (k == 0) {
            // Divergence patience applies at the finest level only: a premature
            // coarse stop would spoil every finer seed (full_pyramid_backend.py:1214-1224).
            a.patience = e->init.divergence_patience;
            a.min_iters = e->init.divergence_min_iterations;
        }
        HIP_CHECK(launch_solve(e->opt.motion_model, e->opt.interpolation, e->opt.spatial_weighting, a, st));
        CHECK(launch_partial(e, k, sl, seed, active, 0, st, nullptr, a.patience, a.min_iters));
        if (k == 0) CHECK(reseed_passes(e, sl, st));
        if (k == 0) HIP_CHECK(hipEventRecord(s->ev_solve1[ring], st));
        CHECK(launch_median_level(e, k, sl, e->opt.outlier_policy == kMedian && k > 0, st));
    }
    if (temporal) {
        // update_temporal_state (full_pyramid.wgsl:519-565): the filtered output
        // is merged under the replace policy without prediction, the raw state
        // otherwise; with prediction the history first becomes the previous
        // measurement where a new one is accepted.
        SlotLevel& fine = s->lv[0];
        const int filtered = e->opt.outlier_policy == kReplace && !velocity ? 1 : 0;
        if (velocity)
            HIP_CHECK(launch_temporal_save_previous(e->lv[0].count(), e->opt.motion_model, fine.state,
                                                    level_output(e, fine), e->hist_state, e->hist_ages,
                                                    e->prev_state, e->prev_ages, st));
        HIP_CHECK(launch_temporal_merge(e->lv[0].count(), e->opt.motion_model, fine.state, level_output(e, fine),
                                        filtered, e->init.temporal_max_age, e->hist_state, e->hist_ages,
                                        s->counts + 1, st, adaptive ? s->counts + 3 : nullptr));
        if (correct) {
            // A frame that skipped the coarse level leaves no coarse measurement.
            const int c = e->levels - 1;
            SlotLevel& coarse = s->lv[c];
            HIP_CHECK(launch_temporal_merge(e->lv[c].count(), e->opt.motion_model, L > c ? coarse.state : nullptr,
                                            level_output(e, coarse), filtered, e->init.temporal_max_age,
                                            e->coarse_state, e->coarse_ages, nullptr, st));
        }
        HIP_CHECK(hipMemcpyAsync(s->pinned_counts, s->counts, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipEventRecord(e->temporal_done, st));
        e->temporal_pending = true;
    }
    return LDIC_OK;
}

ZnccArgs zncc_args(const ldic_engine* e, const SlotLevel& sl, const float* state, float* out) {
    const Level& L = e->lv[0];
    ZnccArgs z{};
    z.width = L.width;
    z.height = L.height;
    z.grid_w = L.gw;
    z.grid_h = L.gh;
    z.S = e->opt.subset_size;
    z.D = e->opt.subset_stride;
    z.interpolation = e->opt.interpolation;
    z.reference = L.reference;
    z.state = state;
    z.coeff = sl.coeff;
    z.target = sl.target;
    z.out = out;
    return z;
}

// Seeded finest-level registration (solve_seeded, icgn.wgsl:1001-1008) from
// the seeds in lv[0].seed / lv[0].active, with the optional ZNCC fallback:
// subsets that did not converge, or whose ZNCC in the last solver iteration is
// below init.fallback_zncc, are searched again around their seed and re-solved from
// the ZNCC seed while every other subset keeps its state (retry_seeded,
// icgn.wgsl:1010-1017).  include_zncc leaves the final ZNCC of the output in
// slot->zncc.
int enqueue_seeded(ldic_engine* e, Slot* s, bool include_zncc) {
    hipStream_t st = s->stream;
    const Level& L = e->lv[0];
    SlotLevel& sl = s->lv[0];
    const int model = e->opt.motion_model;
    if (e->opt.interpolation == kBspline) HIP_CHECK(launch_prefilter(sl.target, L.width, L.height, e->fir, sl.coeff, st));
    if (s->ring_count == kEventRing) CHECK(drain_ring(s));
    const int ring = (s->ring_first + s->ring_count) % kEventRing;
    s->ring_count += 1;
    HIP_CHECK(hipEventRecord(s->ev_solve0[ring], st));
    const bool fallback = e->init.fallback_zncc > -1.0f;
    SolveArgs first = solve_args(e, 0, sl, sl.seed, sl.active, 0);
    // The fallback criterion uses the ZNCC the solver evaluates in its last
    // iteration (at the warp it then updates by less than the tolerance).
    if (fallback) first.zncc_out = s->zncc;
    HIP_CHECK(launch_solve(model, e->opt.interpolation, e->opt.spatial_weighting, first, st));
    CHECK(launch_partial(e, 0, sl, sl.seed, sl.active, 0, st, first.zncc_out));
    if (fallback) {
        HIP_CHECK(launch_retry_mask(L.count(), model, sl.state, s->zncc, e->init.fallback_zncc,
                                    L.roi ? L.selected : nullptr, sl.mask, st));
        SearchArgs q = search_args(e, 0, sl, e->init.fallback_radius);
        q.prior = sl.seed;
        q.prior_stride = state_stride(model);
        q.mask = sl.mask;
        q.seed = sl.retry_seed;
        q.active = sl.retry_active;
        HIP_CHECK(launch_zncc_search(q, st));
        HIP_CHECK(launch_solve(model, e->opt.interpolation, e->opt.spatial_weighting,
                               solve_args(e, 0, sl, sl.retry_seed, sl.retry_active, 1), st));
        CHECK(launch_partial(e, 0, sl, sl.retry_seed, sl.retry_active, 1, st));
    }
    CHECK(reseed_passes(e, sl, st));
    HIP_CHECK(hipEventRecord(s->ev_solve1[ring], st));
    CHECK(launch_median_level(e, 0, sl, false, st));
    if (include_zncc) HIP_CHECK(launch_affine_zncc(zncc_args(e, sl, level_output(e, sl), s->zncc), st));
    return LDIC_OK;
}

int validate_initialization(const ldic_engine* e, const ldic_initialization* in) {
    if (!in) return set_error(LDIC_ERR_ARGUMENT, "initialization is NULL");
    const int S = e->opt.subset_size;
    if (in->kind != LDIC_INIT_ZERO && in->kind != LDIC_INIT_ZNCC && in->kind != LDIC_INIT_TEMPORAL)
        return set_error(LDIC_ERR_ARGUMENT, "initialization kind must be LDIC_INIT_ZERO, _ZNCC or _TEMPORAL");
    if (in->temporal_max_age > 0xFFFFFFFEu)
        return set_error(LDIC_ERR_ARGUMENT, "temporal_max_age must be at most 0xFFFFFFFE");
    if (in->temporal_prediction != 0 && in->temporal_prediction != 1)
        return set_error(LDIC_ERR_ARGUMENT, "temporal_prediction must be 0 (none) or 1 (velocity)");
    if (in->temporal_prediction && in->kind != LDIC_INIT_TEMPORAL)
        return set_error(LDIC_ERR_ARGUMENT, "temporal_prediction requires temporal initialization");
    if (in->temporal_fine_correction && (in->kind != LDIC_INIT_TEMPORAL || e->opt.pyramid_levels != 2))
        return set_error(LDIC_ERR_ARGUMENT, "temporal_fine_correction requires temporal initialization and two levels");
    if (in->divergence_patience < 0 || in->divergence_patience > 30)
        return set_error(LDIC_ERR_ARGUMENT, "divergence_patience must be 0 (off) or an integer from 1 to 30");
    if (in->adaptive_pyramid && (in->kind != LDIC_INIT_TEMPORAL || e->opt.pyramid_levels != 2))
        return set_error(LDIC_ERR_ARGUMENT, "adaptive pyramid requires temporal initialization and two levels");
    if (in->adaptive_pyramid && !(in->adaptive_acceptance > 0.0f && in->adaptive_acceptance <= 1.0f))
        return set_error(LDIC_ERR_ARGUMENT, "adaptive_acceptance must be finite and in (0, 1]");
    if (in->reseed_passes < 0 || in->reseed_passes > 8)
        return set_error(LDIC_ERR_ARGUMENT, "reseed_passes must be an integer from 0 to 8");
    if (in->reseed_passes && e->opt.motion_model == kQuadratic)
        return set_error(LDIC_ERR_ARGUMENT, "neighbour re-seeding supports translation and affine");
    if (in->divergence_patience && (in->divergence_min_iterations < 1 || in->divergence_min_iterations > 30))
        return set_error(LDIC_ERR_ARGUMENT, "divergence_min_iterations must be an integer from 1 to 30");
    const bool search = in->kind == LDIC_INIT_ZNCC, fallback = in->fallback_zncc > -1.0f;
    if (!std::isfinite(in->fallback_zncc) || in->fallback_zncc > 1.0f)
        return set_error(LDIC_ERR_ARGUMENT, "fallback_zncc must be finite and at most 1 (<= -1 disables it)");
    if ((search || fallback) && S > kMaxSearchSubset)
        return set_error(LDIC_ERR_ARGUMENT, "the ZNCC search supports subset sizes up to %d", kMaxSearchSubset);
    const int radii[2] = {search ? in->search_radius : 0, fallback ? in->fallback_radius : 0};
    for (int r : radii) {
        if (r < 0 || r > kMaxSearchRadius)
            return set_error(LDIC_ERR_ARGUMENT, "search radius %d is outside 0..%d", r, kMaxSearchRadius);
        if (zncc_search_lds_bytes(S, r) > kMaxSearchLds)
            return set_error(LDIC_ERR_ARGUMENT,
                             "a ZNCC search of radius %d with subset %d needs %zu bytes of LDS (limit %zu); "
                             "search on a coarser pyramid level", r, S, zncc_search_lds_bytes(S, r), kMaxSearchLds);
    }
    return LDIC_OK;
}

// Host staging copies (caller buffer <-> pinned staging) of whole frames and fields
// are split over a small pool of persistent threads: one thread's memcpy reaches a
// fraction of the host's memory bandwidth, and these copies sit on the host path
// between two frames in flight.
class HostCopier {
  public:
    static HostCopier& get() {
        static HostCopier pool;
        return pool;
    }
    // Rows of `row_bytes` from src (pitch src_pitch) to dst (pitch dst_pitch).
    void copy2d(void* dst, size_t dst_pitch, const void* src, size_t src_pitch, size_t row_bytes, size_t rows) {
        const size_t bytes = row_bytes * rows;
        const int parts = workers_.empty() || bytes < kMinParallelBytes ? 1 : (int)workers_.size() + 1;
        if (parts == 1) {
            run(dst, dst_pitch, src, src_pitch, row_bytes, 0, rows);
            return;
        }
        std::unique_lock<std::mutex> lock(job_mutex_);  // one job at a time
        {
            std::lock_guard<std::mutex> g(m_);
            job_ = Job{dst, dst_pitch, src, src_pitch, row_bytes, rows, parts};
            remaining_ = parts - 1;
            ++generation_;
        }
        cv_.notify_all();
        run_part(job_, 0);
        std::unique_lock<std::mutex> g(m_);
        done_cv_.wait(g, [&] { return remaining_ == 0; });
    }

  private:
    struct Job {
        void* dst; size_t dst_pitch; const void* src; size_t src_pitch; size_t row_bytes, rows; int parts;
    };
    static constexpr size_t kMinParallelBytes = 1u << 20;

    HostCopier() {
        unsigned hw = std::thread::hardware_concurrency();
        int n = hw >= 16 ? 7 : hw >= 8 ? 3 : hw >= 4 ? 1 : 0;
        for (int i = 0; i < n; ++i) workers_.emplace_back([this, i] { loop(i + 1); });
    }
    ~HostCopier() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : workers_) t.join();
    }
    static void run(void* dst, size_t dst_pitch, const void* src, size_t src_pitch, size_t row_bytes, size_t r0,
                    size_t r1) {
        auto* d = static_cast<uint8_t*>(dst);
        auto* q = static_cast<const uint8_t*>(src);
        if (dst_pitch == row_bytes && src_pitch == row_bytes) {
            std::memcpy(d + r0 * row_bytes, q + r0 * row_bytes, (r1 - r0) * row_bytes);
            return;
        }
        for (size_t r = r0; r < r1; ++r) std::memcpy(d + r * dst_pitch, q + r * src_pitch, row_bytes);
    }
    static void run_part(const Job& j, int part) {
        // Contiguous copies split by bytes, pitched ones by rows.
        if (j.dst_pitch == j.row_bytes && j.src_pitch == j.row_bytes) {
            const size_t total = j.row_bytes * j.rows;
            const size_t chunk = (total / j.parts + 4095) & ~(size_t)4095;
            const size_t b0 = std::min(total, chunk * part), b1 = std::min(total, b0 + chunk);
            if (b1 > b0)
                std::memcpy(static_cast<uint8_t*>(j.dst) + b0, static_cast<const uint8_t*>(j.src) + b0, b1 - b0);
            return;
        }
        const size_t per = (j.rows + j.parts - 1) / j.parts;
        const size_t r0 = std::min(j.rows, per * part), r1 = std::min(j.rows, r0 + per);
        run(j.dst, j.dst_pitch, j.src, j.src_pitch, j.row_bytes, r0, r1);
    }
    void loop(int part) {
        uint64_t seen = 0;
        for (;;) {
            Job job;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || generation_ != seen; });
                if (stop_) return;
                seen = generation_;
                job = job_;
            }
            if (part < job.parts) run_part(job, part);
            std::lock_guard<std::mutex> g(m_);
            if (part < job.parts && --remaining_ == 0) done_cv_.notify_one();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex job_mutex_, m_;
    std::condition_variable cv_, done_cv_;
    Job job_{};
    int remaining_ = 0;
    uint64_t generation_ = 0;
    bool stop_ = false;
};

int copy_in(ldic_engine* e, Slot* s, const uint8_t* target, int64_t row_stride) {
    const Level& L = e->lv[0];
    double t0 = now_ms();
    HostCopier::get().copy2d(s->pinned_in, (size_t)L.width, target, (size_t)row_stride, (size_t)L.width,
                             (size_t)L.height);
    HIP_CHECK(hipEventRecord(s->ev_begin, s->stream));
    HIP_CHECK(hipMemcpyAsync(s->lv[0].target, s->pinned_in, (size_t)L.width * L.height, hipMemcpyHostToDevice,
                             s->stream));
    s->upload_ms = now_ms() - t0;
    return LDIC_OK;
}

int check_ready(ldic_engine* e) {
    if (!e) return set_error(LDIC_ERR_ARGUMENT, "engine is NULL");
    if (!e->ready) return set_error(LDIC_ERR_STATE, "call ldic_set_reference before registering targets");
    return LDIC_OK;
}

int check_slot(ldic_engine* e, int slot) {
    CHECK(check_ready(e));
    if (slot < 0 || slot >= (int)e->slots.size())
        return set_error(LDIC_ERR_ARGUMENT, "slot %d is outside the %zu allocated slots", slot, e->slots.size());
    return LDIC_OK;
}

// A batch captured this engine's buffers and solver settings.
int check_unbatched(const ldic_engine* e) {
    if (e && e->batch_refs > 0)
        return set_error(LDIC_ERR_STATE, "the engine belongs to a batch; destroy the batch (ldic_batch_destroy) first");
    return LDIC_OK;
}

// The reference buffers are engine-wide: a slot may replace them only when no other
// slot has queued work that reads them.  (Separate references in flight need one
// engine per reference.)
int check_others_idle(const ldic_engine* e, int slot) {
    for (int j = 0; j < (int)e->slots.size(); ++j)
        if (j != slot && e->slots[j]->pending)
            return set_error(LDIC_ERR_STATE,
                             "slot %d has queued work that reads the current reference; wait for it "
                             "(ldic_retire or ldic_slot_wait) before replacing the reference on slot %d",
                             j, slot);
    return LDIC_OK;
}

// The adaptive pyramid's feedback from a retired frame (full_pyramid_backend.py:
// 1781-1791): the finest level alone next time if the merged finest subsets reach
// the acceptance fraction of the reference-valid ones and none expired.
void adaptive_feedback(ldic_engine* e, Slot* s) {
    if (!s->feedback_pending) return;
    s->feedback_pending = false;
    const double accepted = (double)s->pinned_counts[1];
    const bool good = e->adaptive_population > 0 &&
                      accepted / (double)e->adaptive_population >= (double)e->init.adaptive_acceptance &&
                      s->pinned_counts[3] == 0;
    e->adaptive_next = good ? 1 : 2;
}

int finish(ldic_engine* e, Slot* s, float* out_state, ldic_timing* t, double wall_start) {
    double t0 = now_ms();
    HIP_CHECK(hipEventSynchronize(s->ev_end));
    CHECK(drain_ring(s));
    s->pending = false;
    adaptive_feedback(e, s);
    if (out_state) {
        const size_t bytes = sizeof(float) * (size_t)output_floats(e);
        HostCopier::get().copy2d(out_state, bytes, s->pinned_out, bytes, bytes, 1);
    }
    if (t) {
        const bool summary = s->summary_ready && out_state;
        t->accepted = summary ? (int64_t)s->pinned_summary[0] : -1;
        t->replaced = summary ? (int64_t)s->pinned_summary[1] : -1;
        t->sum_translation_x = summary ? s->pinned_summary[2] : 0.0;
        t->sum_translation_y = summary ? s->pinned_summary[3] : 0.0;
        t->sum_iterations = summary ? s->pinned_summary[4] : 0.0;
        float gpu = 0.0f;
        HIP_CHECK(hipEventElapsedTime(&gpu, s->ev_begin, s->ev_end));
        t->upload_ms = s->upload_ms;
        t->gpu_ms = gpu;
        t->solve_ms = s->drained_solve_ms;
        t->frames = s->drained_frames;
        t->temporal_seeded = e->init.kind == LDIC_INIT_TEMPORAL ? (int32_t)s->pinned_counts[0] : 0;
        t->temporal_updates = e->init.kind == LDIC_INIT_TEMPORAL ? (int32_t)s->pinned_counts[1] : 0;
        t->velocity_predicted = e->init.kind == LDIC_INIT_TEMPORAL ? (int32_t)s->pinned_counts[2] : 0;
        t->levels_used = s->levels_used;
        t->temporal_expired = s->adaptive_frame ? (int32_t)s->pinned_counts[3] : 0;
        t->readback_ms = now_ms() - t0;
        t->wall_ms = now_ms() - wall_start;
    }
    s->drained_solve_ms = 0.0;
    s->drained_frames = 0;
    s->summary_ready = false;
    return LDIC_OK;
}

// Queue the readback of a finest field into the slot's pinned staging, with its
// summary (launch_field_summary) so the host result needs no pass over the field.
int enqueue_readback(ldic_engine* e, Slot* s, const float* out) {
    const int n = e->lv[0].count();
    HIP_CHECK(launch_field_summary(out, n, e->opt.motion_model, s->summary, s->stream));
    HIP_CHECK(hipMemcpyAsync(s->pinned_out, out, sizeof(float) * (size_t)output_floats(e), hipMemcpyDeviceToHost,
                             s->stream));
    HIP_CHECK(hipMemcpyAsync(s->pinned_summary, s->summary, 5 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    s->summary_ready = true;
    return LDIC_OK;
}

int submit_frame(ldic_engine* e, Slot* s, const uint8_t* target, int64_t row_stride) {
    if (s->pending) return set_error(LDIC_ERR_STATE, "slot still holds an unretired frame");
    CHECK(copy_in(e, s, target, row_stride));
    CHECK(enqueue_frame(e, s));
    CHECK(enqueue_readback(e, s, level_output(e, s->lv[0])));
    HIP_CHECK(hipEventRecord(s->ev_end, s->stream));
    s->pending = true;
    s->resident = false;
    s->submitted_at = now_ms();
    return LDIC_OK;
}

void gaussian_taps(double w[7]) {
    double sum = 0.0;
    for (int i = 0; i < 7; ++i) {
        double o = (double)(i - 3);
        w[i] = std::exp(-0.5 * o * o);
    }
    for (int i = 0; i < 7; ++i) sum += w[i];
    for (int i = 0; i < 7; ++i) w[i] /= sum;
}

// 31 float32 FIR taps: Gaussian(sigma=1) convolved with sqrt(3) z^|k|, |k|<=12
// (wgpu_backend.py:527-541).
void fir_taps(float out[31]) {
    double g[7];
    gaussian_taps(g);
    double w[31] = {0};
    const double z = std::sqrt(3.0) - 2.0;
    for (int gi = 0; gi < 7; ++gi)
        for (int k = -12; k <= 12; ++k) w[(gi - 3) + k + 15] += g[gi] * std::sqrt(3.0) * std::pow(z, std::abs(k));
    for (int i = 0; i < 31; ++i) out[i] = (float)w[i];
}

int validate_options(const ldic_options* o) {
    if (!o) return set_error(LDIC_ERR_ARGUMENT, "options is NULL");
    if (o->motion_model < 0 || o->motion_model > 2) return set_error(LDIC_ERR_ARGUMENT, "motion_model must be 0, 1 or 2");
    if (o->interpolation < 0 || o->interpolation > 2) return set_error(LDIC_ERR_ARGUMENT, "interpolation must be 0, 1 or 2");
    if (o->subset_size < 4 || o->subset_size > 46340) return set_error(LDIC_ERR_ARGUMENT, "subset_size must be 4..46340");
    if (o->subset_stride < 1 || o->subset_stride > o->subset_size)
        return set_error(LDIC_ERR_ARGUMENT, "subset_stride must be between 1 and subset_size");
    if (o->spatial_weighting != 0 && o->spatial_weighting != 1)
        return set_error(LDIC_ERR_ARGUMENT, "spatial_weighting must be 0 or 1");
    if (!(o->gaussian_sigma > 0.0f) || !std::isfinite(o->gaussian_sigma))
        return set_error(LDIC_ERR_ARGUMENT, "gaussian_sigma must be finite and positive");
    if (o->outlier_policy < 0 || o->outlier_policy > 3) return set_error(LDIC_ERR_ARGUMENT, "outlier_policy must be 0..3");
    if (o->median_filter_size != 3 && o->median_filter_size != 5 && o->median_filter_size != 7 &&
        o->median_filter_size != 9)
        return set_error(LDIC_ERR_ARGUMENT, "median_filter_size must be 3, 5, 7 or 9");
    if (o->median_filter_scope != 0 && o->median_filter_scope != 1)
        return set_error(LDIC_ERR_ARGUMENT, "median_filter_scope must be 0 or 1");
    if (o->pyramid_levels < 1 || o->pyramid_levels > LDIC_MAX_LEVELS)
        return set_error(LDIC_ERR_ARGUMENT, "pyramid_levels must be 1..4");
    return LDIC_OK;
}

int validate_config(const ldic_level_config* c) {
    if (c->iterations < 1 || c->iterations > 30) return set_error(LDIC_ERR_ARGUMENT, "iterations must be 1..30");
    if (!(c->tolerance > 0.0f) || !std::isfinite(c->tolerance))
        return set_error(LDIC_ERR_ARGUMENT, "tolerance must be finite and positive");
    if (!(c->zncc_min >= 0.0f && c->zncc_min <= 1.0f)) return set_error(LDIC_ERR_ARGUMENT, "zncc_min must be in [0, 1]");
    return LDIC_OK;
}

int prepare_level(ldic_engine* e, int k, hipStream_t st) {
    Level& L = e->lv[k];
    double *G = e->prep_g, *A = e->prep_a, *B = e->prep_b, *GX = e->prep_gx, *GY = e->prep_gy;
    double* weights = e->weights;
    const int w = L.width, h = L.height;
    const bool bspline = e->opt.interpolation == kBspline;
    if (bspline) {
        HIP_CHECK(launch_gaussian(L.image, w, h, e->gauss7, B, G, st));
        HIP_CHECK(hipMemcpyAsync(A, G, sizeof(double) * (size_t)w * h, hipMemcpyDeviceToDevice, st));
        HIP_CHECK(launch_bspline(A, B, w, h, st));
        HIP_CHECK(launch_gradients(A, G, w, h, 1, GX, GY, L.reference, st));
    } else {
        HIP_CHECK(launch_to_double(L.image, w, h, G, st));
        HIP_CHECK(launch_gradients(G, G, w, h, 0, GX, GY, L.reference, st));
    }
    PrepArgs p{};
    p.width = w;
    p.height = h;
    p.grid_w = L.gw;
    p.grid_h = L.gh;
    p.S = e->opt.subset_size;
    p.D = e->opt.subset_stride;
    p.model = e->opt.motion_model;
    p.weighted = e->opt.spatial_weighting;
    p.direct = bspline ? 0 : 1;
    p.intensity = G;
    p.grad_x = GX;
    p.grad_y = GY;
    p.weights = weights;
    p.weight_sum = e->weight_sum;
    p.selected = L.roi ? L.selected : nullptr;
    p.reference = L.reference;
    p.metadata = L.metadata;
    p.inverse = L.inverse;
    p.valid = L.valid;
    HIP_CHECK(launch_subset_prep(p, st));
    return LDIC_OK;
}

// Reference pyramid and per-level preparation, queued on `st`.  The finest
// image is already in lv[0].image.
int enqueue_reference(ldic_engine* e, hipStream_t st) {
    for (int k = 1; k < e->levels; ++k)
        HIP_CHECK(launch_reduce(e->lv[k - 1].image, e->lv[k - 1].width, e->lv[k - 1].height, e->lv[k].image, st));
    for (int k = 0; k < e->levels; ++k) CHECK(prepare_level(e, k, st));
    return LDIC_OK;
}

int set_reference_impl(ldic_engine* e, const uint8_t* image, int width, int height, int64_t row_stride,
                       const ldic_level_config* configs, const ldic_level_roi* rois, uint8_t* out_valid) {
    if (!image || !configs) return set_error(LDIC_ERR_ARGUMENT, "image and configs are required");
    if (width < 16 || height < 16) return set_error(LDIC_ERR_ARGUMENT, "image dimensions must be at least 16");
    // The solver addresses the 16-byte reference plane and the target
    // coefficients by 32-bit byte offsets.
    if ((int64_t)width * height >= (int64_t)1 << 28)
        return set_error(LDIC_ERR_ARGUMENT, "images must have fewer than 2^28 pixels (%dx%d)", width, height);
    if (row_stride < width) return set_error(LDIC_ERR_ARGUMENT, "row_stride is smaller than width");
    const int S = e->opt.subset_size, D = e->opt.subset_stride;
    // Validate every level before touching the previous reference.
    int w = width, h = height;
    int dims[LDIC_MAX_LEVELS][4];
    for (int k = 0; k < e->opt.pyramid_levels; ++k) {
        CHECK(validate_config(&configs[k]));
        if (k > 0) {
            w /= 2;
            h /= 2;
        }
        if (w < S || h < S)
            return set_error(LDIC_ERR_ARGUMENT, "level %d (%dx%d) is smaller than the subset size %d", k, w, h, S);
        dims[k][0] = w;
        dims[k][1] = h;
        dims[k][2] = 1 + (w - S) / D;
        dims[k][3] = 1 + (h - S) / D;
    }
    HIP_CHECK(hipSetDevice(e->device));
    e->ready = false;
    free_slots(e);
    e->ref_arena.release();
    e->levels = e->opt.pyramid_levels;
    const int M = metadata_stride(e->opt.motion_model), P = param_count(e->opt.motion_model);
    for (int k = 0; k < e->levels; ++k) {
        Level& L = e->lv[k];
        L = Level();
        L.width = dims[k][0];
        L.height = dims[k][1];
        L.gw = dims[k][2];
        L.gh = dims[k][3];
        L.config = configs[k];
        size_t px = (size_t)L.width * L.height, n = (size_t)L.count();
        CHECK(e->ref_arena.alloc(&L.image, px));
        CHECK(e->ref_arena.alloc(&L.reference, px));
        CHECK(e->ref_arena.alloc(&L.metadata, n * M));
        CHECK(e->ref_arena.alloc(&L.inverse, n * P * P));
        CHECK(e->ref_arena.alloc(&L.valid, n));
        const ldic_level_roi* r = rois ? &rois[k] : nullptr;
        if (r && r->selected) {
            L.roi = true;
            CHECK(e->ref_arena.alloc(&L.selected, n));
            HIP_CHECK(hipMemcpy(L.selected, r->selected, n, hipMemcpyHostToDevice));
            CHECK(e->ref_arena.alloc(&L.roi_words, n * 4));
            if (r->neighbour_words) {
                HIP_CHECK(hipMemcpy(L.roi_words, r->neighbour_words, n * 4 * sizeof(float), hipMemcpyHostToDevice));
            } else {
                HIP_CHECK(hipMemset(L.roi_words, 0, n * 4 * sizeof(float)));
            }
        }
        if (r && r->seed_links) {
            CHECK(e->ref_arena.alloc(&L.seed_links, n));
            HIP_CHECK(hipMemcpy(L.seed_links, r->seed_links, n * sizeof(uint32_t), hipMemcpyHostToDevice));
        }
    }
    // Integration weights of one subset (Gaussian mode).
    double* d_weights = nullptr;
    e->weights = nullptr;
    if (e->opt.spatial_weighting == 1) {
        std::vector<double> wts((size_t)S * S);
        double half = 0.5 * (S - 1), sigma = e->opt.gaussian_sigma;
        for (int r = 0; r < S; ++r)
            for (int c = 0; c < S; ++c) {
                double lx = c - half, ly = r - half;
                wts[(size_t)r * S + c] = std::exp(-0.5 * (lx * lx + ly * ly) / (sigma * sigma));
            }
        double sum = 0.0;
        for (double v : wts) sum += v;
        e->weight_sum = sum;
        CHECK(e->ref_arena.alloc(&d_weights, wts.size()));
        HIP_CHECK(hipMemcpy(d_weights, wts.data(), wts.size() * sizeof(double), hipMemcpyHostToDevice));
        e->weights = d_weights;
    } else {
        e->weight_sum = (double)S * S;
    }
    // Upload the reference, build its pyramid and prepare every level.
    hipStream_t st = e->prep_stream;
    const Level& L0 = e->lv[0];
    size_t px0 = (size_t)L0.width * L0.height;
    CHECK(e->ref_arena.alloc(&e->prep_g, px0));
    CHECK(e->ref_arena.alloc(&e->prep_a, px0));
    CHECK(e->ref_arena.alloc(&e->prep_b, px0));
    CHECK(e->ref_arena.alloc(&e->prep_gx, px0));
    CHECK(e->ref_arena.alloc(&e->prep_gy, px0));
    CHECK(e->ref_arena.alloc(&e->hist_state, (size_t)L0.count() * state_stride(e->opt.motion_model)));
    CHECK(e->ref_arena.alloc(&e->hist_ages, (size_t)L0.count()));
    CHECK(e->ref_arena.alloc(&e->prev_state, (size_t)L0.count() * state_stride(e->opt.motion_model)));
    CHECK(e->ref_arena.alloc(&e->prev_ages, (size_t)L0.count()));
    CHECK(e->ref_arena.alloc(&e->prediction, (size_t)L0.count() * state_stride(e->opt.motion_model)));
    const int coarse_count = e->lv[e->levels - 1].count();
    CHECK(e->ref_arena.alloc(&e->coarse_state, (size_t)coarse_count * state_stride(e->opt.motion_model)));
    CHECK(e->ref_arena.alloc(&e->coarse_ages, (size_t)coarse_count));
    e->temporal_links = nullptr;
    HIP_CHECK(reset_histories(e, st));
    HIP_CHECK(hipMemcpy2DAsync(L0.image, L0.width, image, (size_t)row_stride, L0.width, L0.height,
                               hipMemcpyHostToDevice, st));
    int rc = enqueue_reference(e, st);
    hipError_t sync = hipStreamSynchronize(st);
    if (rc != LDIC_OK) return rc;
    if (sync != hipSuccess) return set_error(LDIC_ERR_DEVICE, "reference preparation failed: %s", hipGetErrorString(sync));
    std::vector<uint8_t>& valid = e->valid_host;
    valid.assign((size_t)L0.count(), 0);
    HIP_CHECK(hipMemcpy(valid.data(), L0.valid, valid.size(), hipMemcpyDeviceToHost));
    if (out_valid) std::memcpy(out_valid, valid.data(), valid.size());
    e->adaptive_population = 0;
    for (uint8_t v : valid) e->adaptive_population += v != 0;
    e->fine_links = nullptr;
    if (L0.roi) {
        CHECK(e->ref_arena.alloc(&e->fine_links, (size_t)L0.count()));
        HIP_CHECK(launch_links_from_selected(L0.count(), L0.selected, e->fine_links, st));
        HIP_CHECK(hipStreamSynchronize(st));
    }
    CHECK(ensure_slots(e, 1));
    e->ready = true;
    return LDIC_OK;
}

}  // namespace

extern "C" {

int ldic_abi_version(void) { return LDIC_ABI_VERSION; }

const char* ldic_last_error(void) { return g_error.c_str(); }

int ldic_create(const ldic_options* options, ldic_engine** out_engine) {
    if (!out_engine) return set_error(LDIC_ERR_ARGUMENT, "out_engine is NULL");
    *out_engine = nullptr;
    CHECK(validate_options(options));
    int count = 0;
    HIP_CHECK(hipGetDeviceCount(&count));
    if (options->device_ordinal < 0 || options->device_ordinal >= count)
        return set_error(LDIC_ERR_ARGUMENT, "device_ordinal %d is not one of the %d HIP devices", options->device_ordinal,
                         count);
    ldic_engine* e = new ldic_engine();
    e->opt = *options;
    e->device = options->device_ordinal;
    int rc = LDIC_OK;
    auto fail = [&](int code) {
        ldic_destroy(e);
        return code;
    };
    hipError_t he = hipSetDevice(e->device);
    if (he == hipSuccess) he = hipGetDeviceProperties(&e->prop, e->device);
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&e->prep_stream, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&e->temporal_done, hipEventDisableTiming);
    if (he != hipSuccess) return fail(set_error(LDIC_ERR_DEVICE, "device set-up failed: %s", hipGetErrorString(he)));
    float taps[31];
    fir_taps(taps);
    if ((rc = e->const_arena.alloc(&e->fir, 31)) != LDIC_OK) return fail(rc);
    he = hipMemcpy(e->fir, taps, sizeof taps, hipMemcpyHostToDevice);
    if (he != hipSuccess) return fail(set_error(LDIC_ERR_DEVICE, "FIR upload failed: %s", hipGetErrorString(he)));
    gaussian_taps(e->gauss7);
    e->solve_lds = solver_lds_bytes(options->subset_size, options->motion_model, &e->tile_cap);
    *out_engine = e;
    return LDIC_OK;
}

void ldic_destroy(ldic_engine* e) {
    if (!e) return;
    // A batch still holds the engine's device addresses: destroying it would leave the
    // batch launching on freed memory.  The engine stays; the error says why.
    if (check_unbatched(e) != LDIC_OK) return;
    (void)hipSetDevice(e->device);
    free_slots(e);
    e->ref_arena.release();
    e->const_arena.release();
    if (e->prep_stream) (void)hipStreamDestroy(e->prep_stream);
    if (e->temporal_done) (void)hipEventDestroy(e->temporal_done);
    delete e;
}

int ldic_device_name(ldic_engine* e, char* buffer, int32_t capacity) {
    if (!e || !buffer || capacity <= 0) return set_error(LDIC_ERR_ARGUMENT, "invalid buffer");
    std::snprintf(buffer, (size_t)capacity, "%s (%s, %d CUs)", e->prop.name, e->prop.gcnArchName,
                  e->prop.multiProcessorCount);
    return LDIC_OK;
}

int ldic_set_reference(ldic_engine* e, const uint8_t* image, int32_t width, int32_t height, int64_t row_stride,
                       const ldic_level_config* configs, const ldic_level_roi* rois, uint8_t* out_reference_valid) {
    if (!e) return set_error(LDIC_ERR_ARGUMENT, "engine is NULL");
    CHECK(check_unbatched(e));
    return set_reference_impl(e, image, width, height, row_stride, configs, rois, out_reference_valid);
}

int ldic_level_shape(ldic_engine* e, int32_t level, int32_t* width, int32_t* height, int32_t* gw, int32_t* gh) {
    CHECK(check_ready(e));
    if (level < 0 || level >= e->levels) return set_error(LDIC_ERR_ARGUMENT, "level out of range");
    const Level& L = e->lv[level];
    if (width) *width = L.width;
    if (height) *height = L.height;
    if (gw) *gw = L.gw;
    if (gh) *gh = L.gh;
    return LDIC_OK;
}

int ldic_set_level_config(ldic_engine* e, int32_t level, const ldic_level_config* c) {
    CHECK(check_unbatched(e));
    CHECK(check_ready(e));
    if (level < 0 || level >= e->levels || !c) return set_error(LDIC_ERR_ARGUMENT, "level out of range");
    CHECK(validate_config(c));
    e->lv[level].config = *c;
    return LDIC_OK;
}

int ldic_set_slots(ldic_engine* e, int32_t count) {
    CHECK(check_unbatched(e));
    CHECK(check_ready(e));
    if (count < 1 || count > 64) return set_error(LDIC_ERR_ARGUMENT, "slot count must be 1..64");
    return ensure_slots(e, count);
}

int ldic_register(ldic_engine* e, const uint8_t* target, int64_t row_stride, float* out_state, ldic_timing* t) {
    CHECK(check_unbatched(e));
    double wall = now_ms();
    CHECK(check_ready(e));
    if (!target || !out_state) return set_error(LDIC_ERR_ARGUMENT, "target and out_state are required");
    if (row_stride < e->lv[0].width) return set_error(LDIC_ERR_ARGUMENT, "row_stride is smaller than width");
    HIP_CHECK(hipSetDevice(e->device));
    Slot* s = e->slots[0];
    CHECK(submit_frame(e, s, target, row_stride));
    return finish(e, s, out_state, t, wall);
}

int ldic_register_pair(ldic_engine* e, const uint8_t* reference, int64_t reference_stride, const uint8_t* target,
                       int64_t target_stride, float* out_state, uint8_t* out_reference_valid, ldic_timing* t) {
    CHECK(check_unbatched(e));
    double wall = now_ms();
    CHECK(check_ready(e));
    if (!reference || !target || !out_state) return set_error(LDIC_ERR_ARGUMENT, "reference, target and out_state are required");
    if (has_partial(e))
        return set_error(LDIC_ERR_STATE, "consecutive-pair mode does not support partial-support ROI subsets");
    if (e->init.adaptive_pyramid)
        return set_error(LDIC_ERR_STATE, "consecutive-pair mode does not support the adaptive pyramid");
    const Level& L0 = e->lv[0];
    if (reference_stride < L0.width || target_stride < L0.width)
        return set_error(LDIC_ERR_ARGUMENT, "row strides are smaller than the width");
    HIP_CHECK(hipSetDevice(e->device));
    Slot* s = e->slots[0];
    if (s->pending) return set_error(LDIC_ERR_STATE, "slot 0 holds an unretired frame");
    CHECK(check_others_idle(e, 0));
    // A new reference starts a new temporal history.
    if (e->init.kind == LDIC_INIT_TEMPORAL) HIP_CHECK(reset_histories(e, s->stream));
    for (int y = 0; y < L0.height; ++y)
        std::memcpy(s->pinned_ref + (size_t)y * L0.width, reference + (size_t)y * reference_stride, (size_t)L0.width);
    HIP_CHECK(hipMemcpyAsync(e->lv[0].image, s->pinned_ref, (size_t)L0.width * L0.height, hipMemcpyHostToDevice,
                             s->stream));
    CHECK(enqueue_reference(e, s->stream));
    CHECK(submit_frame(e, s, target, target_stride));
    if (out_reference_valid)
        HIP_CHECK(hipMemcpyAsync(s->pinned_valid, L0.valid, (size_t)L0.count(), hipMemcpyDeviceToHost, s->stream));
    HIP_CHECK(hipEventRecord(s->ev_end, s->stream));
    CHECK(finish(e, s, out_state, t, wall));
    if (out_reference_valid) std::memcpy(out_reference_valid, s->pinned_valid, (size_t)L0.count());
    return LDIC_OK;
}

int ldic_register_pair_resident(ldic_engine* e, int32_t slot, const uint8_t* device_reference,
                                const uint8_t* device_target, int64_t row_stride) {
    CHECK(check_unbatched(e));
    CHECK(check_slot(e, slot));
    if (!device_reference || !device_target) return set_error(LDIC_ERR_ARGUMENT, "device pointers are required");
    if (has_partial(e))
        return set_error(LDIC_ERR_STATE, "consecutive-pair mode does not support partial-support ROI subsets");
    if (e->init.adaptive_pyramid)
        return set_error(LDIC_ERR_STATE, "consecutive-pair mode does not support the adaptive pyramid");
    Slot* s = e->slots[slot];
    if (s->pending && !s->resident) return set_error(LDIC_ERR_STATE, "slot %d holds an unretired frame", slot);
    CHECK(check_others_idle(e, slot));
    HIP_CHECK(hipSetDevice(e->device));
    const Level& L0 = e->lv[0];
    if (!s->pending) HIP_CHECK(hipEventRecord(s->ev_begin, s->stream));
    if (e->init.kind == LDIC_INIT_TEMPORAL) HIP_CHECK(reset_histories(e, s->stream));
    HIP_CHECK(hipMemcpy2DAsync(L0.image, L0.width, device_reference, (size_t)row_stride, L0.width, L0.height,
                               hipMemcpyDeviceToDevice, s->stream));
    CHECK(enqueue_reference(e, s->stream));
    HIP_CHECK(hipMemcpy2DAsync(s->lv[0].target, L0.width, device_target, (size_t)row_stride, L0.width, L0.height,
                               hipMemcpyDeviceToDevice, s->stream));
    CHECK(enqueue_frame(e, s));
    HIP_CHECK(hipEventRecord(s->ev_end, s->stream));
    s->upload_ms = 0.0;
    s->pending = true;
    s->resident = true;
    s->submitted_at = now_ms();
    return LDIC_OK;
}

int ldic_submit(ldic_engine* e, int32_t slot, const uint8_t* target, int64_t row_stride) {
    CHECK(check_unbatched(e));
    CHECK(check_slot(e, slot));
    if (!target) return set_error(LDIC_ERR_ARGUMENT, "target is NULL");
    if (row_stride < e->lv[0].width) return set_error(LDIC_ERR_ARGUMENT, "row_stride is smaller than width");
    HIP_CHECK(hipSetDevice(e->device));
    return submit_frame(e, e->slots[slot], target, row_stride);
}

int ldic_retire(ldic_engine* e, int32_t slot, float* out_state, ldic_timing* t) {
    CHECK(check_slot(e, slot));
    Slot* s = e->slots[slot];
    if (!s->pending || s->resident) return set_error(LDIC_ERR_STATE, "slot %d holds no submitted frame", slot);
    HIP_CHECK(hipSetDevice(e->device));
    return finish(e, s, out_state, t, s->submitted_at - s->upload_ms);
}

int ldic_register_seeded(ldic_engine* e, const uint8_t* target, int64_t row_stride, const float* seed_state,
                         const uint8_t* active, int32_t include_zncc, float* out_state, float* out_zncc,
                         ldic_timing* t) {
    CHECK(check_unbatched(e));
    double wall = now_ms();
    CHECK(check_ready(e));
    if (!target || !seed_state || !active || !out_state)
        return set_error(LDIC_ERR_ARGUMENT, "target, seed_state, active and out_state are required");
    if (include_zncc && (e->opt.motion_model != kAffine || !out_zncc))
        return set_error(LDIC_ERR_ARGUMENT, "include_zncc requires the affine model and an output buffer");
    if (include_zncc && (e->opt.spatial_weighting != 0 || has_partial(e)))
        return set_error(LDIC_ERR_ARGUMENT, "include_zncc requires uniform full windows");
    HIP_CHECK(hipSetDevice(e->device));
    Slot* s = e->slots[0];
    if (s->pending) return set_error(LDIC_ERR_STATE, "slot 0 holds an unretired frame");
    const Level& L = e->lv[0];
    const int n = L.count(), stride = state_stride(e->opt.motion_model);
    SlotLevel& sl = s->lv[0];
    CHECK(copy_in(e, s, target, row_stride));
    HIP_CHECK(hipMemcpyAsync(sl.seed, seed_state, sizeof(float) * (size_t)n * stride, hipMemcpyHostToDevice, s->stream));
    std::vector<uint32_t> act((size_t)n);
    for (int i = 0; i < n; ++i) act[i] = active[i] ? 1u : 0u;
    HIP_CHECK(hipMemcpyAsync(sl.active, act.data(), sizeof(uint32_t) * (size_t)n, hipMemcpyHostToDevice, s->stream));
    CHECK(enqueue_seeded(e, s, include_zncc != 0));
    const float* out = level_output(e, sl);
    if (include_zncc)
        HIP_CHECK(hipMemcpyAsync(s->pinned_zncc, s->zncc, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, s->stream));
    CHECK(enqueue_readback(e, s, out));
    HIP_CHECK(hipEventRecord(s->ev_end, s->stream));
    s->pending = true;
    s->resident = false;
    CHECK(finish(e, s, out_state, t, wall));
    if (include_zncc) {
        for (int i = 0; i < n; ++i) out_zncc[i] = s->pinned_zncc[i] < -1.0f ? NAN : s->pinned_zncc[i];
    }
    return LDIC_OK;
}

int ldic_register_resident(ldic_engine* e, int32_t slot, const uint8_t* device_target, int64_t row_stride) {
    CHECK(check_unbatched(e));
    CHECK(check_slot(e, slot));
    if (!device_target) return set_error(LDIC_ERR_ARGUMENT, "device_target is NULL");
    Slot* s = e->slots[slot];
    if (s->pending && !s->resident) return set_error(LDIC_ERR_STATE, "slot %d holds an unretired frame", slot);
    HIP_CHECK(hipSetDevice(e->device));
    const Level& L = e->lv[0];
    s->upload_ms = 0.0;
    if (!s->pending) HIP_CHECK(hipEventRecord(s->ev_begin, s->stream));
    HIP_CHECK(hipMemcpy2DAsync(s->lv[0].target, L.width, device_target, (size_t)row_stride, L.width, L.height,
                               hipMemcpyDeviceToDevice, s->stream));
    CHECK(enqueue_frame(e, s));
    HIP_CHECK(hipEventRecord(s->ev_end, s->stream));
    s->pending = true;
    s->resident = true;
    s->submitted_at = now_ms();
    return LDIC_OK;
}

int ldic_register_seeded_resident(ldic_engine* e, int32_t slot, const uint8_t* device_target, int64_t row_stride,
                                  const float* device_seed_state, const uint32_t* device_active) {
    CHECK(check_unbatched(e));
    CHECK(check_slot(e, slot));
    if (!device_target || !device_seed_state || !device_active)
        return set_error(LDIC_ERR_ARGUMENT, "device target, seed_state and active are required");
    Slot* s = e->slots[slot];
    if (s->pending && !s->resident) return set_error(LDIC_ERR_STATE, "slot %d holds an unretired frame", slot);
    HIP_CHECK(hipSetDevice(e->device));
    const Level& L = e->lv[0];
    SlotLevel& sl = s->lv[0];
    const size_t n = (size_t)L.count();
    s->upload_ms = 0.0;
    if (!s->pending) HIP_CHECK(hipEventRecord(s->ev_begin, s->stream));
    HIP_CHECK(hipMemcpy2DAsync(sl.target, L.width, device_target, (size_t)row_stride, L.width, L.height,
                               hipMemcpyDeviceToDevice, s->stream));
    HIP_CHECK(hipMemcpyAsync(sl.seed, device_seed_state, sizeof(float) * n * state_stride(e->opt.motion_model),
                             hipMemcpyDeviceToDevice, s->stream));
    HIP_CHECK(hipMemcpyAsync(sl.active, device_active, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s->stream));
    CHECK(enqueue_seeded(e, s, false));
    HIP_CHECK(hipEventRecord(s->ev_end, s->stream));
    s->pending = true;
    s->resident = true;
    s->submitted_at = now_ms();
    return LDIC_OK;
}

int ldic_prepare_reference_resident(ldic_engine* e, int32_t slot, const uint8_t* device_reference,
                                    int64_t row_stride) {
    CHECK(check_unbatched(e));
    CHECK(check_slot(e, slot));
    if (!device_reference) return set_error(LDIC_ERR_ARGUMENT, "device_reference is NULL");
    if (has_partial(e))
        return set_error(LDIC_ERR_STATE, "a new reference would discard the partial-support statistics");
    Slot* s = e->slots[slot];
    if (s->pending && !s->resident) return set_error(LDIC_ERR_STATE, "slot %d holds an unretired frame", slot);
    CHECK(check_others_idle(e, slot));
    if (row_stride < e->lv[0].width) return set_error(LDIC_ERR_ARGUMENT, "row_stride is smaller than width");
    HIP_CHECK(hipSetDevice(e->device));
    const Level& L0 = e->lv[0];
    if (!s->pending) HIP_CHECK(hipEventRecord(s->ev_begin, s->stream));
    HIP_CHECK(hipMemcpy2DAsync(L0.image, L0.width, device_reference, (size_t)row_stride, L0.width, L0.height,
                               hipMemcpyDeviceToDevice, s->stream));
    CHECK(enqueue_reference(e, s->stream));
    HIP_CHECK(hipEventRecord(s->ev_end, s->stream));
    s->pending = true;
    s->resident = true;
    s->submitted_at = now_ms();
    return LDIC_OK;
}

int ldic_set_initialization(ldic_engine* e, const ldic_initialization* init) {
    CHECK(check_unbatched(e));
    if (!e) return set_error(LDIC_ERR_ARGUMENT, "engine is NULL");
    CHECK(validate_initialization(e, init));
    for (Slot* s : e->slots)
        if (s->pending) return set_error(LDIC_ERR_STATE, "a slot holds an unretired frame");
    const bool restart = init->kind == LDIC_INIT_TEMPORAL && e->init.kind != LDIC_INIT_TEMPORAL;
    e->init = *init;
    if (restart && e->ready) CHECK(ldic_reset_temporal(e));
    return LDIC_OK;
}

int ldic_prepare_partial(ldic_engine* e, int32_t level, int32_t count, const uint32_t* indices,
                         const uint32_t* support_words, const float* boxes, float* out_diagnostics) {
    CHECK(check_unbatched(e));
    CHECK(check_ready(e));
    if (level < 0 || level >= e->levels) return set_error(LDIC_ERR_ARGUMENT, "level out of range");
    if (count < 0 || (count > 0 && (!indices || !support_words || !boxes)))
        return set_error(LDIC_ERR_ARGUMENT, "partial tables are required");
    if (count > 0 && e->opt.motion_model == kQuadratic)
        return set_error(LDIC_ERR_ARGUMENT, "partial supports exist for translation and affine only");
    for (Slot* s : e->slots)
        if (s->pending) return set_error(LDIC_ERR_STATE, "a slot holds an unretired frame");
    Level& L = e->lv[level];
    const int S = e->opt.subset_size, W = (S * S + 31) / 32;
    for (int i = 0; i < count; ++i)
        if (indices[i] >= (uint32_t)L.count()) return set_error(LDIC_ERR_ARGUMENT, "partial index %u out of range", indices[i]);
    HIP_CHECK(hipSetDevice(e->device));
    L.partial_count = 0;
    if (count == 0) return LDIC_OK;
    float* diag = nullptr;
    CHECK(e->ref_arena.alloc(&L.partial_list, (size_t)count));
    CHECK(e->ref_arena.alloc(&L.support, (size_t)count * W));
    CHECK(e->ref_arena.alloc(&L.support_boxes, (size_t)count));
    CHECK(e->ref_arena.alloc(&diag, (size_t)count * 4));
    HIP_CHECK(hipMemcpy(L.partial_list, indices, sizeof(uint32_t) * count, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(L.support, support_words, sizeof(uint32_t) * (size_t)count * W, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemcpy(L.support_boxes, boxes, sizeof(float4) * count, hipMemcpyHostToDevice));
    PartialStatsArgs pa{};
    pa.width = L.width;
    pa.grid_w = L.gw;
    pa.S = S;
    pa.D = e->opt.subset_stride;
    pa.model = e->opt.motion_model;
    pa.count = count;
    pa.reference = L.reference;
    pa.weights = e->opt.spatial_weighting == 1 ? e->weights : nullptr;
    pa.indices = L.partial_list;
    pa.support = L.support;
    pa.metadata = L.metadata;
    pa.inverse = L.inverse;
    pa.diagnostics = diag;
    HIP_CHECK(launch_partial_stats(pa, e->prep_stream));
    std::vector<float> d((size_t)count * 4);
    HIP_CHECK(hipMemcpyAsync(d.data(), diag, d.size() * sizeof(float), hipMemcpyDeviceToHost, e->prep_stream));
    HIP_CHECK(hipStreamSynchronize(e->prep_stream));
    if (out_diagnostics) std::memcpy(out_diagnostics, d.data(), d.size() * sizeof(float));
    L.partial_count = count;
    if (level == 0 && e->valid_host.size() == (size_t)L.count())
        for (int i = 0; i < count; ++i) {
            // A partial subset is reference-valid when its support is admitted.
            const uint8_t admitted = d[(size_t)i * 4] != 0.0f ? 1 : 0;
            e->adaptive_population += (int64_t)admitted - (int64_t)(e->valid_host[indices[i]] != 0);
            e->valid_host[indices[i]] = admitted;
        }
    return LDIC_OK;
}

int ldic_reset_temporal(ldic_engine* e) {
    CHECK(check_unbatched(e));
    CHECK(check_ready(e));
    for (Slot* s : e->slots)
        if (s->pending) return set_error(LDIC_ERR_STATE, "a slot holds an unretired frame");
    HIP_CHECK(hipSetDevice(e->device));
    HIP_CHECK(reset_histories(e, e->prep_stream));
    HIP_CHECK(hipStreamSynchronize(e->prep_stream));
    return LDIC_OK;
}

int ldic_set_temporal_links(ldic_engine* e, const uint32_t* coarse_links) {
    CHECK(check_unbatched(e));
    CHECK(check_ready(e));
    HIP_CHECK(hipSetDevice(e->device));
    if (!coarse_links) {
        e->temporal_links = nullptr;
        return LDIC_OK;
    }
    const size_t n = (size_t)e->lv[e->levels - 1].count();
    uint32_t* d = nullptr;
    CHECK(e->ref_arena.alloc(&d, n));
    HIP_CHECK(hipMemcpy(d, coarse_links, n * sizeof(uint32_t), hipMemcpyHostToDevice));
    e->temporal_links = d;
    return LDIC_OK;
}

int ldic_export_search(ldic_engine* e, int32_t slot, int32_t level, float* shift, float* score) {
    CHECK(check_slot(e, slot));
    if (level < 0 || level >= e->levels) return set_error(LDIC_ERR_ARGUMENT, "level out of range");
    Slot* s = e->slots[slot];
    HIP_CHECK(hipSetDevice(e->device));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    const size_t n = (size_t)e->lv[level].count();
    if (shift) HIP_CHECK(hipMemcpy(shift, s->lv[level].shift, sizeof(float) * 2 * n, hipMemcpyDeviceToHost));
    if (score) HIP_CHECK(hipMemcpy(score, s->lv[level].score, sizeof(float) * n, hipMemcpyDeviceToHost));
    return LDIC_OK;
}

int ldic_slot_wait(ldic_engine* e, int32_t slot, ldic_timing* t) {
    CHECK(check_slot(e, slot));
    Slot* s = e->slots[slot];
    if (!s->pending) return LDIC_OK;
    HIP_CHECK(hipSetDevice(e->device));
    return finish(e, s, nullptr, t, s->submitted_at);
}

int ldic_slot_output(ldic_engine* e, int32_t slot, float* out_state) {
    return ldic_export_level_state(e, slot, 0, out_state);
}

int ldic_export_level_state(ldic_engine* e, int32_t slot, int32_t level, float* out_state) {
    CHECK(check_slot(e, slot));
    if (level < 0 || level >= e->levels || !out_state) return set_error(LDIC_ERR_ARGUMENT, "invalid level or buffer");
    Slot* s = e->slots[slot];
    HIP_CHECK(hipSetDevice(e->device));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    if (s->pending) {
        CHECK(drain_ring(s));
        s->drained_solve_ms = 0.0;
        s->drained_frames = 0;
        adaptive_feedback(e, s);
    }
    s->pending = false;
    size_t n = (size_t)e->lv[level].count() * state_stride(e->opt.motion_model);
    HIP_CHECK(hipMemcpy(out_state, level_output(e, s->lv[level]), n * sizeof(float), hipMemcpyDeviceToHost));
    return LDIC_OK;
}

int ldic_export_reference(ldic_engine* e, int32_t level, float* reference4, float* metadata, float* inverse) {
    CHECK(check_ready(e));
    if (level < 0 || level >= e->levels) return set_error(LDIC_ERR_ARGUMENT, "level out of range");
    HIP_CHECK(hipSetDevice(e->device));
    const Level& L = e->lv[level];
    size_t n = (size_t)L.count();
    int P = param_count(e->opt.motion_model);
    if (reference4)
        HIP_CHECK(hipMemcpy(reference4, L.reference, sizeof(float4) * (size_t)L.width * L.height, hipMemcpyDeviceToHost));
    if (metadata)
        HIP_CHECK(hipMemcpy(metadata, L.metadata, sizeof(float) * n * metadata_stride(e->opt.motion_model),
                            hipMemcpyDeviceToHost));
    if (inverse) HIP_CHECK(hipMemcpy(inverse, L.inverse, sizeof(float) * n * P * P, hipMemcpyDeviceToHost));
    return LDIC_OK;
}

int ldic_export_coefficients(ldic_engine* e, int32_t slot, int32_t level, float* coefficients) {
    CHECK(check_slot(e, slot));
    if (level < 0 || level >= e->levels || !coefficients) return set_error(LDIC_ERR_ARGUMENT, "invalid level or buffer");
    if (e->opt.interpolation != kBspline) return set_error(LDIC_ERR_STATE, "no coefficient plane in direct modes");
    Slot* s = e->slots[slot];
    HIP_CHECK(hipSetDevice(e->device));
    HIP_CHECK(hipStreamSynchronize(s->stream));
    const Level& L = e->lv[level];
    HIP_CHECK(hipMemcpy(coefficients, s->lv[level].coeff, sizeof(float) * (size_t)L.width * L.height,
                        hipMemcpyDeviceToHost));
    return LDIC_OK;
}

int ldic_op_reduce(const uint8_t* src, int32_t w, int32_t h, uint8_t* dst) {
    if (!src || !dst || w < 2 || h < 2) return set_error(LDIC_ERR_ARGUMENT, "invalid reduction arguments");
    uint8_t *ds = nullptr, *dd = nullptr;
    Arena a;
    int rc = a.alloc(&ds, (size_t)w * h);
    if (rc == LDIC_OK) rc = a.alloc(&dd, (size_t)(w / 2) * (h / 2));
    hipError_t he = hipSuccess;
    if (rc == LDIC_OK) {
        he = hipMemcpy(ds, src, (size_t)w * h, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = launch_reduce(ds, w, h, dd, nullptr);
        if (he == hipSuccess) he = hipMemcpy(dst, dd, (size_t)(w / 2) * (h / 2), hipMemcpyDeviceToHost);
    }
    a.release();
    if (rc != LDIC_OK) return rc;
    if (he != hipSuccess) return set_error(LDIC_ERR_DEVICE, "reduction failed: %s", hipGetErrorString(he));
    return LDIC_OK;
}

int ldic_reduce_device(const uint8_t* device_src, int32_t width, int32_t height, int32_t count, uint8_t* device_dst,
                       void* stream) {
    if (!device_src || !device_dst || width < 2 || height < 2 || count < 1)
        return set_error(LDIC_ERR_ARGUMENT, "invalid reduction arguments");
    const size_t in = (size_t)width * height, out = (size_t)(width / 2) * (height / 2);
    for (int b = 0; b < count; ++b)
        HIP_CHECK(launch_reduce(device_src + b * in, width, height, device_dst + b * out, (hipStream_t)stream));
    return LDIC_OK;
}

int ldic_slot_stream(ldic_engine* e, int32_t slot, void** stream) {
    CHECK(check_slot(e, slot));
    if (!stream) return set_error(LDIC_ERR_ARGUMENT, "stream is NULL");
    *stream = e->slots[slot]->stream;
    return LDIC_OK;
}

int ldic_op_zncc_search(const uint8_t* reference, const uint8_t* target, int32_t width, int32_t height,
                        int32_t subset_size, int32_t subset_stride, int32_t radius, const float* prior,
                        int32_t prior_stride, const uint8_t* mask, float* shift, float* score) {
    if (!reference || !target || !shift || !score) return set_error(LDIC_ERR_ARGUMENT, "images and outputs are required");
    if (subset_size < 4 || subset_size > kMaxSearchSubset || width < subset_size || height < subset_size)
        return set_error(LDIC_ERR_ARGUMENT, "subset_size must be 4..%d and fit the image", kMaxSearchSubset);
    if (subset_stride < 1) return set_error(LDIC_ERR_ARGUMENT, "subset_stride must be positive");
    if (radius < 0 || radius > kMaxSearchRadius || zncc_search_lds_bytes(subset_size, radius) > kMaxSearchLds)
        return set_error(LDIC_ERR_ARGUMENT, "radius %d is outside the supported range for subset %d", radius, subset_size);
    if (prior && prior_stride < 4) return set_error(LDIC_ERR_ARGUMENT, "prior_stride must be at least 4");
    SearchArgs q{};
    q.width = width;
    q.height = height;
    q.grid_w = 1 + (width - subset_size) / subset_stride;
    q.grid_h = 1 + (height - subset_size) / subset_stride;
    q.S = subset_size;
    q.D = subset_stride;
    q.R = radius;
    q.prior_stride = prior_stride;
    const size_t px = (size_t)width * height, n = (size_t)q.grid_w * q.grid_h;
    Arena a;
    uint8_t *dr = nullptr, *dt = nullptr, *dm = nullptr;
    float *dp = nullptr, *ds = nullptr, *dc = nullptr;
    int rc = a.alloc(&dr, px);
    if (rc == LDIC_OK) rc = a.alloc(&dt, px);
    if (rc == LDIC_OK) rc = a.alloc(&ds, 2 * n);
    if (rc == LDIC_OK) rc = a.alloc(&dc, n);
    if (rc == LDIC_OK && prior) rc = a.alloc(&dp, n * prior_stride);
    if (rc == LDIC_OK && mask) rc = a.alloc(&dm, n);
    hipError_t he = hipSuccess;
    if (rc == LDIC_OK) {
        he = hipMemcpy(dr, reference, px, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(dt, target, px, hipMemcpyHostToDevice);
        if (he == hipSuccess && prior) he = hipMemcpy(dp, prior, sizeof(float) * n * prior_stride, hipMemcpyHostToDevice);
        if (he == hipSuccess && mask) he = hipMemcpy(dm, mask, n, hipMemcpyHostToDevice);
        q.reference = dr;
        q.target = dt;
        q.prior = dp;
        q.mask = dm;
        q.shift = ds;
        q.score = dc;
        if (he == hipSuccess) he = launch_zncc_search(q, nullptr);
        if (he == hipSuccess) he = hipMemcpy(shift, ds, sizeof(float) * 2 * n, hipMemcpyDeviceToHost);
        if (he == hipSuccess) he = hipMemcpy(score, dc, sizeof(float) * n, hipMemcpyDeviceToHost);
    }
    a.release();
    if (rc != LDIC_OK) return rc;
    if (he != hipSuccess) return set_error(LDIC_ERR_DEVICE, "ZNCC search failed: %s", hipGetErrorString(he));
    return LDIC_OK;
}

int ldic_op_prefilter(const uint8_t* src, int32_t w, int32_t h, float* coeff) {
    if (!src || !coeff || w < 1 || h < 1) return set_error(LDIC_ERR_ARGUMENT, "invalid prefilter arguments");
    float taps[31];
    fir_taps(taps);
    uint8_t* ds = nullptr;
    float *dc = nullptr, *df = nullptr;
    Arena a;
    int rc = a.alloc(&ds, (size_t)w * h);
    if (rc == LDIC_OK) rc = a.alloc(&dc, (size_t)w * h);
    if (rc == LDIC_OK) rc = a.alloc(&df, 31);
    hipError_t he = hipSuccess;
    if (rc == LDIC_OK) {
        he = hipMemcpy(ds, src, (size_t)w * h, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(df, taps, sizeof taps, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = launch_prefilter(ds, w, h, df, dc, nullptr);
        if (he == hipSuccess) he = hipMemcpy(coeff, dc, sizeof(float) * (size_t)w * h, hipMemcpyDeviceToHost);
    }
    a.release();
    if (rc != LDIC_OK) return rc;
    if (he != hipSuccess) return set_error(LDIC_ERR_DEVICE, "prefilter failed: %s", hipGetErrorString(he));
    return LDIC_OK;
}

int ldic_op_median(const ldic_median_op* c, const float* state, const float* metadata, const float* roi_words,
                   float* filtered) {
    if (!c || !state || !metadata || !filtered) return set_error(LDIC_ERR_ARGUMENT, "config, state, metadata and "
                                                                     "filtered are required");
    if (c->grid_width < 1 || c->grid_height < 1 || c->subset_stride < 1)
        return set_error(LDIC_ERR_ARGUMENT, "grid and stride must be positive");
    if (c->motion_model < 0 || c->motion_model > 2 || c->outlier_policy < 0 || c->outlier_policy > 3)
        return set_error(LDIC_ERR_ARGUMENT, "unknown motion model or outlier policy");
    if (c->median_filter_size < 3 || c->median_filter_size > 9 || !(c->median_filter_size & 1) ||
        (c->median_filter_scope != 0 && c->median_filter_scope != 1))
        return set_error(LDIC_ERR_ARGUMENT, "median_filter_size must be 3, 5, 7 or 9 and the scope 0 or 1");
    const size_t n = (size_t)c->grid_width * c->grid_height;
    const size_t stride = (size_t)state_stride(c->motion_model), mstride = (size_t)metadata_stride(c->motion_model);
    Arena a;
    float *ds = nullptr, *dm = nullptr, *dw = nullptr, *df = nullptr;
    int rc = a.alloc(&ds, n * stride);
    if (rc == LDIC_OK) rc = a.alloc(&dm, n * mstride);
    if (rc == LDIC_OK) rc = a.alloc(&df, n * stride);
    if (rc == LDIC_OK && roi_words) rc = a.alloc(&dw, n * 4);
    hipError_t he = hipSuccess;
    if (rc == LDIC_OK) {
        he = hipMemcpy(ds, state, sizeof(float) * n * stride, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemcpy(dm, metadata, sizeof(float) * n * mstride, hipMemcpyHostToDevice);
        if (he == hipSuccess && dw) he = hipMemcpy(dw, roi_words, sizeof(float) * n * 4, hipMemcpyHostToDevice);
        if (he == hipSuccess) he = hipMemset(df, 0, sizeof(float) * n * stride);
        if (he == hipSuccess) {
            MedianArgs m{};
            m.grid_w = c->grid_width;
            m.grid_count = (int)n;
            m.D = c->subset_stride;
            m.model = c->motion_model;
            m.policy = c->outlier_policy;
            m.size = c->median_filter_size;
            m.scope = c->median_filter_scope;
            m.roi_enabled = dw ? 1 : 0;
            m.seed_available = c->seed_available ? 1 : 0;
            m.state = ds;
            m.metadata = dm;
            m.roi_words = dw;
            m.filtered = df;
            he = c->outlier_policy == kDisabled ? hipMemcpy(df, ds, sizeof(float) * n * stride, hipMemcpyDeviceToDevice)
                                                 : launch_median(m, nullptr);
        }
        if (he == hipSuccess) he = hipMemcpy(filtered, df, sizeof(float) * n * stride, hipMemcpyDeviceToHost);
    }
    a.release();
    if (rc != LDIC_OK) return rc;
    if (he != hipSuccess) return set_error(LDIC_ERR_DEVICE, "median pass failed: %s", hipGetErrorString(he));
    return LDIC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batched heterogeneous registration (no counterpart in the reference, which
// registers one pair per engine).  Several engines (one image pair each: own
// size, subset size, step, warp model and pyramid depth) share one frame
// sequence on one stream: levels are aligned on the finest one and processed
// coarse to fine; at each level the solves of all pairs with the same solver
// variant run in one persistent launch (k_solve_batched), whose grid size is
// chosen by timing candidates on the device (the autotuner).  The whole frame
// is then captured once into a HIP graph and replayed.
// ---------------------------------------------------------------------------
struct ldic_batch {
    struct Group {
        BatchLaunch launch{};
        std::vector<int> members;
        // Block prefix sums for 1, 2 and 4 subsets per workgroup (index log2(waves)).
        int* block_start[3] = {nullptr, nullptr, nullptr};
        int total[3] = {0, 0, 0};
        std::vector<std::pair<int, int>> candidates;   // (waves per workgroup, grid) tried by the autotuner
        std::vector<float> timings;
    };
    std::vector<ldic_engine*> engines;
    std::vector<const uint8_t*> targets;
    int device = 0;
    int max_levels = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // Branches of a frame: pair p's own kernels run on aux[p % kBatchBranches], and the
    // solver groups of a level on aux[g % kBatchBranches]; fork_ev / join_ev order them
    // against the main stream (in a captured frame they become the graph's edges).
    static constexpr int kBatchBranches = 4;
    hipStream_t aux[kBatchBranches] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[kBatchBranches] = {};
    std::vector<std::vector<Group>> steps;  // steps[j]: groups at finest-aligned level j
    Arena arena;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool launched = false;
    bool holds = false;  // the engines' batch_refs count this batch
};

namespace {

int batch_pre_solve(ldic_batch* b, int p, int k, hipStream_t st, const float** seed, const uint32_t** active) {
    ldic_engine* e = b->engines[p];
    Slot* s = e->slots[0];
    SlotLevel& sl = s->lv[k];
    const Level& lev = e->lv[k];
    *seed = nullptr;
    *active = nullptr;
    if (k < e->levels - 1) {
        const Level& src = e->lv[k + 1];
        SeedArgs sa{};
        sa.src_grid_w = src.gw;
        sa.src_grid_h = src.gh;
        sa.dst_grid_w = lev.gw;
        sa.dst_count = lev.count();
        sa.src_scale = (float)(1 << (k + 1));
        sa.dst_scale = (float)(1 << k);
        sa.S = (float)e->opt.subset_size;
        sa.D = (float)e->opt.subset_stride;
        sa.model = e->opt.motion_model;
        sa.track_availability = 0;
        sa.src_state = level_output(e, s->lv[k + 1]);
        sa.permissions = lev.seed_links;
        sa.dst_seed = sl.seed;
        sa.dst_active = sl.active;
        HIP_CHECK(launch_transfer_seed(sa, st));
        *seed = sl.seed;
        *active = sl.active;
    } else if (e->init.kind == LDIC_INIT_ZNCC) {
        SearchArgs q = search_args(e, k, sl, e->init.search_radius);
        q.mask = lev.roi ? lev.selected : nullptr;
        q.seed = sl.seed;
        q.active = sl.active;
        HIP_CHECK(launch_zncc_search(q, st));
        *seed = sl.seed;
        *active = sl.active;
    }
    if (e->opt.interpolation == kBspline) HIP_CHECK(launch_prefilter(sl.target, lev.width, lev.height, e->fir, sl.coeff, st));
    return LDIC_OK;
}

// One frame of the batch, queued on its stream (also what the graph captures).
// One frame, as three kinds of phases separated by barriers on the main stream:
// per-pair work (copy and reductions, then per level the seed transfer or search
// and the prefilter) and per-pair outlier passes, each pair on its own branch
// stream; and the level's solver groups, each group on a branch stream.  Pairs
// and groups are independent, so the branches overlap the small kernels of
// different pairs and the tails of the groups' persistent launches.
int batch_enqueue(ldic_batch* b) {
    hipStream_t st = b->stream;
    constexpr int nb = ldic_batch::kBatchBranches;
    auto fork = [&]() -> int {
        HIP_CHECK(hipEventRecord(b->fork_ev, st));
        for (int i = 0; i < nb; ++i) HIP_CHECK(hipStreamWaitEvent(b->aux[i], b->fork_ev, 0));
        return LDIC_OK;
    };
    auto join = [&]() -> int {
        for (int i = 0; i < nb; ++i) {
            HIP_CHECK(hipEventRecord(b->join_ev[i], b->aux[i]));
            HIP_CHECK(hipStreamWaitEvent(st, b->join_ev[i], 0));
        }
        return LDIC_OK;
    };
    auto branch = [&](size_t i) { return b->aux[i % nb]; };
    CHECK(fork());
    for (size_t p = 0; p < b->engines.size(); ++p) {
        ldic_engine* e = b->engines[p];
        const Level& L0 = e->lv[0];
        Slot* s = e->slots[0];
        hipStream_t ps = branch(p);
        HIP_CHECK(hipMemcpy2DAsync(s->lv[0].target, L0.width, b->targets[p], L0.width, L0.width, L0.height,
                                   hipMemcpyDeviceToDevice, ps));
        for (int k = 1; k < e->levels; ++k)
            HIP_CHECK(launch_reduce(s->lv[k - 1].target, e->lv[k - 1].width, e->lv[k - 1].height, s->lv[k].target, ps));
    }
    for (int j = b->max_levels - 1; j >= 0; --j) {
        // A pair's kernels stay on one branch, in order; it is already forked here.
        for (size_t p = 0; p < b->engines.size(); ++p) {
            if (b->engines[p]->levels <= j) continue;
            const float* seed;
            const uint32_t* active;
            CHECK(batch_pre_solve(b, (int)p, j, branch(p), &seed, &active));
        }
        CHECK(join());
        CHECK(fork());
        for (size_t g = 0; g < b->steps[j].size(); ++g) HIP_CHECK(launch_solve_batched(b->steps[j][g].launch, branch(g)));
        CHECK(join());
        CHECK(fork());
        for (size_t p = 0; p < b->engines.size(); ++p) {
            ldic_engine* e = b->engines[p];
            if (e->levels <= j) continue;
            CHECK(launch_median_level(e, j, e->slots[0]->lv[j], e->opt.outlier_policy == kMedian && j > 0, branch(p)));
        }
    }
    CHECK(join());
    return LDIC_OK;
}

// Point a group's launch at the block prefix sums of `waves` subsets per
// workgroup, with `grid` persistent workgroups.
void batch_shape(ldic_batch::Group& g, int waves, int grid) {
    const int k = waves == 1 ? 0 : waves == 2 ? 1 : 2;
    BatchLaunch& l = g.launch;
    l.waves = waves;
    l.block_start = g.block_start[k];
    l.total = g.total[k];
    l.lds = (size_t)l.wave_stride * sizeof(float) * waves;
    l.grid = grid;
}

int batch_build(ldic_batch* b) {
    b->max_levels = 0;
    for (ldic_engine* e : b->engines) b->max_levels = std::max(b->max_levels, e->levels);
    b->steps.assign(b->max_levels, {});
    for (int j = 0; j < b->max_levels; ++j) {
        // Group the pairs present at level j by solver variant.
        std::vector<ldic_batch::Group>& groups = b->steps[j];
        for (size_t p = 0; p < b->engines.size(); ++p) {
            ldic_engine* e = b->engines[p];
            if (e->levels <= j) continue;
            const int ppl = solver_pixels_per_lane(e->opt.subset_size, e->opt.motion_model);
            const int pitch = solver_tile_pitch(e->opt.subset_size, e->opt.motion_model, e->opt.interpolation,
                                                e->opt.spatial_weighting);
            ldic_batch::Group* g = nullptr;
            for (auto& cand : groups) {
                const BatchLaunch& l = cand.launch;
                if (l.model == e->opt.motion_model && l.interp == e->opt.interpolation &&
                    l.weighted == e->opt.spatial_weighting && l.ppl == ppl && l.pitch == pitch)
                    g = &cand;
            }
            if (!g) {
                groups.emplace_back();
                g = &groups.back();
                g->launch.model = e->opt.motion_model;
                g->launch.interp = e->opt.interpolation;
                g->launch.weighted = e->opt.spatial_weighting;
                g->launch.ppl = ppl;
                g->launch.pitch = pitch;
            }
            g->members.push_back((int)p);
        }
        for (ldic_batch::Group& g : groups) {
            std::vector<SolveArgs> args;
            std::vector<int> start[3];
            for (auto& v : start) v.assign(1, 0);
            int cap = 0;
            for (int p : g.members) {
                ldic_engine* e = b->engines[p];
                SlotLevel& sl = e->slots[0]->lv[j];
                const bool seeded = j < e->levels - 1 || e->init.kind == LDIC_INIT_ZNCC;
                args.push_back(solve_args(e, j, sl, seeded ? sl.seed : nullptr, seeded ? sl.active : nullptr, 0));
                for (int k = 0; k < 3; ++k) {
                    const int w = 1 << k;
                    start[k].push_back(start[k].back() + (e->lv[j].count() + w - 1) / w);
                }
                cap = std::max(cap, e->tile_cap);
            }
            SolveArgs* d_args = nullptr;
            CHECK(b->arena.alloc(&d_args, args.size()));
            HIP_CHECK(hipMemcpy(d_args, args.data(), sizeof(SolveArgs) * args.size(), hipMemcpyHostToDevice));
            for (int k = 0; k < 3; ++k) {
                CHECK(b->arena.alloc(&g.block_start[k], start[k].size()));
                HIP_CHECK(hipMemcpy(g.block_start[k], start[k].data(), sizeof(int) * start[k].size(),
                                    hipMemcpyHostToDevice));
                g.total[k] = start[k].back();
            }
            BatchLaunch& l = g.launch;
            l.pairs = d_args;
            l.npairs = (int)args.size();
            l.wave_stride = cap;
            batch_shape(g, kWavesPerBlock, g.total[2]);
            // Candidates: 1, 2 or 4 waves per workgroup, each with grids of 1-8 full
            // waves of 4-wave-equivalent workgroups per CU and the plain grid.
            g.candidates.clear();
            const int cus = b->engines[0]->prop.multiProcessorCount;
            for (int k = 2; k >= 0; --k) {
                const int w = 1 << k;
                for (int per : {1, 2, 3, 4, 8}) {
                    const int grid = cus * per * (kWavesPerBlock / w);
                    if (grid < g.total[k]) g.candidates.emplace_back(w, grid);
                }
                g.candidates.emplace_back(w, g.total[k]);
            }
        }
    }
    return LDIC_OK;
}

// Times every grid candidate of every group on the device (each solve is
// idempotent for fixed seeds, so the frame's own buffers are reused) and keeps
// the finest of the fastest (below).
constexpr float kTuneSlack = 0.03f;

int batch_tune(ldic_batch* b) {
    CHECK(batch_enqueue(b));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    for (int j = b->max_levels - 1; j >= 0; --j) {
        for (ldic_batch::Group& g : b->steps[j]) {
            g.timings.clear();
            float best = 1e30f;
            for (const auto& cand : g.candidates) {
                batch_shape(g, cand.first, cand.second);
                HIP_CHECK(launch_solve_batched(g.launch, b->stream));  // warm
                HIP_CHECK(hipEventRecord(b->ev0, b->stream));
                HIP_CHECK(launch_solve_batched(g.launch, b->stream));
                HIP_CHECK(hipEventRecord(b->ev1, b->stream));
                HIP_CHECK(hipEventSynchronize(b->ev1));
                float ms = 0.0f;
                HIP_CHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                g.timings.push_back(ms);
                best = std::min(best, ms);
            }
            // Among the shapes within kTuneSlack of the fastest (timed alone), the most
            // workgroups, then the fewest waves per workgroup: in the frame the groups of a
            // level run side by side on the graph's branches, and a finer static partition
            // of the subsets finishes more evenly when some CUs start late.
            std::pair<int, int> best_shape(g.launch.waves, g.launch.grid);
            bool found = false;
            for (size_t i = 0; i < g.candidates.size(); ++i) {
                const auto& cand = g.candidates[i];
                if (g.timings[i] > best * (1.0f + kTuneSlack)) continue;
                if (!found || cand.second > best_shape.second ||
                    (cand.second == best_shape.second && cand.first < best_shape.first)) {
                    best_shape = cand;
                    found = true;
                }
            }
            batch_shape(g, best_shape.first, best_shape.second);
        }
    }
    return LDIC_OK;
}

void batch_release_graph(ldic_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

}  // namespace

extern "C" {

int ldic_batch_create(ldic_engine* const* engines, int32_t count, const uint8_t* const* device_targets,
                      ldic_batch** out) {
    if (!out || !engines || !device_targets || count < 1) return set_error(LDIC_ERR_ARGUMENT, "invalid batch arguments");
    *out = nullptr;
    for (int i = 0; i < count; ++i) {
        CHECK(check_ready(engines[i]));
        if (!device_targets[i]) return set_error(LDIC_ERR_ARGUMENT, "device target %d is NULL", i);
        if (engines[i]->device != engines[0]->device)
            return set_error(LDIC_ERR_ARGUMENT, "all engines of a batch must use one device");
        if (engines[i]->init.kind == LDIC_INIT_TEMPORAL)
            return set_error(LDIC_ERR_ARGUMENT, "temporal initialisation is not supported in a batch");
        if (has_partial(engines[i]))
            return set_error(LDIC_ERR_ARGUMENT, "partial-support ROI subsets are not supported in a batch");
        if (engines[i]->init.divergence_patience || engines[i]->init.reseed_passes)
            return set_error(LDIC_ERR_ARGUMENT, "divergence patience and re-seeding are not supported in a batch");
        for (int k = 0; k < i; ++k)
            if (engines[k] == engines[i]) return set_error(LDIC_ERR_ARGUMENT, "engine %d appears twice", i);
    }
    ldic_batch* b = new ldic_batch();
    b->engines.assign(engines, engines + count);
    b->targets.assign(device_targets, device_targets + count);
    b->device = engines[0]->device;
    auto fail = [&](int rc) {
        ldic_batch_destroy(b);
        return rc;
    };
    hipError_t he = hipSetDevice(b->device);
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreate(&b->ev0);
    if (he == hipSuccess) he = hipEventCreate(&b->ev1);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming);
    for (int i = 0; i < ldic_batch::kBatchBranches && he == hipSuccess; ++i) {
        he = hipStreamCreateWithFlags(&b->aux[i], hipStreamNonBlocking);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&b->join_ev[i], hipEventDisableTiming);
    }
    if (he != hipSuccess) return fail(set_error(LDIC_ERR_DEVICE, "batch set-up failed: %s", hipGetErrorString(he)));
    for (ldic_engine* e : b->engines)
        if (e->slots[0]->pending) return fail(set_error(LDIC_ERR_STATE, "an engine's slot 0 holds a pending frame"));
    int rc = batch_build(b);
    if (rc != LDIC_OK) return fail(rc);
    for (ldic_engine* e : b->engines) ++e->batch_refs;
    b->holds = true;
    *out = b;
    return LDIC_OK;
}

int ldic_batch_tune(ldic_batch* b) {
    if (!b) return set_error(LDIC_ERR_ARGUMENT, "batch is NULL");
    HIP_CHECK(hipSetDevice(b->device));
    batch_release_graph(b);
    CHECK(batch_tune(b));
    // Capture one frame with the chosen grids.
    HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(b);
    hipGraph_t graph = nullptr;
    hipError_t he = hipStreamEndCapture(b->stream, &graph);
    if (rc != LDIC_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (he != hipSuccess) return set_error(LDIC_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(he));
    b->graph = graph;
    HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return LDIC_OK;
}

int ldic_batch_launch(ldic_batch* b, int32_t use_graph) {
    if (!b) return set_error(LDIC_ERR_ARGUMENT, "batch is NULL");
    if (use_graph && !b->exec) return set_error(LDIC_ERR_STATE, "call ldic_batch_tune before launching the graph");
    HIP_CHECK(hipSetDevice(b->device));
    HIP_CHECK(hipEventRecord(b->ev0, b->stream));
    if (use_graph) {
        HIP_CHECK(hipGraphLaunch(b->exec, b->stream));
    } else {
        CHECK(batch_enqueue(b));
    }
    HIP_CHECK(hipEventRecord(b->ev1, b->stream));
    b->launched = true;
    return LDIC_OK;
}

int ldic_batch_wait(ldic_batch* b, double* gpu_ms) {
    if (!b) return set_error(LDIC_ERR_ARGUMENT, "batch is NULL");
    HIP_CHECK(hipSetDevice(b->device));
    HIP_CHECK(hipStreamSynchronize(b->stream));
    if (gpu_ms) {
        float ms = 0.0f;
        if (b->launched) HIP_CHECK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        *gpu_ms = ms;
    }
    return LDIC_OK;
}

int ldic_batch_describe(ldic_batch* b, char* buffer, int32_t capacity) {
    if (!b || !buffer || capacity <= 0) return set_error(LDIC_ERR_ARGUMENT, "invalid buffer");
    std::string text = "[";
    for (int j = b->max_levels - 1; j >= 0; --j)
        for (const ldic_batch::Group& g : b->steps[j]) {
            char item[4096];
            std::string times;
            for (size_t i = 0; i < g.timings.size(); ++i) {
                char t[64];
                std::snprintf(t, sizeof t, "%s[%d,%d,%.4f]", i ? "," : "", g.candidates[i].first,
                              g.candidates[i].second, g.timings[i]);
                times += t;
            }
            std::snprintf(item, sizeof item,
                          "%s{\"level\":%d,\"model\":%d,\"ppl\":%d,\"pairs\":%d,\"waves\":%d,\"blocks\":%d,"
                          "\"grid\":%d,\"candidates_ms\":[%s]}",
                          text.size() > 1 ? "," : "", j, g.launch.model, g.launch.ppl, g.launch.npairs, g.launch.waves,
                          g.launch.total, g.launch.grid, times.c_str());
            text += item;
        }
    text += "]";
    std::snprintf(buffer, (size_t)capacity, "%s", text.c_str());
    return (int)text.size() < capacity ? LDIC_OK : set_error(LDIC_ERR_ARGUMENT, "buffer too small (%zu)", text.size());
}

void ldic_batch_destroy(ldic_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_release_graph(b);
    b->arena.release();
    for (int i = 0; i < ldic_batch::kBatchBranches; ++i)
        if (b->aux[i]) (void)hipStreamSynchronize(b->aux[i]);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
    for (int i = 0; i < ldic_batch::kBatchBranches; ++i) {
        if (b->join_ev[i]) (void)hipEventDestroy(b->join_ev[i]);
        if (b->aux[i]) (void)hipStreamDestroy(b->aux[i]);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->holds)
        for (ldic_engine* e : b->engines) --e->batch_refs;
    delete b;
}

}  // extern "C"
