// Launch interface of the live-dic gfx950 kernels (host side).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ldic {

enum Model : int { kAffine = 0, kTranslation = 1, kQuadratic = 2 };
enum Interp : int { kBspline = 0, kLinear = 1, kKeys = 2 };
enum Policy : int { kFlag = 0, kReplace = 1, kDisabled = 2, kMedian = 3 };

constexpr int kWavesPerBlock = 4;   // most subsets (waves) per persistent batched-solver workgroup
// Subsets (waves) per k_solve workgroup.  One: a wave that converges early
// frees its LDS slice and registers at once instead of holding them until the
// slowest wave of its block has finished.
constexpr int kSolveBlockWaves = 1;

// Sets the calling thread's ldic_last_error() message and returns `code`.
int report_error(int code, const char* message);

inline int param_count(int model) { return model == kQuadratic ? 12 : 6; }
inline int state_stride(int model) { return model == kQuadratic ? 16 : 8; }
inline int metadata_stride(int model) { return model == kQuadratic ? 28 : 16; }

struct SolveArgs {
    int width, height, grid_w, grid_h;
    int S, D;
    int iterations;
    float tolerance, zncc_min;
    float inv_var;          // 1 / sigma^2 of the Gaussian integration weight
    int roi_enabled;
    int preserve_disabled;
    int tile_cap;           // LDS floats available to each wave
    const float4* reference;
    const float* metadata;
    const float* inverse;
    const float* coeff;     // B-spline coefficient plane (B-spline mode)
    const uint8_t* target;  // raw target (direct interpolants)
    const float* seed;      // optional initial state, same layout as the output
    const uint32_t* active; // optional per-subset enable flags
    float* state;
    float* zncc_out;        // optional: ZNCC of the last iteration (at its pre-update warp), NaN if none
    // Partial-support pass (weighted variants only): the listed subsets, their
    // support bits (ceil(S*S/32) words each) and local support boxes.
    const uint32_t* partial_list;
    int partial_count;
    const uint32_t* support;
    const float4* support_boxes;
    // Divergence patience (0: off) and the iteration from which it counts.
    int patience, min_iters;
};

struct SeedArgs {
    int src_grid_w, src_grid_h, dst_grid_w, dst_count;
    float src_scale, dst_scale, S, D;
    int model;
    int track_availability;
    const float* src_state;
    const uint32_t* permissions;  // optional; 31 when absent
    float* dst_seed;
    uint32_t* dst_active;
    uint32_t* seeded_count;       // optional: += subsets that received a weighted seed and are selected
    int projection;               // temporal projection: with track_availability, also mark unselected seeds
};

// Fine-level temporal correction (refine_temporal_seed, full_pyramid.wgsl:451-515).
struct RefineArgs {
    SeedArgs edge;                // coarse -> fine geometry; src_state is the coarse history
    const float* fine_history;
    const uint32_t* fine_ages;
    const uint32_t* coarse_ages;
    uint32_t max_age;
};

struct MedianArgs {
    int grid_w, grid_count, D, model, policy, size, scope, roi_enabled, seed_available;
    const float* state;
    const float* metadata;
    const float* roi_words;
    float* filtered;
};

struct ZnccArgs {
    int width, height, grid_w, grid_h, S, D, interpolation;
    const float4* reference;
    const float* state;
    const float* coeff;
    const uint8_t* target;
    float* out;
};

struct SearchArgs {
    int width, height, grid_w, grid_h, S, D, R;
    const uint8_t* reference;  // uint8 planes of one level, row pitch = width
    const uint8_t* target;
    const float* prior;        // optional packed state: search around (prior[0], prior[3])
    int prior_stride;
    const uint8_t* mask;       // optional: search only where non-zero
    float* shift;              // grid * 2: best shift with sub-pixel refinement
    float* score;              // grid: its ZNCC, -2 when no shift is valid, NaN when skipped
    float* seed;               // optional packed state: translation from the search, other terms from the prior
    int seed_stride;
    uint32_t* active;          // optional: 1 where searched, 0 where skipped
};

struct PrepArgs {
    int width, height, grid_w, grid_h, S, D, model, weighted;
    int direct;                 // 1 for the linear and Keys interpolants
    const double* intensity;    // Gaussian-filtered (B-spline) or raw intensity, float64
    const double* grad_x;
    const double* grad_y;
    const double* weights;      // S*S integration weights (Gaussian mode) or null
    double weight_sum;
    const uint8_t* selected;    // optional ROI selection
    float4* reference;          // receives the centred intensity when D == S
    float* metadata;
    float* inverse;
    uint8_t* valid;             // reference validity before the ROI override
};

// Float64 statistics of partial-support subsets written into a level's metadata and
// inverse rows (translation or affine; 16-float rows, 36-float inverses).
struct PartialStatsArgs {
    int width, grid_w, S, D, model, count;
    const float4* reference;   // prepared plane (f, gx, gy, 0)
    const double* weights;     // S*S integration weights, or null (uniform)
    const uint32_t* indices;   // flat subset index of each partial subset
    const uint32_t* support;   // ceil(S*S/32) words per subset
    float* metadata;
    float* inverse;
    float* diagnostics;        // 4 per subset: admitted, rank, normalised condition, RMS contrast
};
hipError_t launch_partial_stats(const PartialStatsArgs& a, hipStream_t s);

// Pick the register-cached solver variant for a subset width.
int solver_pixels_per_lane(int S, int model);
// Compile-time LDS row pitch of the solver variant (0: the staged width).
int solver_tile_pitch(int S, int model, int interp, int weighted);
size_t solver_lds_bytes(int S, int model, int* tile_cap);

hipError_t launch_reduce(const uint8_t* src, int w, int h, uint8_t* dst, hipStream_t s);
hipError_t launch_prefilter(const uint8_t* src, int w, int h, const float* fir31, float* coeff, hipStream_t s);
hipError_t launch_solve(int model, int interp, int weighted, const SolveArgs& a, hipStream_t s);

// One persistent launch over the subset grids of several pairs that share
// model, interpolant, register bucket (ppl) and weighting.  pairs and
// block_start (npairs + 1 prefix sums of ceil(grid / 4)) are device arrays;
// wave_stride is the LDS floats per wave (the largest tile_cap of the pairs).
struct BatchLaunch {
    int model, interp, weighted, ppl, pitch;
    int waves;  // subsets (waves) per workgroup, 1, 2 or 4; block_start counts blocks of this many
    const SolveArgs* pairs;
    const int* block_start;
    int npairs, total, grid;
    size_t lds;
    int wave_stride;
};
hipError_t launch_solve_batched(const BatchLaunch& b, hipStream_t s);
hipError_t launch_transfer_seed(const SeedArgs& a, hipStream_t s);
hipError_t launch_median(const MedianArgs& a, hipStream_t s);
// Temporal history (full_pyramid.wgsl:519-575, full_pyramid_backend.py:660-677).
hipError_t launch_temporal_reset(int count, int model, float* history, uint32_t* ages, hipStream_t s);
hipError_t launch_links_from_selected(int count, const uint8_t* selected, uint32_t* links, hipStream_t s);
hipError_t launch_temporal_merge(int count, int model, const float* raw, const float* output, int use_filtered,
                                 uint32_t max_age, float* history, uint32_t* ages, uint32_t* accepted_count,
                                 hipStream_t s, uint32_t* expired_count = nullptr);
hipError_t launch_fill_u32(int count, uint32_t value, uint32_t* p, hipStream_t s);
// Five float64 values: accepted count, replaced count, and the sums of tx, ty and
// iterations over the accepted subsets of a packed state field.  `out` holds
// kSummaryDoubles doubles (the summary, then per-workgroup partials).
constexpr int kSummaryMaxBlocks = 256;
constexpr int kSummaryDoubles = 8 + 5 * kSummaryMaxBlocks;
hipError_t launch_field_summary(const float* state, int count, int model, double* out, hipStream_t s);
// Velocity prediction (full_pyramid.wgsl:83-148): extrapolate the history from
// the previous accepted measurement; save the measurement before the merge.
hipError_t launch_temporal_predict(int count, int model, const float* history, const float* previous,
                                   const uint32_t* ages, const uint32_t* previous_ages, float* prediction,
                                   uint32_t* predicted_count, hipStream_t s);
hipError_t launch_temporal_save_previous(int count, int model, const float* raw, const float* output,
                                         const float* history, const uint32_t* ages, float* previous,
                                         uint32_t* previous_ages, hipStream_t s);
hipError_t launch_temporal_refine(const RefineArgs& a, hipStream_t s);
// dst[index[i] * row + k] = src[i * row + k]
hipError_t launch_scatter_rows(int count, int row, const uint32_t* index, const float* src, float* dst,
                               hipStream_t s);
hipError_t launch_affine_zncc(const ZnccArgs& a, hipStream_t s);
// Neighbour re-seeding of unaccepted subsets from their nearest accepted
// subset (jump flooding); writes seed and active for a preserving retry.
hipError_t launch_reseed(int grid_w, int grid_h, int model, float D, const float* state, const float* metadata,
                         const uint8_t* selected, uint32_t* near_a, uint32_t* near_b, float* seed, uint32_t* active,
                         hipStream_t s);

// Exhaustive integer ZNCC search; LDS bytes per workgroup for (S, R).
size_t zncc_search_lds_bytes(int S, int R);
hipError_t launch_zncc_search(const SearchArgs& a, hipStream_t s);
// mask[i] = selected[i] && (status != 1 || zncc[i] < zncc_min); zncc nullable.
hipError_t launch_retry_mask(int count, int model, const float* state, const float* zncc, float zncc_min,
                             const uint8_t* selected, uint8_t* mask, hipStream_t s);

// Reference preparation (float64).
hipError_t launch_gaussian(const uint8_t* src, int w, int h, const double* w7, double* tmp, double* out,
                           hipStream_t s);
hipError_t launch_to_double(const uint8_t* src, int w, int h, double* out, hipStream_t s);
hipError_t launch_bspline(double* plane, double* tmp, int w, int h, hipStream_t s);
hipError_t launch_gradients(const double* coeff, const double* intensity, int w, int h, int bspline,
                            double* gx, double* gy, float4* reference, hipStream_t s);
hipError_t launch_subset_prep(const PrepArgs& a, hipStream_t s);

}  // namespace ldic
