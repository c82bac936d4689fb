/*
 * live-dic for MI355X: C ABI of the gfx950 IC-GN engine (liblivedic.so).
 *
 * The reference has no native library: its Python supervisor drives WGSL
 * compute pipelines through wgpu-py.  Each entry point below replaces the part
 * of that supervisor-to-device boundary named in its comment (paths relative to
 * DIC/LucasKanadeICGN/python/src/live_dic/).  Plain pointers and sizes only;
 * host arrays are borrowed for the duration of the call.  Every function
 * returns 0 on success or a negative LDIC_ERR_* code; ldic_last_error() then
 * holds a message for the calling thread.  An engine is not thread-safe.
 */
#ifndef LIVEDIC_H
#define LIVEDIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDIC_ABI_VERSION 1
#define LDIC_MAX_LEVELS 4

enum {
    LDIC_OK = 0,
    LDIC_ERR_ARGUMENT = -1,  /* invalid argument; no device work was done */
    LDIC_ERR_STATE = -2,     /* call out of order (no reference, closed slot) */
    LDIC_ERR_DEVICE = -3,    /* HIP runtime or kernel launch failure */
    LDIC_ERR_MEMORY = -4     /* device or pinned allocation failed */
};

/* Motion model, interpolation and outlier codes follow the WGSL override
 * constants (motion.py:9-12, interpolation.py:12-16, outlier.py:9-14). */
typedef struct ldic_options {
    int32_t motion_model;       /* 0 affine, 1 translation, 2 quadratic */
    int32_t interpolation;      /* 0 B-spline, 1 bilinear, 2 Keys cubic */
    int32_t subset_size;        /* S */
    int32_t subset_stride;      /* D, 1..S */
    int32_t spatial_weighting;  /* 0 uniform, 1 Gaussian */
    float gaussian_sigma;       /* pixels, used when spatial_weighting == 1 */
    int32_t outlier_policy;     /* 0 flag, 1 replace, 2 disabled, 3 median */
    int32_t median_filter_size; /* 3, 5, 7 or 9 */
    int32_t median_filter_scope;/* 0 rejected, 1 all */
    int32_t pyramid_levels;     /* 1..4 */
    int32_t device_ordinal;     /* HIP device index */
} ldic_options;

typedef struct ldic_level_config {
    int32_t iterations; /* 1..30 */
    float tolerance;
    float zncc_min;
} ldic_level_config;

/* Static ROI tables of one level, built by the host planner (roi.py:178-263,
 * 358-437).  All pointers may be NULL: no ROI at that level. */
typedef struct ldic_level_roi {
    const uint8_t* selected;        /* grid_h*grid_w flags: full window inside the mask */
    const float* neighbour_words;   /* grid*4: neighbour permission words (metadata tail) */
    const uint32_t* seed_links;     /* grid: transfer permissions into this level */
} ldic_level_roi;

typedef struct ldic_timing {
    double upload_ms;    /* host copy into pinned staging and the device copy issue */
    double gpu_ms;       /* device time from the first upload to the last kernel (HIP events) */
    double solve_ms;     /* summed device time of the finest-level solver kernel over `frames` */
    double readback_ms;  /* wait and copy of the packed state into the caller's buffer */
    double wall_ms;      /* whole call */
    int32_t frames;      /* frames completed by this call (resident slots may queue several) */
    int32_t temporal_seeded;   /* temporal mode: coarsest subsets seeded from the history (last frame) */
    int32_t temporal_updates;  /* temporal mode: finest subsets merged into the history (last frame) */
    int32_t velocity_predicted; /* velocity prediction: finest subsets extrapolated (last frame) */
    int32_t levels_used;       /* pyramid levels the last frame solved (the adaptive pyramid may use 1) */
    int32_t temporal_expired;  /* adaptive pyramid: finest history entries that expired (last frame) */
    /* Summary of the finest field copied into out_state, computed on the device with it
     * (the statistics of RegistrationResult, engine.py:105-193): subsets with status > 0.5
     * and > 1.5, and float64 sums of tx, ty and iterations over the former.  accepted and
     * replaced are -1 when the call returned no field. */
    int64_t accepted;
    int64_t replaced;
    double sum_translation_x;
    double sum_translation_y;
    double sum_iterations;
} ldic_timing;

/* Device selection and kernel set-up: ICGNWGPU.__init__ device/pipeline
 * creation, wgpu_backend.py:889-1152, and ICGNWGPUFullPyramid.__init__,
 * full_pyramid_backend.py:202-431. */
int ldic_create(const ldic_options* options, struct ldic_engine** out_engine);
/* ICGNWGPU.close / ICGNWGPUFullPyramid.close (wgpu_backend.py:3078-3093,
 * full_pyramid_backend.py:1904-1919).  NULL is accepted. */
void ldic_destroy(struct ldic_engine* engine);
const char* ldic_last_error(void);
int ldic_abi_version(void);
/* adapter_info / device_name attributes (wgpu_backend.py:113-144). */
int ldic_device_name(struct ldic_engine* engine, char* buffer, int32_t capacity);

/* set_reference: the reference pyramid, then per level the Gaussian, the
 * recursive B-spline prefilter, gradients, per-subset statistics and inverse
 * Hessians, all on the device (replaces _reduce_half and _reference_arrays on
 * the host, wgpu_backend.py:318-334, 589-823, and set_reference
 * full_pyramid_backend.py:1135-1368).  configs and rois hold one entry per
 * level, finest first.  out_reference_valid (fine grid, nullable) receives the
 * reference validity before ROI exclusion.  Images are at least 16x16 and have
 * fewer than 2^28 pixels. */
int ldic_set_reference(struct ldic_engine* engine, const uint8_t* image, int32_t width, int32_t height,
                       int64_t row_stride, const ldic_level_config* configs, const ldic_level_roi* rois,
                       uint8_t* out_reference_valid);
int ldic_level_shape(struct ldic_engine* engine, int32_t level, int32_t* width, int32_t* height,
                     int32_t* grid_width, int32_t* grid_height);
int ldic_set_level_config(struct ldic_engine* engine, int32_t level, const ldic_level_config* config);

/* register: upload, target pyramid, per-level prefilter, solve, outlier pass
 * and seed transfer, readback of the finest packed state (gh*gw*8 or *16
 * floats) -- full_pyramid_backend.py:1436-1569, 1693-1832. */
int ldic_register(struct ldic_engine* engine, const uint8_t* target, int64_t row_stride, float* out_state,
                  ldic_timing* timing);

/* Consecutive-pair registration: the reference is replaced by `reference` and
 * prepared on the device in the same queued sequence as the target's solve.
 * The reference does this only in its browser path (web/core/reference-gpu.js:
 * 90-117 driving shaders/reference.wgsl, then engine.registerPair); the native
 * supervisor would call set_reference then register.  The dimensions, options,
 * ROI tables and level configurations of the last ldic_set_reference are kept.
 * out_reference_valid (fine grid, nullable) receives the pair's validity. */
int ldic_register_pair(struct ldic_engine* engine, const uint8_t* reference, int64_t reference_stride,
                       const uint8_t* target, int64_t target_stride, float* out_state,
                       uint8_t* out_reference_valid, ldic_timing* timing);
/* The reference buffers are engine-wide: both pair calls, and
 * ldic_prepare_reference_resident, refuse (LDIC_ERR_STATE) while another slot
 * has queued work; references in flight on several slots need one engine each. */
int ldic_register_pair_resident(struct ldic_engine* engine, int32_t slot, const uint8_t* device_reference,
                                const uint8_t* device_target, int64_t row_stride);

/* register_stream slots (full_pyramid_backend.py:1834-1902): submit copies the
 * target into the slot's pinned staging and queues its work; retire waits and
 * copies the result out, in submission order chosen by the caller. */
int ldic_set_slots(struct ldic_engine* engine, int32_t count);
int ldic_submit(struct ldic_engine* engine, int32_t slot, const uint8_t* target, int64_t row_stride);
int ldic_retire(struct ldic_engine* engine, int32_t slot, float* out_state, ldic_timing* timing);

/* register_seeded (wgpu_backend.py:2773-2819, solve_seeded icgn.wgsl:1001-1008)
 * on the finest level, with optional final affine ZNCC (affine_zncc.wgsl).
 * out_zncc receives NaN where no correlation is defined. */
int ldic_register_seeded(struct ldic_engine* engine, const uint8_t* target, int64_t row_stride,
                         const float* seed_state, const uint8_t* active, int32_t include_zncc, float* out_state,
                         float* out_zncc, ldic_timing* timing);

/* ---- Seeding (no counterpart in the reference).  The reference seeds the
 * coarsest pyramid level with zeros, a temporal prediction, a Fourier estimate
 * or an external CPM-HOG matcher (full_pyramid_backend.py:1693-1716, seeds
 * written to the coarsest level's seed_state with its active_mask); its
 * nearest ZNCC formulation scores a few candidate translations
 * (shaders/spatial_seed.wgsl:65-190).  This engine adds an exhaustive integer
 * ZNCC search on the uint8 level planes (exact integer moments, float64
 * quotient) in two places: as the coarsest-level initialiser, and as the
 * fallback of seeded registration. ---- */
enum { LDIC_INIT_ZERO = 0, LDIC_INIT_ZNCC = 1, LDIC_INIT_TEMPORAL = 2 };

typedef struct ldic_initialization {
    int32_t kind;            /* LDIC_INIT_ZERO, LDIC_INIT_ZNCC (coarsest level searched over +-search_radius) or
                                LDIC_INIT_TEMPORAL: the last accepted finest field, projected onto the coarsest
                                grid, seeds every frame (initialization="temporal",
                                full_pyramid.wgsl:437-575, full_pyramid_backend.py:433-881) */
    int32_t search_radius;   /* coarsest-level pixels, 0..64 */
    float fallback_zncc;     /* seeded registration: subsets that do not converge, or whose ZNCC in the last
                                solver iteration is below this, are searched again around their seed and
                                re-solved; <= -1 disables */
    int32_t fallback_radius; /* finest-level pixels around the seed, 0..64 */
    uint32_t temporal_max_age; /* LDIC_INIT_TEMPORAL: history expires after more than this many rejected frames;
                                  0xFFFFFFFE keeps it (temporal_max_age=None) */
    int32_t temporal_prediction;      /* LDIC_INIT_TEMPORAL: 1 extrapolates subsets with two consecutive fresh
                                         accepted measurements (temporal_prediction="velocity",
                                         full_pyramid.wgsl:83-137); 0 none */
    int32_t temporal_fine_correction; /* LDIC_INIT_TEMPORAL, two levels: the finest seed is the finest history
                                         plus the coarse solve's correction of the coarse history
                                         (refine_temporal_seed, full_pyramid.wgsl:451-515) */
    int32_t divergence_patience;      /* 1..30: an unconverged finest-level solve stops after this many
                                         iterations without a decrease of its change (icgn.wgsl:103-111);
                                         0 off */
    int32_t divergence_min_iterations; /* 1..30: iterations before the patience counts (default 4) */
    int32_t adaptive_pyramid; /* LDIC_INIT_TEMPORAL, two levels (pyramid_policy="adaptive",
                                 full_pyramid_backend.py:1775-1791): after a frame whose merged finest subsets
                                 reach adaptive_acceptance of the reference-valid ones and none expired, the
                                 next frame solves the finest level only, seeded straight from the history;
                                 otherwise both levels.  One frame in flight at a time. */
    float adaptive_acceptance; /* (0, 1] */
    int32_t reseed_passes;    /* seeded registration, 0..8: after the fallback, subsets still unaccepted are
                                 seeded from the warp of their nearest accepted subset (one with at least two
                                 accepted neighbours), extrapolated to their centre, and re-solved; translation
                                 and affine.  An addition without a reference counterpart. */
} ldic_initialization;

/* Partial-support subsets of one level (roi_policy="partial"; roi.py:266-355,
 * partial.py:26-130): their list, support bits (ceil(S*S/32) little-endian
 * words each, row-major pixels) and local support boxes (xmin, ymin, xmax,
 * ymax), from ldic_roi_planner_partial.  The device computes each support's
 * float64 statistics from the prepared reference of that level and overwrites
 * its metadata row and inverse Hessian (validity -4 when admitted, 0 and a zero
 * inverse otherwise); out_diagnostics (nullable, 4 floats per subset) receives
 * admitted (1/0), Hessian rank, normalised condition and weighted RMS contrast.
 * Call after ldic_set_reference; each frame then solves these subsets in a
 * partial pass after the full-window pass (solve_partial*, icgn.wgsl:1040-1085).
 * Translation and affine only. */
int ldic_prepare_partial(struct ldic_engine* engine, int32_t level, int32_t count, const uint32_t* indices,
                         const uint32_t* support_words, const float* boxes, float* out_diagnostics);

/* Forget the temporal history (reset_temporal_state, full_pyramid_backend.py:
 * 660-677).  ldic_set_reference and a change of initialisation also reset it.
 * Frames must be registered in chronological order in temporal mode. */
int ldic_reset_temporal(struct ldic_engine* engine);
/* ROI links of the temporal projection from the finest grid onto the
 * coarsest (seed_permissions with source scale 1, full_pyramid_backend.py:
 * 608-620); NULL restores "all links".  Call after ldic_set_reference. */
int ldic_set_temporal_links(struct ldic_engine* engine, const uint32_t* coarse_links);

/* Takes effect for the next queued frame; rejected while a slot is pending.
 * The search needs (2R+S)(ceil((2R+1)/4)+ceil(S/4))*4 + S*ceil(S/4)*4 +
 * 8(2R+1)(2R+S-1) bytes of LDS, at most 64 KiB: search large displacements
 * on a coarser level. */
int ldic_set_initialization(struct ldic_engine* engine, const ldic_initialization* init);

/* Seeded registration of a resident frame: seeds (packed state layout) and
 * active flags (uint32) are device pointers, e.g. the output of the learned
 * seed network.  Finish with ldic_slot_wait / ldic_slot_output. */
int ldic_register_seeded_resident(struct ldic_engine* engine, int32_t slot, const uint8_t* device_target,
                                  int64_t row_stride, const float* device_seed_state, const uint32_t* device_active);

/* Queues the preparation of a new reference (device plane) on a slot's stream,
 * in order with that slot's resident registrations: the per-pair reference of
 * seeded consecutive-pair registration.  Same preparation as
 * ldic_set_reference (full_pyramid_backend.py:1135-1368); dimensions, options,
 * ROI tables and level configurations of the last ldic_set_reference are kept. */
int ldic_prepare_reference_resident(struct ldic_engine* engine, int32_t slot, const uint8_t* device_reference,
                                    int64_t row_stride);

/* The last search of a slot's level: shift (grid*2, level pixels) and ZNCC
 * score (grid; -2 where no shift had a valid window, NaN where skipped). */
int ldic_export_search(struct ldic_engine* engine, int32_t slot, int32_t level, float* shift, float* score);

/* Stand-alone search on host planes (parity tests): every integer shift in
 * [-radius, radius]^2 around the prior translation (packed state, components 0
 * and 3; NULL for none) of every subset with mask != 0 (NULL: all). */
int ldic_op_zncc_search(const uint8_t* reference, const uint8_t* target, int32_t width, int32_t height,
                        int32_t subset_size, int32_t subset_stride, int32_t radius, const float* prior,
                        int32_t prior_stride, const uint8_t* mask, float* shift, float* score);

/* ---- Batched heterogeneous registration (no counterpart in the reference,
 * which registers one pair per engine).  Each engine, prepared with
 * ldic_set_reference, contributes one pair: its own size, subset size, step,
 * warp model and pyramid depth.  A frame of the batch runs every pair's
 * sequence (full_pyramid_backend.py:1436-1569) on one stream, levels aligned on
 * the finest and processed coarse to fine; the solves of all pairs that share
 * a solver variant run in one persistent launch whose grid size is picked by
 * timing candidates on the device.  The targets are device planes (row pitch =
 * width) read at every launch; results stay in each engine's slot 0 and are
 * read with ldic_slot_output after ldic_batch_wait.  While the batch exists its
 * engines refuse (LDIC_ERR_STATE) every call that would re-prepare, reconfigure or
 * queue work on them, and ldic_destroy leaves them alive with that error set:
 * destroy the batch first. ---- */
typedef struct ldic_batch ldic_batch;
int ldic_batch_create(struct ldic_engine* const* engines, int32_t count, const uint8_t* const* device_targets,
                      ldic_batch** out_batch);
/* Runs one frame eagerly, times the grid candidates of every persistent solve,
 * keeps the fastest and captures one frame into a HIP graph. */
int ldic_batch_tune(ldic_batch* batch);
/* Queues one frame: the graph (use_graph != 0, after ldic_batch_tune) or the
 * same launches issued one by one. */
int ldic_batch_launch(ldic_batch* batch, int32_t use_graph);
/* Waits for the queued frames; gpu_ms (nullable) is the device time of the last one. */
int ldic_batch_wait(ldic_batch* batch, double* gpu_ms);
/* JSON description of the persistent launches and the autotuner's timings. */
int ldic_batch_describe(ldic_batch* batch, char* buffer, int32_t capacity);
void ldic_batch_destroy(ldic_batch* batch);

/* Resident frames for throughput measurement: the target is already in device
 * memory (row pitch row_stride bytes).  Queues the same per-frame work on the
 * slot's stream without host copies.  Further resident frames may be queued on
 * the same slot before waiting (stream order keeps them sequential);
 * ldic_slot_wait blocks until all are done and reports their summed solver time. */
int ldic_register_resident(struct ldic_engine* engine, int32_t slot, const uint8_t* device_target,
                           int64_t row_stride);
int ldic_slot_wait(struct ldic_engine* engine, int32_t slot, ldic_timing* timing);
int ldic_slot_output(struct ldic_engine* engine, int32_t slot, float* out_state);

/* Parity hooks: the prepared reference of one level (the arrays returned by
 * _reference_arrays, wgpu_backend.py:589-823) and the target coefficient
 * plane of a slot's last frame (prefilter_y output, icgn.wgsl:181-214). */
int ldic_export_reference(struct ldic_engine* engine, int32_t level, float* reference4, float* metadata,
                          float* inverse);
int ldic_export_coefficients(struct ldic_engine* engine, int32_t slot, int32_t level, float* coefficients);
int ldic_export_level_state(struct ldic_engine* engine, int32_t slot, int32_t level, float* out_state);

/* Stand-alone operators on host buffers (parity tests): reduce_fused
 * (full_pyramid.wgsl:215-249) and prefilter_x/prefilter_y (icgn.wgsl:145-214). */
int ldic_op_reduce(const uint8_t* source, int32_t width, int32_t height, uint8_t* destination);
int ldic_op_prefilter(const uint8_t* source, int32_t width, int32_t height, float* coefficients);

/* The outlier pass over a given packed state: the median_filter / UOD entry of
 * icgn.wgsl:1325-1605 dispatched on a written state buffer, as the reference's
 * own tests drive it (python/tests/test_gpu_regressions.py:860-905).  state and
 * filtered hold grid_width*grid_height*(8|16) floats; metadata 16 (28 quadratic)
 * floats per subset, of which the validity word (index 1) is read; roi_words is
 * NULL or 4 floats per subset (the ROI tail of the metadata).  Codes follow
 * ldic_options; seed_available marks a pass over a coarser pyramid level whose
 * median also fills unusable subsets for the seed transfer. */
typedef struct ldic_median_op {
    int32_t grid_width, grid_height, subset_stride, motion_model;
    int32_t outlier_policy, median_filter_size, median_filter_scope, seed_available;
} ldic_median_op;
int ldic_op_median(const ldic_median_op* config, const float* state, const float* metadata, const float* roi_words,
                   float* filtered);

/* ---- ROI planning on the device: the static per-level tables of the reference's
 * host planner (roi.py:63-97 selection and reduced masks, :178-263 neighbour and
 * median link words, :266-355 partial supports, :358-437 seed links), computed
 * from the boolean mask by HIP kernels.  A planner holds the mask of every
 * pyramid level (reduced on the device, roi.py:89-97) and their summed-area
 * tables; plan() fills one level's tables; seed links read two planned levels. ---- */
typedef struct ldic_roi_planner ldic_roi_planner;
typedef struct ldic_roi_plan_options {
    int32_t subset_size, subset_stride;
    int32_t median_filter_size;  /* 3: link bits only; 5-9: also the 81-bit median words */
    int32_t partial;             /* roi_policy="partial": supports of incomplete windows (S <= 64) */
    int32_t visibility;          /* roi_truncation="visibility" */
} ldic_roi_plan_options;
/* mask: width*height bytes, non-zero = allowed. */
int ldic_roi_planner_create(const uint8_t* mask, int32_t width, int32_t height, int32_t levels,
                            int32_t device_ordinal, ldic_roi_planner** out_planner);
/* Size, contents (nullable) and "every pixel allowed" flag of one level's mask. */
int ldic_roi_planner_mask(ldic_roi_planner* planner, int32_t level, int32_t* width, int32_t* height,
                          uint8_t* out_mask, int32_t* all_allowed);
/* One level's tables over its gh*gw grid: selection, 3x3 link bits, median words
 * (gh*gw*4, nullable; filled for median_filter_size > 3), support coverage (nullable;
 * 1 for full windows, the support fraction of partial ones) and the number of
 * partial subsets, fetched by ldic_roi_planner_partial. */
int ldic_roi_planner_plan(ldic_roi_planner* planner, int32_t level, const ldic_roi_plan_options* options,
                          uint8_t* out_selected, uint32_t* out_neighbour_bits, float* out_median_words,
                          float* out_coverage, int32_t* out_partial_count);
/* Partial subsets of the last plan of `level`, in subset order: flat indices, support
 * bits (ceil(S*S/32) words each, row-major, pixel 0 in bit 0) and boxes (local
 * xmin, ymin, xmax, ymax). */
int ldic_roi_planner_partial(ldic_roi_planner* planner, int32_t level, uint32_t* out_indices,
                             uint32_t* out_support_words, float* out_support_boxes);
/* Transfer links into destination_level from source_level (both planned): the
 * pyramid transfer (source coarser) or the temporal projection (source finer). */
int ldic_roi_planner_seed_links(ldic_roi_planner* planner, int32_t source_level, int32_t destination_level,
                                uint32_t* out_links);
void ldic_roi_planner_destroy(ldic_roi_planner* planner);

/* ---- Workload generation and device memory (no counterpart in the reference,
 * whose benchmarks read PNG fixtures; used for the synthetic speckle pairs of
 * the measurements and for training data of the learned seed). ---- */

/* Displacement u(X) in reference coordinates: a quadratic polynomial about
 * (origin_x, origin_y) in the quadratic state order plus one sinusoid per
 * component, amplitude * sin(kx X + ky Y + phase). */
typedef struct ldic_field {
    int32_t enabled;
    double origin_x, origin_y;
    double poly[12];
    double wave[2][4];
} ldic_field;

/* Gaussian-blob speckle texture: background + sum amplitude exp(-r^2/2 sigma^2).
 * Blobs (x, y, amplitude, 0 as float32, image coordinates) are sorted into
 * square cells of `cell` pixels starting at cell_origin; cell_start holds
 * cells_x*cells_y+1 offsets.  Both live on the device.  Blobs further than
 * `cell` pixels from a point are ignored. */
typedef struct ldic_speckle {
    const float* device_blobs;
    const int32_t* device_cell_start;
    int32_t cells_x, cells_y, cell;
    float cell_origin;  /* image coordinate of the first cell's corner (both axes) */
    float sigma, background;
} ldic_speckle;

/* Renders the deformed texture (field NULL: the reference) with Gaussian noise
 * of noise_sigma grey levels, rounded and clipped to uint8, into device memory.
 * device_float_out (nullable) receives the noiseless float image. */
int ldic_render_speckle(const ldic_speckle* speckle, const ldic_field* field, int32_t width, int32_t height,
                        float noise_sigma, uint64_t noise_seed, uint8_t* device_out, float* device_float_out,
                        void* stream);

/* Batched rendering: `count` images of width x height into device_out (image b
 * at b*width*height).  device_fields (nullable: all undeformed) holds one field
 * per image (enabled == 0 for an undeformed one); device_offsets (nullable)
 * holds per-image texture offsets (x, y) so that crops of one large texture
 * can be rendered.  Image b draws its noise from noise_seed advanced by b. */
int ldic_render_speckle_batch(const ldic_speckle* speckle, const ldic_field* device_fields,
                              const int32_t* device_offsets, int32_t count, int32_t width, int32_t height,
                              float noise_sigma, uint64_t noise_seed, uint8_t* device_out, void* stream);

/* reduce_fused (full_pyramid.wgsl:215-249) of `count` device images of
 * width x height (image b at b*width*height) into (width/2) x (height/2)
 * images, queued on `stream`. */
int ldic_reduce_device(const uint8_t* device_src, int32_t width, int32_t height, int32_t count, uint8_t* device_dst,
                       void* stream);

/* The HIP stream of a slot, so that device work of the caller (the learned
 * seed network) is queued in order with the slot's registrations. */
int ldic_slot_stream(struct ldic_engine* engine, int32_t slot, void** stream);

/* ---- Learned seed network (no counterpart in the reference; the reference's
 * nearest initialisers are listed above ldic_set_initialization).  Kernels of
 * a small correlation CNN, forward and training: bfloat16 activations and
 * weights (uint16_t storage, NHWC with a channel stride and offset so that a
 * concatenation is a set of channel slices of one buffer), float32
 * accumulation on the matrix cores, float32 gradients.  All pointers are
 * device pointers; work is queued on `stream`. ---- */

/* 3x3 convolution, padding 1, stride 1 or 2.  Weights are (out_channels, KP)
 * with KP = 9*in_channels rounded up to 16, column k = tap*in_channels + ci,
 * tap = 3*ky + kx; the padding columns are zero. */
typedef struct ldic_nn_conv {
    int32_t batch, height, width;
    int32_t in_channels, in_stride, in_offset;
    int32_t out_height, out_width;   /* (height-1)/stride+1, (width-1)/stride+1 */
    int32_t out_channels, out_stride, out_offset;
    int32_t stride;
    int32_t out_channel_tile;        /* forward: output channels per workgroup, 32, 64 or 128; 0 picks */
} ldic_nn_conv;

/* y = act(conv(x) + bias): bfloat16 into y, or float32 into y_f32 when y_f32
 * is not NULL; leaky != 0 applies the leaky ReLU (slope 0.1). */
int ldic_nn_conv_forward(const ldic_nn_conv* conv, const uint16_t* x, const uint16_t* w, const float* bias,
                         int32_t leaky, uint16_t* y, float* y_f32, void* stream);
/* The same with the GEMM columns split over `splits` workgroup slices (for
 * layers with few output pixels): each slice stores its partial sums in its own
 * part of `workspace` (splits * batch * out_height * out_width * out_channels
 * floats), then one pass adds the parts in order, with bias and activation, so
 * the result is the same on every run. */
int ldic_nn_conv_forward_split(const ldic_nn_conv* conv, const uint16_t* x, const uint16_t* w, const float* bias,
                               int32_t leaky, uint16_t* y, float* y_f32, float* workspace, int32_t splits,
                               void* stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the split-K
 * forward above with its partial sums in a library-owned static scratch of
 * 1,048,576 floats (4 MiB) per device instead of a workspace argument; it calls
 * no allocator.  splits * batch * out_height * out_width * out_channels must not
 * exceed that (LDIC_ERR_ARGUMENT otherwise).  All calls on one device share its
 * scratch, which every call overwrites before reading: calls that overlap on
 * one device race. */
int ldic_nn_conv_forward_split_static(const ldic_nn_conv* conv, const uint16_t* x, const uint16_t* w,
                                      const float* bias, int32_t leaky, uint16_t* y, float* y_f32, int32_t splits,
                                      void* stream);
/* Gradients from dy (float32, layout of the output, taken after the
 * activation's backward): dx (nullable; float32 with its own channel stride
 * and offset, overwritten or accumulated) needs the transposed weights; dw
 * and db (nullable) are accumulated. */
int ldic_nn_conv_backward(const ldic_nn_conv* conv, const uint16_t* x, const uint16_t* w_transposed, const float* dy,
                          float* dx, int32_t dx_stride, int32_t dx_offset, int32_t accumulate, float* dw, float* db,
                          void* stream);
/* w_transposed is (in_channels, KPT): column tap*out_channels + co holds
 * w[co][tap*in_channels + ci]; KPT = 9*out_channels rounded up to 16. */
int ldic_nn_transpose_weights(int32_t in_channels, int32_t out_channels, const uint16_t* w, uint16_t* w_transposed,
                              void* stream);
/* dy *= (y > 0 ? 1 : 0.1) over `count` elements of identical layout. */
int ldic_nn_leaky_backward(size_t count, const uint16_t* y, float* dy, void* stream);

/* Local correlation out[p][d] = scale * sum_c f1[p][c] f2[p+d][c] over the
 * (2 radius + 1)^2 offsets d (dy outer), zero outside the image. */
typedef struct ldic_nn_corr {
    int32_t batch, height, width, channels, radius;
    int32_t f1_stride, f1_offset, f2_stride, f2_offset, out_stride, out_offset;
    float scale;
} ldic_nn_corr;
int ldic_nn_corr_forward(const ldic_nn_corr* corr, const uint16_t* f1, const uint16_t* f2, uint16_t* out,
                         void* stream);
/* Accumulates the gradients of f1 and f2 (either nullable) from dc. */
int ldic_nn_corr_backward(const ldic_nn_corr* corr, const uint16_t* f1, const uint16_t* f2, const float* dc,
                          int32_t dc_stride, int32_t dc_offset, float* df1, int32_t df1_stride, int32_t df1_offset,
                          float* df2, int32_t df2_stride, int32_t df2_offset, void* stream);

/* Bilinear warp of f (dense NHWC) by a float32 flow (pixels of f), zero
 * outside; the backward accumulates into d_f only (the flow is a constant). */
int ldic_nn_warp_forward(int32_t batch, int32_t height, int32_t width, int32_t channels, const uint16_t* f,
                         const float* flow, uint16_t* out, void* stream);
int ldic_nn_warp_backward(int32_t batch, int32_t height, int32_t width, int32_t channels, const float* flow,
                          const float* d_out, float* d_f, void* stream);
/* fine(x, y) = 2 * bilinear(coarse)(x/2, y/2); optional bfloat16 copy into a
 * channel slice. */
int ldic_nn_upsample_flow(int32_t batch, int32_t coarse_height, int32_t coarse_width, int32_t fine_height,
                          int32_t fine_width, const float* coarse, float* fine, uint16_t* fine_bf16,
                          int32_t bf16_stride, int32_t bf16_offset, void* stream);
int ldic_nn_copy_slice(size_t pixels, int32_t channels, const uint16_t* src, int32_t src_stride, int32_t src_offset,
                       uint16_t* dst, int32_t dst_stride, int32_t dst_offset, void* stream);
int ldic_nn_add_slice(size_t pixels, int32_t channels, const float* src, int32_t src_stride, int32_t src_offset,
                      float* dst, int32_t dst_stride, int32_t dst_offset, void* stream);
/* uint8 -> bfloat16 as (v - 128) / 64; float32 -> bfloat16. */
int ldic_nn_images_to_bf16(size_t count, const uint8_t* src, uint16_t* dst, void* stream);
int ldic_nn_to_bf16(size_t count, const float* src, uint16_t* dst, void* stream);
/* Huber loss (delta 1) of `count` flow vectors: grad = weight/count *
 * clamp(flow - truth, -1, 1); the loss is added into *loss. */
int ldic_nn_flow_loss(int32_t count, const float* flow, const float* truth, float weight, float* grad, float* loss,
                      void* stream);
/* Adam with bias correction (step >= 1); param_bf16 (nullable) follows. */
int ldic_nn_adam(size_t count, float* param, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
                 float eps, int32_t step, uint16_t* param_bf16, void* stream);

/* Seeds of a subset grid from a flow field (flow pixel j at finest-level
 * coordinate j * scale): translation at the subset centre, optional
 * gradients from central differences; active = selected (nullable: all). */
typedef struct ldic_nn_seed {
    int32_t grid_width, grid_height, subset_size, subset_stride;
    int32_t flow_width, flow_height;
    float scale;
    int32_t state_stride, gradients;
    const uint8_t* selected;
} ldic_nn_seed;
int ldic_nn_seed_from_flow(const ldic_nn_seed* seed, const float* flow, float* state, uint32_t* active, void* stream);

int ldic_device_alloc(size_t bytes, void** out);
int ldic_device_free(void* pointer);
int ldic_copy_to_device(void* device_dst, const void* host_src, size_t bytes);
int ldic_copy_to_host(void* host_dst, const void* device_src, size_t bytes);
int ldic_device_synchronize(void);

/* Device-memory helpers of the learned seeder's host layer (no counterpart in the
 * reference; they replace the framework tensors it would otherwise need):
 * ldic_device_zero clears `bytes` bytes on `stream` with a fill kernel (so a
 * captured stream records a kernel node, not a memset node);
 * ldic_device_copy_async copies device to device on `stream`;
 * ldic_stream_synchronize waits for `stream` (NULL: the null stream). */
int ldic_device_zero(void* device_ptr, size_t bytes, void* stream);
int ldic_device_copy_async(void* device_dst, const void* device_src, size_t bytes, void* stream);
int ldic_stream_synchronize(void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LIVEDIC_H */
