/*
 * CPU oracle for the live-dic IC-GN hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the arithmetic of the reference WGSL kernels
 * that the native `live_dic` supervisor runs per target frame.  It is the checker
 * for the HIP library: only tests, the smoke check and the benchmark's CPU
 * baseline leg may load it.  The product never links or calls it.
 *
 * Every function names the reference lines it follows (paths relative to
 * DIC/LucasKanadeICGN/python/src/live_dic/).  Floating-point expressions keep
 * the WGSL evaluation order and are compiled without FMA contraction
 * (-ffp-contract=off) so that one rounding happens where WGSL has one operator.
 *
 * Summation order in the solver follows the portable WGSL kernel: a strided
 * per-lane accumulation over SOLVER_WORKGROUP_SIZE lanes followed by the
 * shared-memory tree with strides n/2 .. 1 (shaders/icgn.wgsl:711-799).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAX_LANES 128

/* shaders/icgn.wgsl:128-143 and shaders/full_pyramid.wgsl:150-158: single
 * reflection without edge repeat, followed by the halo clamp. */
static int mirror_clamp(int c, int size) {
    int r = c;
    if (c < 0) r = -c;
    else if (c >= size) r = 2 * size - 2 - c;
    if (r < 0) r = 0;
    if (r > size - 1) r = size - 1;
    return r;
}

/* ------------------------------------------------------------------------- */
/* Integer 5x5 pyramid reduction: shaders/full_pyramid.wgsl:215-249.           */
/* ------------------------------------------------------------------------- */
void ora_reduce(const uint8_t* src, int w, int h, uint8_t* dst) {
    static const uint32_t taps[5] = {1u, 4u, 6u, 4u, 1u};
    int dw = w / 2, dh = h / 2;
    for (int y = 0; y < dh; ++y) {
        for (int x = 0; x < dw; ++x) {
            uint32_t v = 0;
            for (int j = 0; j < 5; ++j) {
                int sy = mirror_clamp(2 * y + j - 2, h);
                for (int i = 0; i < 5; ++i) {
                    int sx = mirror_clamp(2 * x + i - 2, w);
                    v += taps[j] * taps[i] * (uint32_t)src[(size_t)sy * w + sx];
                }
            }
            dst[(size_t)y * dw + x] = (uint8_t)(v / 256u);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Target prefilter: shaders/icgn.wgsl:145-214 (25 central taps of the 31    */
/* stored FIR weights, index tap + 15).                                      */
/* ------------------------------------------------------------------------- */
void ora_prefilter(const uint8_t* src, int w, int h, const float* fir31, float* coeff, float* scratch) {
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            float r = 0.0f;
            for (int tap = -12; tap <= 12; ++tap) {
                int sx = mirror_clamp(x + tap, w);
                r = r + fir31[tap + 15] * (float)src[(size_t)y * w + sx];
            }
            scratch[(size_t)y * w + x] = r;
        }
    }
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            float r = 0.0f;
            for (int tap = -12; tap <= 12; ++tap) {
                int sy = mirror_clamp(y + tap, h);
                r = r + fir31[tap + 15] * scratch[(size_t)sy * w + x];
            }
            coeff[(size_t)y * w + x] = r;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* IC-GN solver: shaders/icgn.wgsl:216-958.                                  */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t width, height, grid_width, grid_height;
    int32_t subset_size, subset_stride;
    int32_t motion_model;      /* 0 affine, 1 translation, 2 quadratic */
    int32_t interpolation;     /* 0 B-spline coefficients, 1 bilinear, 2 Keys cubic */
    int32_t spatial_weighting; /* 0 uniform, 1 Gaussian */
    int32_t roi_enabled;
    float gaussian_sigma;
    int32_t iterations;
    float tolerance, zncc_min;
    int32_t lanes;             /* SOLVER_WORKGROUP_SIZE, 16 by default */
    int32_t preserve_disabled;
    int32_t divergence_patience;       /* DIVERGENCE_PATIENCE, 0 off (icgn.wgsl:100-111) */
    int32_t divergence_min_iterations; /* DIVERGENCE_MIN_ITERATIONS */
} ora_solve_cfg;

typedef struct {
    const ora_solve_cfg* c;
    const float* reference; /* vec4 per pixel: f, gx, gy, 0 */
    const float* metadata;  /* 16 or 28 floats per subset */
    const float* inverse;   /* 36 or 144 floats per subset */
    const float* coeff;     /* B-spline coefficient plane */
    const uint8_t* target;  /* raw target for the direct interpolants */
} solve_ctx;

/* icgn.wgsl:230-240 */
static void cubic_weights(float t, float w[4]) {
    float t2 = t * t;
    float t3 = t2 * t;
    float inv = 1.0f - t;
    w[0] = inv * inv * inv / 6.0f;
    w[1] = (3.0f * t3 - 6.0f * t2 + 4.0f) / 6.0f;
    w[2] = (-3.0f * t3 + 3.0f * t2 + 3.0f * t + 1.0f) / 6.0f;
    w[3] = t3 / 6.0f;
}

/* icgn.wgsl:283-291 */
static void keys_weights(float t, float w[4]) {
    float t2 = t * t;
    float t3 = t2 * t;
    w[0] = -0.5f * t + t2 - 0.5f * t3;
    w[1] = 1.0f - 2.5f * t2 + 1.5f * t3;
    w[2] = 0.5f * t + 2.0f * t2 - 1.5f * t3;
    w[3] = -0.5f * t2 + 0.5f * t3;
}

/* Out-of-plane indices can only arise for divergent quadratic warps between
 * the nine checked points; clamp them into the plane. */
static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static float coeff_at(const solve_ctx* s, int x, int y) {
    int w = s->c->width, h = s->c->height;
    return s->coeff[(size_t)clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)];
}

static float source_at(const solve_ctx* s, int x, int y) {
    int w = s->c->width, h = s->c->height;
    return (float)s->target[(size_t)clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)];
}

static float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }

/* icgn.wgsl:242-309 */
static float sample_target(const solve_ctx* s, float px, float py) {
    float fx = floorf(px), fy = floorf(py);
    int bx = (int)fx, by = (int)fy;
    float tx = px - fx, ty = py - fy;
    if (s->c->interpolation == 1) {
        float top = mixf(source_at(s, bx, by), source_at(s, bx + 1, by), tx);
        float bottom = mixf(source_at(s, bx, by + 1), source_at(s, bx + 1, by + 1), tx);
        return mixf(top, bottom, ty);
    }
    float wx[4], wy[4];
    if (s->c->interpolation == 2) {
        keys_weights(tx, wx);
        keys_weights(ty, wy);
    } else {
        cubic_weights(tx, wx);
        cubic_weights(ty, wy);
    }
    float value = 0.0f;
    for (int yy = 0; yy < 4; ++yy) {
        float row = 0.0f;
        for (int xx = 0; xx < 4; ++xx) {
            float v = s->c->interpolation == 2 ? source_at(s, bx + xx - 1, by + yy - 1)
                                               : coeff_at(s, bx + xx - 1, by + yy - 1);
            row = row + wx[xx] * v;
        }
        value = value + wy[yy] * row;
    }
    return value;
}

static float metadata_value(const solve_ctx* s, int subset, int component) {
    int stride = s->c->motion_model == 2 ? 28 : 16;
    return s->metadata[(size_t)subset * stride + component];
}

static float hessian_value(const solve_ctx* s, int subset, int component) {
    int stride = s->c->motion_model == 2 ? 144 : 36;
    return s->inverse[(size_t)subset * stride + component];
}

/* icgn.wgsl:1607-1629 */
static float half_extent(const ora_solve_cfg* c) { return 0.5f * ((float)c->subset_size - 1.0f); }

static float spatial_weight(const ora_solve_cfg* c, float lx, float ly) {
    if (c->spatial_weighting == 1) {
        float inverse_variance = 1.0f / (c->gaussian_sigma * c->gaussian_sigma);
        return expf(-0.5f * (lx * lx + ly * ly) * inverse_variance);
    }
    return 1.0f;
}

static float integration_weight_sum(const solve_ctx* s, int subset, int partial) {
    if (s->c->spatial_weighting == 1 || partial) {
        int mean_offset = s->c->motion_model == 2 ? 26 : 14;
        return metadata_value(s, subset, mean_offset + 1);
    }
    return (float)(s->c->subset_size * s->c->subset_size);
}

static float length2(float a, float b) { return sqrtf(a * a + b * b); }

/* icgn.wgsl:515-549 */
static void compose_standard(int model, float T[2], float L[4], const float d[6], float det) {
    if (model == 1) {
        T[0] = T[0] - d[0];
        T[1] = T[1] - d[3];
        L[0] = L[1] = L[2] = L[3] = 0.0f;
        return;
    }
    float i11 = (1.0f + d[5]) / det;
    float i12 = -d[2] / det;
    float i21 = -d[4] / det;
    float i22 = (1.0f + d[1]) / det;
    float w11 = 1.0f + L[0], w12 = L[1], w21 = L[2], w22 = 1.0f + L[3];
    float itx = -(i11 * d[0] + i12 * d[3]);
    float ity = -(i21 * d[0] + i22 * d[3]);
    float ntx = T[0] + (w11 * itx + w12 * ity);
    float nty = T[1] + (w21 * itx + w22 * ity);
    float u11 = w11 * i11 + w12 * i21;
    float u12 = w11 * i12 + w12 * i22;
    float u21 = w21 * i11 + w22 * i21;
    float u22 = w21 * i12 + w22 * i22;
    T[0] = ntx;
    T[1] = nty;
    L[0] = u11 - 1.0f;
    L[1] = u12;
    L[2] = u21;
    L[3] = u22 - 1.0f;
}

/* icgn.wgsl:393-435 */
static void quadratic_warp_matrix(const float p[12], float m[6][6]) {
    float u = p[0], ux = p[1], uy = p[2], v = p[3], vx = p[4], vy = p[5];
    float uxx = p[6], uxy = p[7], uyy = p[8], vxx = p[9], vxy = p[10], vyy = p[11];
    m[0][0] = 1.0f + 2.0f * ux + ux * ux + u * uxx;
    m[0][1] = 2.0f * u * uxy + 2.0f * (1.0f + ux) * uy;
    m[0][2] = uy * uy + u * uyy;
    m[0][3] = 2.0f * u * (1.0f + ux);
    m[0][4] = 2.0f * u * uy;
    m[0][5] = u * u;
    m[1][0] = 0.5f * (v * uxx + 2.0f * (1.0f + ux) * vx + u * vxx);
    m[1][1] = 1.0f + uy * vx + ux * vy + v * uxy + u * vxy + vy + ux;
    m[1][2] = 0.5f * (v * uyy + 2.0f * uy * (1.0f + vy) + u * vyy);
    m[1][3] = v + v * ux + u * vx;
    m[1][4] = u + v * uy + u * vy;
    m[1][5] = u * v;
    m[2][0] = vx * vx + v * vxx;
    m[2][1] = 2.0f * v * vxy + 2.0f * vx * (1.0f + vy);
    m[2][2] = 1.0f + 2.0f * vy + vy * vy + v * vyy;
    m[2][3] = 2.0f * v * vx;
    m[2][4] = 2.0f * v * (1.0f + vy);
    m[2][5] = v * v;
    m[3][0] = 0.5f * uxx; m[3][1] = uxy; m[3][2] = 0.5f * uyy; m[3][3] = 1.0f + ux; m[3][4] = uy; m[3][5] = u;
    m[4][0] = 0.5f * vxx; m[4][1] = vxy; m[4][2] = 0.5f * vyy; m[4][3] = vx; m[4][4] = 1.0f + vy; m[4][5] = v;
    m[5][0] = 0.0f; m[5][1] = 0.0f; m[5][2] = 0.0f; m[5][3] = 0.0f; m[5][4] = 0.0f; m[5][5] = 1.0f;
}

/* icgn.wgsl:437-513; returns 0 when a pivot is below 1e-8. */
static int compose_quadratic(float T[2], float L[4], float C[6], const float d[12]) {
    float inc[6][6];
    float a[6][12];
    quadratic_warp_matrix(d, inc);
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            a[r][c] = inc[r][c];
            a[r][6 + c] = r == c ? 1.0f : 0.0f;
        }
    for (int col = 0; col < 6; ++col) {
        int pr = col;
        float ps = fabsf(a[col][col]);
        for (int cand = col + 1; cand < 6; ++cand) {
            float cs = fabsf(a[cand][col]);
            if (cs > ps) { pr = cand; ps = cs; }
        }
        if (ps < 1e-8f) return 0;
        if (pr != col) {
            float tmp[12];
            memcpy(tmp, a[col], sizeof tmp);
            memcpy(a[col], a[pr], sizeof tmp);
            memcpy(a[pr], tmp, sizeof tmp);
        }
        float pivot = a[col][col];
        for (int e = 0; e < 12; ++e) a[col][e] /= pivot;
        for (int r = 0; r < 6; ++r) {
            if (r != col) {
                float f = a[r][col];
                for (int e = 0; e < 12; ++e) a[r][e] -= f * a[col][e];
            }
        }
    }
    float cu[6] = {0.5f * C[0], C[1], 0.5f * C[2], 1.0f + L[0], L[1], T[0]};
    float cv[6] = {0.5f * C[3], C[4], 0.5f * C[5], L[2], 1.0f + L[3], T[1]};
    float uu[6], uv[6];
    for (int col = 0; col < 6; ++col) {
        uu[col] = 0.0f;
        uv[col] = 0.0f;
        for (int e = 0; e < 6; ++e) {
            uu[col] += cu[e] * a[e][6 + col];
            uv[col] += cv[e] * a[e][6 + col];
        }
    }
    C[0] = 2.0f * uu[0]; C[1] = uu[1]; C[2] = 2.0f * uu[2];
    C[3] = 2.0f * uv[0]; C[4] = uv[1]; C[5] = 2.0f * uv[2];
    T[0] = uu[5];
    T[1] = uv[5];
    L[0] = uu[3] - 1.0f;
    L[1] = uu[4];
    L[2] = uv[3];
    L[3] = uv[4] - 1.0f;
    return 1;
}

/* icgn.wgsl:551-586 */
static float standard_change(const ora_solve_cfg* c, const float d[6]) {
    if (c->motion_model == 1) return length2(d[0], d[3]);
    float m = 0.0f;
    float he = half_extent(c);
    for (int corner = 0; corner < 4; ++corner) {
        float x = (corner == 1 || corner == 3) ? he : -he;
        float y = corner >= 2 ? he : -he;
        float ch = length2(d[0] + d[1] * x + d[2] * y, d[3] + d[4] * x + d[5] * y);
        m = fmaxf(m, ch);
    }
    return m;
}

static float quadratic_change(const ora_solve_cfg* c, const float d[12]) {
    float m = 0.0f;
    float he = half_extent(c);
    for (int p = 0; p < 9; ++p) {
        float x = (float)(p % 3 - 1) * he;
        float y = (float)(p / 3 - 1) * he;
        float ch = length2(d[0] + d[1] * x + d[2] * y + 0.5f * d[6] * x * x + d[7] * x * y + 0.5f * d[8] * y * y,
                           d[3] + d[4] * x + d[5] * y + 0.5f * d[9] * x * x + d[10] * x * y + 0.5f * d[11] * y * y);
        m = fmaxf(m, ch);
    }
    return m;
}

/* divergence_status, icgn.wgsl:103-111. */
static float divergence_status(const ora_solve_cfg* c, float change, int count, float status, float* previous,
                               int* stalled) {
    if (c->divergence_patience == 0 || status != 0.0f) return status;
    if (count >= c->divergence_min_iterations) *stalled = change >= *previous ? *stalled + 1 : 0;
    *previous = change;
    return *stalled >= c->divergence_patience ? 3.0f : status;
}

static float convergence_status(const ora_solve_cfg* c, float change, float zncc) {
    if (change >= c->tolerance) return 0.0f;
    return zncc >= c->zncc_min ? 1.0f : -2.0f;
}

/* icgn.wgsl:595-958 for one subset. */
/* support == NULL: the full window.  Otherwise a partial-support subset
 * (solve_partial*, icgn.wgsl:1040-1085): only support pixels (row-major bits,
 * pixel 0 in bit 0 of word 0) enter the sums, the subset must carry metadata
 * validity -4 (icgn.wgsl:611-616), rejection is -1, its own mean and weight
 * sum are used (icgn.wgsl:630-636) and the support box (xmin, ymin, xmax,
 * ymax, local pixels) bounds the warp test (icgn.wgsl:638-642). */
static void solve_subset(const solve_ctx* s, int subset, const float* seed, int enabled, float* out,
                         const uint32_t* support, const float* box) {
    const ora_solve_cfg* c = s->c;
    const int model = c->motion_model;
    const int S = c->subset_size, D = c->subset_stride;
    const int lanes = c->lanes;
    float T[2] = {0.0f, 0.0f}, L[4] = {0.0f, 0.0f, 0.0f, 0.0f}, C[6] = {0, 0, 0, 0, 0, 0};
    if (seed) {
        T[0] = seed[0];
        T[1] = seed[3];
        if (model != 1) {
            L[0] = seed[1]; L[1] = seed[2]; L[2] = seed[4]; L[3] = seed[5];
        }
        if (model == 2) {
            C[0] = seed[6]; C[1] = seed[7]; C[2] = seed[8]; C[3] = seed[9]; C[4] = seed[10]; C[5] = seed[11];
        }
    }
    const int partial = support != NULL;
    int reference_valid = partial ? metadata_value(s, subset, 1) == -4.0f : metadata_value(s, subset, 1) >= 0.5f;
    int valid = enabled && reference_valid;
    float rejected = -1.0f;
    if (c->roi_enabled && !partial) {
        float m1 = metadata_value(s, subset, 1);
        rejected = fminf(-1.0f, m1);
    }
    float status = valid ? 0.0f : rejected;
    float iters = 0.0f;

    const int gx = subset % c->grid_width, gy = subset / c->grid_width;
    const float cx = 0.5f * (float)S + (float)gx * (float)D;
    const float cy = 0.5f * (float)S + (float)gy * (float)D;
    const float he = half_extent(c);
    const int mean_offset = model == 2 ? 26 : 14;
    const float ref_mean = (D != S || partial) ? metadata_value(s, subset, mean_offset) : 0.0f;
    const float W = (float)c->width, H = (float)c->height;
    float bx0 = -he, by0 = -he, bx1 = he, by1 = he;
    if (partial) {
        bx0 = box[0] - he;
        by0 = box[1] - he;
        bx1 = box[2] - he;
        by1 = box[3] - he;
    }

    float previous_change = 3.402823e38f;
    int stalled = 0;
    const int patience = c->divergence_patience;
    for (int it = 0; it < 30; ++it) {
        int inside = 1;
        if (patience && status == 3.0f) {
            /* icgn.wgsl:647: a stopped solve skips the bounds test. */
        } else if (model == 1) {
            float minx = cx - 0.5f + bx0 + T[0], miny = cy - 0.5f + by0 + T[1];
            float maxx = cx - 0.5f + bx1 + T[0], maxy = cy - 0.5f + by1 + T[1];
            inside = minx >= 2.0f && miny >= 2.0f && maxx <= W - 3.0f && maxy <= H - 3.0f;
        } else {
            int count = model == 2 ? 9 : 4;
            for (int p = 0; p < count; ++p) {
                float lx = (p == 1 || p == 3) ? bx1 : bx0;
                float ly = p >= 2 ? by1 : by0;
                if (model == 2) {
                    lx = (float)(p % 3 - 1) * he;
                    ly = (float)(p / 3 - 1) * he;
                }
                float dx = T[0] + (L[0] * lx + L[1] * ly);
                float dy = T[1] + (L[2] * lx + L[3] * ly);
                if (model == 2) {
                    dx += 0.5f * C[0] * lx * lx + C[1] * lx * ly + 0.5f * C[2] * ly * ly;
                    dy += 0.5f * C[3] * lx * lx + C[4] * lx * ly + 0.5f * C[5] * ly * ly;
                }
                float px = cx - 0.5f + lx + dx;
                float py = cy - 0.5f + ly + dy;
                inside = inside && px >= 2.0f && py >= 2.0f && px <= W - 3.0f && py <= H - 3.0f;
            }
        }
        if (!inside) {
            if (!c->roi_enabled || status > -2.5f) status = -1.0f;
        }
        if (it >= c->iterations || fabsf(status) > 0.5f) break;

        /* Per-lane partial sums, vec4 a/b/c and the quadratic pair. */
        float pa[MAX_LANES][4], pb[MAX_LANES][4], pc[MAX_LANES][4];
        float qa[MAX_LANES][4], qb[MAX_LANES][4];
        for (int lane = 0; lane < lanes; ++lane) {
            float sc = 0.0f, sc2 = 0.0f, src = 0.0f;
            float sj[6] = {0, 0, 0, 0, 0, 0};
            float sq[6] = {0, 0, 0, 0, 0, 0};
            for (int pix = lane; pix < S * S; pix += lanes) {
                int col = pix % S, row = pix / S;
                float lx = (float)col - he, ly = (float)row - he;
                float dx = T[0], dy = T[1];
                if (model != 1) {
                    dx = dx + (L[0] * lx + L[1] * ly);
                    dy = dy + (L[2] * lx + L[3] * ly);
                    if (model == 2) {
                        dx += 0.5f * C[0] * lx * lx + C[1] * lx * ly + 0.5f * C[2] * ly * ly;
                        dy += 0.5f * C[3] * lx * lx + C[4] * lx * ly + 0.5f * C[5] * ly * ly;
                    }
                }
                float cur = sample_target(s, cx - 0.5f + lx + dx, cy - 0.5f + ly + dy);
                int pxi = gx * D + col, pyi = gy * D + row;
                const float* rv = s->reference + ((size_t)pyi * c->width + pxi) * 4;
                float f = rv[0] - ref_mean;
                float gxv = rv[1], gyv = rv[2];
                float iw = spatial_weight(c, lx, ly);
                if (partial && !((support[pix >> 5] >> (pix & 31)) & 1u)) iw = 0.0f;
                float wc = iw * cur;
                float gxg = gxv * wc, gyg = gyv * wc;
                sc = sc + wc;
                sc2 = sc2 + wc * cur;
                src = src + iw * f * cur;
                sj[0] = sj[0] + gxg;
                sj[3] = sj[3] + gyg;
                if (model != 1) {
                    sj[1] = sj[1] + gxg * lx;
                    sj[2] = sj[2] + gxg * ly;
                    sj[4] = sj[4] + gyg * lx;
                    sj[5] = sj[5] + gyg * ly;
                    if (model == 2) {
                        float hx2 = 0.5f * lx * lx, xy = lx * ly, hy2 = 0.5f * ly * ly;
                        sq[0] += gxg * hx2; sq[1] += gxg * xy; sq[2] += gxg * hy2;
                        sq[3] += gyg * hx2; sq[4] += gyg * xy; sq[5] += gyg * hy2;
                    }
                }
            }
            pa[lane][0] = sc; pa[lane][1] = sc2; pa[lane][2] = src; pa[lane][3] = sj[0];
            pb[lane][0] = sj[1]; pb[lane][1] = sj[2]; pb[lane][2] = sj[3]; pb[lane][3] = sj[4];
            pc[lane][0] = sj[5]; pc[lane][1] = 0.0f; pc[lane][2] = 0.0f; pc[lane][3] = 0.0f;
            qa[lane][0] = sq[0]; qa[lane][1] = sq[1]; qa[lane][2] = sq[2]; qa[lane][3] = sq[3];
            qb[lane][0] = sq[4]; qb[lane][1] = sq[5]; qb[lane][2] = 0.0f; qb[lane][3] = 0.0f;
        }
        for (int stride = lanes / 2; stride >= 1; stride /= 2) {
            for (int lane = 0; lane < stride; ++lane) {
                for (int k = 0; k < 4; ++k) {
                    pa[lane][k] = pa[lane][k] + pa[lane + stride][k];
                    pb[lane][k] = pb[lane][k] + pb[lane + stride][k];
                    pc[lane][k] = pc[lane][k] + pc[lane + stride][k];
                    qa[lane][k] = qa[lane][k] + qa[lane + stride][k];
                    qb[lane][k] = qb[lane][k] + qb[lane + stride][k];
                }
            }
        }
        float mean = pa[0][0] / integration_weight_sum(s, subset, partial);
        float norm = sqrtf(fmaxf(0.0f, pa[0][1] - pa[0][0] * mean));
        float ref_norm = metadata_value(s, subset, 0);
        iters = (float)(it + 1);
        if (norm < 1e-9f || ref_norm < 1e-9f) {
            status = -2.0f;
            continue;
        }
        float ratio = ref_norm / norm;
        float zncc = pa[0][2] / (ref_norm * norm);
        if (model == 2) {
            float rj[12] = {pa[0][3], pb[0][0], pb[0][1], pb[0][2], pb[0][3], pc[0][0],
                            qa[0][0], qa[0][1], qa[0][2], qa[0][3], qb[0][0], qb[0][1]};
            float res[12], d[12];
            for (int k = 0; k < 12; ++k)
                res[k] = metadata_value(s, subset, 2 + k) - ratio * (rj[k] - mean * metadata_value(s, subset, 14 + k));
            for (int r = 0; r < 12; ++r) {
                float v = 0.0f;
                for (int k = 0; k < 12; ++k) v += hessian_value(s, subset, 12 * r + k) * res[k];
                d[r] = -v;
            }
            if (!compose_quadratic(T, L, C, d)) {
                status = -1.0f;
            } else {
                float ch = quadratic_change(c, d);
                status = divergence_status(c, ch, it + 1, convergence_status(c, ch, zncc), &previous_change,
                                           &stalled);
            }
        } else {
            float rj[6] = {pa[0][3], pb[0][0], pb[0][1], pb[0][2], pb[0][3], pc[0][0]};
            float d[6] = {0, 0, 0, 0, 0, 0};
            if (model == 1) {
                float rx = metadata_value(s, subset, 2) - ratio * (rj[0] - mean * metadata_value(s, subset, 8));
                float ry = metadata_value(s, subset, 5) - ratio * (rj[3] - mean * metadata_value(s, subset, 11));
                d[0] = -(hessian_value(s, subset, 0) * rx + hessian_value(s, subset, 3) * ry);
                d[3] = -(hessian_value(s, subset, 18) * rx + hessian_value(s, subset, 21) * ry);
            } else {
                float res[6];
                for (int k = 0; k < 6; ++k)
                    res[k] = metadata_value(s, subset, 2 + k) - ratio * (rj[k] - mean * metadata_value(s, subset, 8 + k));
                for (int r = 0; r < 6; ++r) {
                    float v = 0.0f;
                    for (int k = 0; k < 6; ++k) v += hessian_value(s, subset, 6 * r + k) * res[k];
                    d[r] = -v;
                }
            }
            float det = model == 1 ? 1.0f : (1.0f + d[1]) * (1.0f + d[5]) - d[2] * d[4];
            if (fabsf(det) < 1e-8f) {
                status = -1.0f;
            } else {
                compose_standard(model, T, L, d, det);
                float ch = standard_change(c, d);
                status = divergence_status(c, ch, it + 1, convergence_status(c, ch, zncc), &previous_change,
                                           &stalled);
            }
        }
    }
    if (patience && status == 3.0f) status = 0.0f; /* icgn.wgsl:928-929 */
    if (model == 2) {
        float v[16] = {T[0], L[0], L[1], T[1], L[2], L[3], C[0], C[1], C[2], C[3], C[4], C[5], status, iters, 0.0f, 0.0f};
        memcpy(out, v, sizeof v);
    } else {
        float v[8] = {T[0], L[0], L[1], T[1], L[2], L[3], status, iters};
        memcpy(out, v, sizeof v);
    }
}

/* Solve every subset: `solve` (seed == NULL), `solve_seeded` and
 * `retry_seeded` (icgn.wgsl:960-1017).  State layout: icgn.wgsl:928-957. */
int ora_solve(const ora_solve_cfg* c, const float* reference, const float* metadata, const float* inverse,
              const float* coeff, const uint8_t* target, const float* seed_state, const uint32_t* active,
              float* state) {
    if (c->lanes < 1 || c->lanes > MAX_LANES || (c->lanes & (c->lanes - 1))) return -1;
    solve_ctx s = {c, reference, metadata, inverse, coeff, target};
    const int n = c->grid_width * c->grid_height;
    const int stride = c->motion_model == 2 ? 16 : 8;
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) {
        const float* seed = seed_state ? seed_state + (size_t)i * stride : NULL;
        int enabled = active ? active[i] != 0u : 1;
        if (!enabled && c->preserve_disabled) continue;
        /* The state slots of seeds are read as (tx, ux, uy, ty, vx, vy, [curvature]). */
        float seed_params[12];
        if (seed) {
            if (c->motion_model == 2) {
                const float* q = seed;
                float tmp[12] = {q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9], q[10], q[11]};
                memcpy(seed_params, tmp, sizeof tmp);
            } else {
                for (int k = 0; k < 6; ++k) seed_params[k] = seed[k];
            }
        }
        solve_subset(&s, i, seed ? seed_params : NULL, enabled, state + (size_t)i * stride, NULL, NULL);
    }
    return 0;
}

/* The partial pass (solve_partial / solve_partial_seeded / retry_partial_seeded,
 * icgn.wgsl:1063-1085) over `count` listed subsets, run after ora_solve on the
 * same state: support words ceil(S*S/32) per entry, boxes 4 floats per entry. */
int ora_solve_partial(const ora_solve_cfg* c, const float* reference, const float* metadata, const float* inverse,
                      const float* coeff, const uint8_t* target, const float* seed_state, const uint32_t* active,
                      int count, const uint32_t* indices, const uint32_t* support_words, const float* boxes,
                      float* state) {
    if (c->lanes < 1 || c->lanes > MAX_LANES || (c->lanes & (c->lanes - 1))) return -1;
    if (c->motion_model == 2) return -1;
    solve_ctx s = {c, reference, metadata, inverse, coeff, target};
    const int n = c->grid_width * c->grid_height;
    const int words = (c->subset_size * c->subset_size + 31) / 32;
    for (int k = 0; k < count; ++k)
        if ((int)indices[k] >= n) return -1;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < count; ++k) {
        const int i = (int)indices[k];
        const float* seed = seed_state ? seed_state + (size_t)i * 8 : NULL;
        int enabled = active ? active[i] != 0u : 1;
        if (!enabled && c->preserve_disabled) continue;
        solve_subset(&s, i, seed, enabled, state + (size_t)i * 8, support_words + (size_t)k * words, boxes + 4 * k);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Coarse-to-fine seed transfer: shaders/full_pyramid.wgsl:163-433.          */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t source_grid_width, source_grid_height;
    int32_t destination_grid_width, destination_grid_count;
    float source_scale, destination_scale;
    float subset_size, subset_stride;
    int32_t motion_model;
    int32_t track_seed_availability;
} ora_seed_cfg;

static void accumulate_seed(const ora_seed_cfg* c, const float* src_state, int source_index, float weight,
                            float spx, float spy, float dpx, float dpy, float ts, float acc[12], float* wsum) {
    int q = c->motion_model == 2;
    int stride = q ? 16 : 8;
    const float* s = src_state + (size_t)source_index * stride;
    float a[4] = {s[0], s[1], s[2], s[3]};
    float b[4] = {s[4], s[5], s[6], s[7]};
    float cc[4] = {0, 0, 0, 0};
    float status = b[2];
    if (q) {
        cc[0] = s[8]; cc[1] = s[9]; cc[2] = s[10]; cc[3] = s[11];
        status = s[12];
    }
    if (status < -0.5f) return;
    float dx = dpx - spx, dy = dpy - spy;
    float tu = a[0] + a[1] * dx + a[2] * dy;
    float tv = a[3] + b[0] * dx + b[1] * dy;
    float ux = a[1], uy = a[2], vx = b[0], vy = b[1];
    if (q) {
        tu += 0.5f * b[2] * dx * dx + b[3] * dx * dy + 0.5f * cc[0] * dy * dy;
        tv += 0.5f * cc[1] * dx * dx + cc[2] * dx * dy + 0.5f * cc[3] * dy * dy;
        ux += b[2] * dx + b[3] * dy;
        uy += b[3] * dx + cc[0] * dy;
        vx += cc[1] * dx + cc[2] * dy;
        vy += cc[2] * dx + cc[3] * dy;
    }
    acc[0] += weight * tu * ts;
    acc[1] += weight * ux;
    acc[2] += weight * uy;
    acc[3] += weight * tv * ts;
    acc[4] += weight * vx;
    acc[5] += weight * vy;
    if (q) {
        float cs = 1.0f / ts;
        acc[6] += weight * b[2] * cs;
        acc[7] += weight * b[3] * cs;
        for (int k = 0; k < 4; ++k) acc[8 + k] += weight * cc[k] * cs;
    }
    *wsum += weight;
}

void ora_transfer_seed(const ora_seed_cfg* c, const float* src_state, const uint32_t* permissions,
                       float* dst_seed, uint32_t* dst_active, float* dst_weight /* nullable */) {
    int q = c->motion_model == 2;
    int stride = q ? 16 : 8;
    for (int di = 0; di < c->destination_grid_count; ++di) {
        uint32_t perm = permissions ? permissions[di] : 31u;
        int dxg = di % c->destination_grid_width, dyg = di / c->destination_grid_width;
        float so = (1.0f - c->source_scale) * 0.5f;
        float dof = (1.0f - c->destination_scale) * 0.5f;
        float dcx = 0.5f * c->subset_size + c->subset_stride * (float)dxg;
        float dcy = 0.5f * c->subset_size + c->subset_stride * (float)dyg;
        float scx = (c->destination_scale * dcx + (dof - so)) / c->source_scale;
        float scy = (c->destination_scale * dcy + (dof - so)) / c->source_scale;
        float sgx = (scx - 0.5f * c->subset_size) / c->subset_stride;
        float sgy = (scy - 0.5f * c->subset_size) / c->subset_stride;
        int mlx = c->source_grid_width - 2 > 0 ? c->source_grid_width - 2 : 0;
        int mly = c->source_grid_height - 2 > 0 ? c->source_grid_height - 2 : 0;
        int lx = clampi((int)floorf(sgx), 0, mlx);
        int ly = clampi((int)floorf(sgy), 0, mly);
        float fx = fminf(fmaxf(sgx - (float)lx, 0.0f), 1.0f);
        float fy = fminf(fmaxf(sgy - (float)ly, 0.0f), 1.0f);
        float ts = c->source_scale / c->destination_scale;
        float acc[12] = {0};
        float wsum = 0.0f;
        for (int corner = 0; corner < 4; ++corner) {
            if ((perm & (1u << corner)) == 0u) continue;
            int ox = corner & 1, oy = corner >> 1;
            int sx = lx + ox, sy = ly + oy;
            if (sx >= c->source_grid_width || sy >= c->source_grid_height) continue;
            int si = sy * c->source_grid_width + sx;
            float wx = ox == 1 ? fx : 1.0f - fx;
            float wy = oy == 1 ? fy : 1.0f - fy;
            float spx = 0.5f * c->subset_size + c->subset_stride * (float)sx;
            float spy = 0.5f * c->subset_size + c->subset_stride * (float)sy;
            accumulate_seed(c, src_state, si, wx * wy, spx, spy, scx, scy, ts, acc, &wsum);
        }
        uint32_t fallback = perm >> 5u;
        if (wsum == 0.0f && fallback != 0u) {
            int si = (int)fallback - 1;
            float st = q ? src_state[(size_t)si * stride + 12] : src_state[(size_t)si * stride + 6];
            if (st >= 0.5f) {
                int sx = si % c->source_grid_width, sy = si / c->source_grid_width;
                float spx = 0.5f * c->subset_size + c->subset_stride * (float)sx;
                float spy = 0.5f * c->subset_size + c->subset_stride * (float)sy;
                accumulate_seed(c, src_state, si, 1.0f, spx, spy, scx, scy, ts, acc, &wsum);
            }
        }
        if (wsum > 0.0f)
            for (int k = 0; k < 12; ++k) acc[k] /= wsum;
        float* o = dst_seed + (size_t)di * stride;
        for (int k = 0; k < stride; ++k) o[k] = 0.0f;
        if (q) {
            for (int k = 0; k < 12; ++k) o[k] = acc[k];
        } else {
            for (int k = 0; k < 6; ++k) o[k] = acc[k];
        }
        if (c->track_seed_availability) {
            float marker = wsum >= 0.99999f ? 0.0f : -1.0f;
            o[q ? 12 : 6] = marker;
        }
        dst_active[di] = (perm >> 4u) & 1u;
        if (dst_weight) dst_weight[di] = wsum;
    }
}

/* ------------------------------------------------------------------------- */
/* UOD / median post-pass: shaders/icgn.wgsl:1087-1605.                      */
/* ------------------------------------------------------------------------- */
typedef struct {
    int32_t grid_width, grid_count;
    int32_t subset_stride;
    int32_t motion_model;
    int32_t outlier_policy;   /* 0 flag, 1 replace, 3 median */
    int32_t median_filter_size;
    int32_t median_filter_scope; /* 0 rejected, 1 all */
    int32_t roi_enabled;
    int32_t median_seed_available;
} ora_median_cfg;

/* icgn.wgsl:1087-1101 (bubble passes, then the middle element or mean of two). */
static float median_nine(float* v, int count) {
    for (int ph = 0; ph < 9; ++ph)
        for (int i = 0; i < 8; ++i)
            if (i + 1 < count && v[i] > v[i + 1]) {
                float t = v[i]; v[i] = v[i + 1]; v[i + 1] = t;
            }
    if (count % 2 == 1) return v[count / 2];
    return 0.5f * (v[count / 2 - 1] + v[count / 2]);
}

static int cmp_float(const void* a, const void* b) {
    float x = *(const float*)a, y = *(const float*)b;
    return (x > y) - (x < y);
}

/* icgn.wgsl:1103-1147: the quickselect returns the order statistics. */
static float median_up_to_81(float* v, int count) {
    qsort(v, (size_t)count, sizeof(float), cmp_float);
    if (count % 2 == 1) return v[count / 2];
    return 0.5f * (v[count / 2 - 1] + v[count / 2]);
}

static float recentered_component(const float* st, int model, int comp, float ox, float oy) {
    const float* a = st;
    const float* b = st + 4;
    if (model == 2) {
        const float* cc = st + 8;
        switch (comp) {
            case 0: return a[0] + a[1] * ox + a[2] * oy + 0.5f * b[2] * ox * ox + b[3] * ox * oy + 0.5f * cc[0] * oy * oy;
            case 1: return a[1] + b[2] * ox + b[3] * oy;
            case 2: return a[2] + b[3] * ox + cc[0] * oy;
            case 3: return a[3] + b[0] * ox + b[1] * oy + 0.5f * cc[1] * ox * ox + cc[2] * ox * oy + 0.5f * cc[3] * oy * oy;
            case 4: return b[0] + cc[1] * ox + cc[2] * oy;
            case 5: return b[1] + cc[2] * ox + cc[3] * oy;
            case 6: return b[2];
            case 7: return b[3];
            case 8: return cc[0];
            case 9: return cc[1];
            case 10: return cc[2];
            default: return cc[3];
        }
    }
    switch (comp) {
        case 0: return a[0] + a[1] * ox + a[2] * oy;
        case 1: return a[1];
        case 2: return a[2];
        case 3: return a[3] + b[0] * ox + b[1] * oy;
        case 4: return b[0];
        default: return b[1];
    }
}

static float status_of(const float* st, int model) { return model == 2 ? st[12] : st[6]; }

/* icgn.wgsl:1150-1323 */
static void median_motion_state(const ora_median_cfg* c, const float* state, const float* metadata,
                                const float* roi_words, int i, float* filtered) {
    int model = c->motion_model;
    int stride = model == 2 ? 16 : 8;
    const float* own = state + (size_t)i * stride;
    float* out = filtered + (size_t)i * stride;
    memcpy(out, own, sizeof(float) * stride);
    float own_status = status_of(own, model);
    int mstride = model == 2 ? 28 : 16;
    float reference_status = metadata[(size_t)i * mstride + 1];
    if ((own_status >= 0.5f && !c->median_filter_scope) || (reference_status < 0.5f && reference_status != -4.0f))
        return;
    int px = i % c->grid_width, py = i / c->grid_width;
    int gh = c->grid_count / c->grid_width;
    int radius = c->median_filter_size / 2;
    static const int order[12] = {0, 3, 1, 2, 4, 5, 6, 7, 8, 9, 10, 11};
    int active_count = model == 1 ? 2 : (model == 2 ? 12 : 6);
    float f[12];
    for (int k = 0; k < 12; ++k) f[k] = k < stride ? own[k] : 0.0f;
    if (model != 2) { /* own_c is zero for the 8-float layout */
        for (int k = 8; k < 12; ++k) f[k] = 0.0f;
    }
    const float* perm = roi_words ? roi_words + (size_t)i * 4 : NULL;
    for (int ai = 0; ai < active_count; ++ai) {
        int comp = order[ai];
        float values[81];
        int count = 0;
        if (own_status >= 0.5f) values[count++] = recentered_component(own, model, comp, 0.0f, 0.0f);
        for (int dy = -radius; dy <= radius; ++dy) {
            for (int dx = -radius; dx <= radius; ++dx) {
                if (dx == 0 && dy == 0) continue;
                int nx = px + dx, ny = py + dy;
                if (nx < 0 || ny < 0 || nx >= c->grid_width || ny >= gh) continue;
                if (c->roi_enabled) {
                    uint32_t bit = (uint32_t)((dy + 1) * 3 + dx + 1);
                    uint32_t word = (uint32_t)perm[0];
                    if (c->median_filter_size > 3) {
                        bit = (uint32_t)((dy + 4) * 9 + dx + 4);
                        word = (uint32_t)perm[bit / 24u];
                    }
                    if ((word & (1u << (bit % 24u))) == 0u) continue;
                }
                int ni = ny * c->grid_width + nx;
                const float* ns = state + (size_t)ni * stride;
                if (status_of(ns, model) < 0.5f) continue;
                values[count++] = recentered_component(ns, model, comp, -(float)c->subset_stride * (float)dx,
                                                       -(float)c->subset_stride * (float)dy);
            }
        }
        int minimum = (c->roi_enabled && c->median_filter_size == 3 && own_status < 0.5f) ? 3 : 4;
        if (count < minimum) return;
        f[comp] = median_up_to_81(values, count);
    }
    for (int k = 0; k < 8; ++k) out[k] = f[k];
    if (model == 2) {
        for (int k = 8; k < 12; ++k) out[k] = f[k];
        out[12] = (c->median_seed_available && own_status < 0.5f) ? 0.0f : own[12];
        out[13] = own[13];
        out[14] = own[14];
        out[15] = own[15];
    } else {
        out[6] = (c->median_seed_available && own_status < 0.5f) ? 0.0f : own[6];
        out[7] = own[7];
    }
}

void ora_median(const ora_median_cfg* c, const float* state, const float* metadata, const float* roi_words,
                float* filtered) {
    int model = c->motion_model;
    int stride = model == 2 ? 16 : 8;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < c->grid_count; ++i) {
        if (c->outlier_policy == 3) {
            median_motion_state(c, state, metadata, roi_words, i, filtered);
            continue;
        }
        const float* own = state + (size_t)i * stride;
        float* out = filtered + (size_t)i * stride;
        memcpy(out, own, sizeof(float) * stride);
        if (status_of(own, model) < 0.5f) continue;
        int px = i % c->grid_width, py = i / c->grid_width;
        int gh = c->grid_count / c->grid_width;
        int ncomp = model == 2 ? 12 : 6;
        float nb[9][12];
        int count = 0;
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dx = -1; dx <= 1; ++dx) {
                if (dx == 0 && dy == 0) continue;
                int nx = px + dx, ny = py + dy;
                if (nx < 0 || ny < 0 || nx >= c->grid_width || ny >= gh) continue;
                /* ROI links for every model: icgn.wgsl's quadratic branch (:1356-1400) has
                 * no ROI test because the reference rejects quadratic + ROI (roi.py:45-46);
                 * the library supports it and honours the links there as for affine. */
                if (c->roi_enabled) {
                    uint32_t perm = (uint32_t)roi_words[(size_t)i * 4];
                    uint32_t bit = (uint32_t)((dy + 1) * 3 + dx + 1);
                    if ((perm & (1u << bit)) == 0u) continue;
                }
                int ni = ny * c->grid_width + nx;
                const float* ns = state + (size_t)ni * stride;
                if (status_of(ns, model) < 0.5f) continue;
                float ox = -(float)c->subset_stride * (float)dx;
                float oy = -(float)c->subset_stride * (float)dy;
                if (model == 2) {
                    for (int k = 0; k < 12; ++k) nb[count][k] = recentered_component(ns, 2, k, ox, oy);
                } else {
                    float p1 = ns[1], p2 = ns[2], p4 = ns[4], p5 = ns[5];
                    nb[count][0] = ns[0] + p1 * ox + p2 * oy;
                    nb[count][1] = p1;
                    nb[count][2] = p2;
                    nb[count][3] = ns[3] + p4 * ox + p5 * oy;
                    nb[count][4] = p4;
                    nb[count][5] = p5;
                }
                count++;
            }
        }
        if (count < 4) continue;
        float filt[12], med[12];
        for (int k = 0; k < ncomp; ++k) filt[k] = med[k] = own[k];
        static const int order[12] = {0, 3, 1, 2, 4, 5, 6, 7, 8, 9, 10, 11};
        int model_count = model == 1 ? 2 : ncomp;
        int median_count = c->outlier_policy == 1 ? model_count : 2;
        float values[9];
        for (int ai = 0; ai < median_count; ++ai) {
            int comp = order[ai];
            for (int k = 0; k < count; ++k) values[k] = nb[k][comp];
            med[comp] = median_nine(values, count);
        }
        int flagged = 0;
        for (int ai = 0; ai < 2; ++ai) {
            int comp = order[ai];
            for (int k = 0; k < count; ++k) values[k] = fabsf(nb[k][comp] - med[comp]);
            float spread = median_nine(values, count);
            float score = fabsf(filt[comp] - med[comp]) / (spread + 0.1f);
            if (score > 2.0f) flagged = 1;
        }
        if (flagged) {
            if (c->outlier_policy == 1)
                for (int ai = 0; ai < model_count; ++ai) filt[order[ai]] = med[order[ai]];
            for (int k = 0; k < ncomp; ++k) out[k] = filt[k];
            if (model == 2) {
                out[12] = 2.0f;
                out[13] = own[13];
                out[14] = 0.0f;
                out[15] = 0.0f;
            } else {
                out[6] = 2.0f;
                out[7] = own[7];
            }
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Final affine ZNCC: shaders/affine_zncc.wgsl:57-99.                        */
/* ------------------------------------------------------------------------- */
void ora_affine_zncc(int w, int h, int gw, int gh, int S, int D, int interpolation, const float* reference,
                     const float* state, const float* coeff, const uint8_t* target, float* out) {
    ora_solve_cfg cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.width = w; cfg.height = h; cfg.interpolation = interpolation;
    solve_ctx s = {&cfg, reference, NULL, NULL, coeff, target};
    float half = ((float)S - 1.0f) * 0.5f;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < gw * gh; ++i) {
        const float* p = state + (size_t)i * 8;
        if (p[6] != 1.0f) {
            out[i] = -2.0f;
            continue;
        }
        int gx = i % gw, gy = i / gw;
        float acc[64][6];
        memset(acc, 0, sizeof acc);
        for (int lane = 0; lane < 64; ++lane) {
            for (int k = lane; k < S * S; k += 64) {
                int px = gx * D + k % S, py = gy * D + k / S;
                float dx = (float)(k % S) - half, dy = (float)(k / S) - half;
                float tx = (float)px + (p[0] + (p[1] * dx + p[2] * dy));
                float ty = (float)py + (p[3] + (p[4] * dx + p[5] * dy));
                int valid = tx >= 2.0f && ty >= 2.0f && tx <= (float)(w - 3) && ty <= (float)(h - 3);
                if (valid) {
                    float g;
                    if (interpolation == 1) {
                        float fx = floorf(tx), fy = floorf(ty);
                        int bx = (int)fx, by = (int)fy;
                        float ax = tx - fx, ay = ty - fy;
                        float a = source_at(&s, bx, by) / 255.0f, b = source_at(&s, bx + 1, by) / 255.0f;
                        float cc = source_at(&s, bx, by + 1) / 255.0f, d = source_at(&s, bx + 1, by + 1) / 255.0f;
                        g = 255.0f * mixf(mixf(a, b, ax), mixf(cc, d, ax), ay);
                    } else {
                        float fx = floorf(tx), fy = floorf(ty);
                        int bx = (int)fx, by = (int)fy;
                        float wx[4], wy[4];
                        if (interpolation == 2) { keys_weights(tx - fx, wx); keys_weights(ty - fy, wy); }
                        else { cubic_weights(tx - fx, wx); cubic_weights(ty - fy, wy); }
                        g = 0.0f;
                        for (int j = 0; j < 4; ++j)
                            for (int ii = 0; ii < 4; ++ii) {
                                float pix = interpolation == 0 ? coeff_at(&s, bx + ii - 1, by + j - 1)
                                                               : source_at(&s, bx + ii - 1, by + j - 1);
                                g += wx[ii] * wy[j] * pix;
                            }
                    }
                    float f = reference[((size_t)py * w + px) * 4];
                    acc[lane][0] += f; acc[lane][1] += g; acc[lane][2] += f * f; acc[lane][3] += g * g;
                    acc[lane][4] += f * g; acc[lane][5] += 1.0f;
                }
            }
        }
        for (int off = 32; off > 0; off /= 2)
            for (int lane = 0; lane < off; ++lane)
                for (int k = 0; k < 6; ++k) acc[lane][k] += acc[lane + off][k];
        float n = acc[0][5];
        float vf = acc[0][2] - acc[0][0] * acc[0][0] / fmaxf(n, 1.0f);
        float vg = acc[0][3] - acc[0][1] * acc[0][1] / fmaxf(n, 1.0f);
        if (n == (float)(S * S) && vf > 1e-9f && vg > 1e-9f) {
            float z = (acc[0][4] - acc[0][0] * acc[0][1] / n) / sqrtf(vf * vg);
            out[i] = fminf(fmaxf(z, -1.0f), 1.0f);
        } else {
            out[i] = -2.0f;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Integer-shift ZNCC seed search.  No reference counterpart: the nearest is  */
/* the candidate scoring of shaders/spatial_seed.wgsl:65-190 (ZNCC of a few   */
/* candidate translations).  Here every shift (dx, dy) in [-R, R]^2 around    */
/* the prior displacement (rounded half-to-even) is scored on the raw uint8   */
/* planes from exact integer moments, n = S^2:                               */
/*   num = n*Sfg - Sf*Sg,  vf = n*Sff - Sf^2,  vg = n*Sgg - Sg^2  (int64)      */
/*   zncc = (float)(num / sqrt((double)vf * (double)vg))                     */
/* A window leaving the image, or vf == 0 or vg == 0, scores -2.  The peak is */
/* the largest float score, ties to the lower row-major index (dy outer).    */
/* A parabola through the peak and its two neighbours per axis gives the     */
/* sub-pixel offset, 0 on the search border or next to an invalid shift.     */
/* mask (nullable): subsets with mask == 0 are skipped (shift 0, score NaN). */
/* ------------------------------------------------------------------------- */
void ora_zncc_search(const uint8_t* ref, const uint8_t* tgt, int w, int h, int gw, int gh, int S, int D, int R,
                     const float* prior, int prior_stride, const uint8_t* mask, float* out_shift, float* out_score) {
    const int K = 2 * R + 1;
    const int64_t n = (int64_t)S * S;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < gw * gh; ++i) {
        if (mask && !mask[i]) {
            out_shift[2 * i] = out_shift[2 * i + 1] = 0.0f;
            out_score[i] = NAN;
            continue;
        }
        const int ox = (i % gw) * D, oy = (i / gw) * D;
        int px = 0, py = 0;
        if (prior) {
            /* A prior that is not finite or lies beyond the image counts as 0. */
            float fx = prior[(size_t)i * prior_stride], fy = prior[(size_t)i * prior_stride + 3];
            px = fabsf(fx) <= (float)w ? (int)rintf(fx) : 0;
            py = fabsf(fy) <= (float)h ? (int)rintf(fy) : 0;
        }
        int64_t sf = 0, sff = 0;
        for (int r = 0; r < S; ++r)
            for (int c = 0; c < S; ++c) {
                int64_t f = ref[(size_t)(oy + r) * w + ox + c];
                sf += f;
                sff += f * f;
            }
        const int64_t vf = n * sff - sf * sf;
        float* scores = (float*)malloc(sizeof(float) * K * K);
        int best = 0;
        for (int s = 0; s < K * K; ++s) {
            const int x0 = ox + px - R + s % K, y0 = oy + py - R + s / K;
            float score = -2.0f;
            if (x0 >= 0 && y0 >= 0 && x0 + S <= w && y0 + S <= h && vf > 0) {
                int64_t sg = 0, sgg = 0, sfg = 0;
                for (int r = 0; r < S; ++r)
                    for (int c = 0; c < S; ++c) {
                        int64_t f = ref[(size_t)(oy + r) * w + ox + c];
                        int64_t g = tgt[(size_t)(y0 + r) * w + x0 + c];
                        sg += g;
                        sgg += g * g;
                        sfg += f * g;
                    }
                const int64_t vg = n * sgg - sg * sg;
                if (vg > 0) score = (float)((double)(n * sfg - sf * sg) / sqrt((double)vf * (double)vg));
            }
            scores[s] = score;
            if (score > scores[best]) best = s;
        }
        const int bx = best % K, by = best / K;
        float subx = 0.0f, suby = 0.0f;
        const float c0 = scores[best];
        if (c0 > -1.5f) {
            if (bx > 0 && bx < K - 1) {
                const float l = scores[best - 1], r = scores[best + 1];
                const float den = (l + r) - 2.0f * c0;
                if (l > -1.5f && r > -1.5f && den < 0.0f) subx = (0.5f * (l - r)) / den;
            }
            if (by > 0 && by < K - 1) {
                const float u = scores[best - K], d = scores[best + K];
                const float den = (u + d) - 2.0f * c0;
                if (u > -1.5f && d > -1.5f && den < 0.0f) suby = (0.5f * (u - d)) / den;
            }
        }
        out_shift[2 * i] = (float)(bx - R + px) + subx;
        out_shift[2 * i + 1] = (float)(by - R + py) + suby;
        out_score[i] = c0;
        free(scores);
    }
}
