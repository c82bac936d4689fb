// Inverse-compositional Gauss-Newton subset solver for gfx950.
//
// One wave64 owns one subset: one wave per workgroup in k_solve, four
// independent waves per workgroup in the persistent k_solve_batched, no block
// barriers.  Translation/affine subsets up to 16x16 keep their reference tile
// (f, gx, gy) in VGPRs; larger ones and every quadratic subset re-read it from
// L1/L2 each iteration.  The target footprint around the current warp is
// staged in the wave's LDS slice and re-staged only when the warp leaves it, so
// the per-iteration B-spline gathers (16 per pixel) hit LDS instead of L2.
// Sums are reduced with DPP inside the wave and every lane evaluates the
// (wave-uniform) increment, composition and convergence test, so no lane waits
// on another.
//
// Equations: shaders/icgn.wgsl:216-958 (sampling 230-309, increments 321-384,
// compositions 386-549, change metrics 551-593, solve loop 595-958).
#include "ldic_device.h"
#include "ldic_kernels.h"

namespace ldic {

// Cubic B-spline weights times 6, icgn.wgsl:230-240.  The solver's sums are
// invariant to a common scale of the sampled values (the mean, norm, ratio and
// ZNCC all cancel it), so the samplers leave out the 1/36 of the two axes and
// the norm threshold is scaled instead (kSampleScale).
// w[2](t) = w[1](1 - t) (the kernel is even), so w[2] reuses the powers of
// s = 1 - t that w[0] needs: 9 operations per axis.
__device__ __forceinline__ void bspline_weights6(float t, float w[4]) {
    const float s = 1.0f - t;
    const float t2 = t * t, t3 = t2 * t, s2 = s * s, s3 = s2 * s;
    w[0] = s3;
    w[1] = fmaf(3.0f, t3, fmaf(-6.0f, t2, 4.0f));
    w[2] = fmaf(3.0f, s3, fmaf(-6.0f, s2, 4.0f));
    w[3] = t3;
}

template <int INTERP>
constexpr float kSampleScale = INTERP == kBspline ? 36.0f : 1.0f;

__device__ __forceinline__ void keys_weights(float t, float w[4]) {
    float t2 = t * t, t3 = t2 * t;
    w[0] = -0.5f * t + t2 - 0.5f * t3;
    w[1] = 1.0f - 2.5f * t2 + 1.5f * t3;
    w[2] = 0.5f * t + 2.0f * t2 - 1.5f * t3;
    w[3] = -0.5f * t2 + 0.5f * t3;
}

// Target footprint held in LDS: box origin (x0, y0) and pitch.
struct Tile {
    int x0, y0, w, h;
    bool valid;
};

// Floor split of a positive coordinate: v_cvt_flr_i32_f32 and v_fract_f32 give
// (int)floor(x) and x - floor(x) exactly.
__device__ __forceinline__ void floor_split(float x, int& i, float& f) {
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(i) : "v"(x));
    f = __builtin_amdgcn_fractf(x);
}

// origin + iy * TP + ix for a tile with compile-time row pitch TP and a
// wave-uniform origin (the callers fold the staged box's corner and the
// stencil's -1 offsets into it), so (ix, iy) are image coordinates: one
// v_mad_u32_u24 and one shift-add (written out because the compiler folds the
// origin into the product and falls back to a full 32-bit multiply).  Image
// coordinates are non-negative and below 2^24.
template <int TP>
__device__ __forceinline__ const float* tile_tap(const float* origin, int ix, int iy) {
    int idx;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(idx) : "v"(iy), "n"(TP), "v"(ix));
    return origin + idx;
}

// The same with the wave-uniform pitch of the staged box (a scalar operand), for
// tiles without a compile-time pitch; pitches are below 2^24 too.
__device__ __forceinline__ const float* tile_tap_dyn(const float* origin, int ix, int iy, int pitch) {
    int idx;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(idx) : "v"(iy), "s"(pitch), "v"(ix));
    return origin + idx;
}

// TP > 0: compile-time row pitch, so the 16 taps are immediate offsets from one
// per-pixel LDS address; TP == 0: the pitch is the staged box width.
template <int INTERP, int TP>
__device__ __forceinline__ float tile_sample(const float* tile, const Tile& t, float px, float py) {
    const int pitch = TP > 0 ? TP : t.w;
    int ix, iy;
    float tx, ty;
    floor_split(px, ix, tx);
    floor_split(py, iy, ty);
    if (INTERP == kLinear) {
        const float* b = tile + (__mul24(iy - t.y0, pitch) + (ix - t.x0));
        float top = b[0] * (1.0f - tx) + b[1] * tx;
        float bottom = b[pitch] * (1.0f - tx) + b[pitch + 1] * tx;
        return top * (1.0f - ty) + bottom * ty;
    }
    float wx[4], wy[4];
    if (INTERP == kKeys) {
        keys_weights(tx, wx);
        keys_weights(ty, wy);
    } else {
        bspline_weights6(tx, wx);
        bspline_weights6(ty, wy);
    }
    const float* b;
    if constexpr (TP > 0)
        b = tile_tap<TP>(tile - (TP * (t.y0 + 1) + t.x0 + 1), ix, iy);
    else
        b = tile_tap_dyn(tile - (pitch * (t.y0 + 1) + t.x0 + 1), ix, iy, pitch);
    float v;
    const float* r = b;  // row pointers by addition (no multiply when the pitch is a register)
#pragma unroll
    for (int yy = 0; yy < 4; ++yy) {
        float row = wx[0] * r[0];
#pragma unroll
        for (int xx = 1; xx < 4; ++xx) row += wx[xx] * r[xx];
        v = yy == 0 ? wy[0] * row : v + wy[yy] * row;
        r += pitch;
    }
    return v;
}

// Direct gathers with clamped indices, for footprints that do not fit LDS.
template <int INTERP>
__device__ __forceinline__ float global_sample(const SolveArgs& a, float px, float py) {
    float fx = floorf(px), fy = floorf(py);
    int ix = (int)fx, iy = (int)fy;
    float tx = px - fx, ty = py - fy;
    const int W = a.width, H = a.height;
    auto at = [&](int x, int y) -> float {
        x = clampi(x, 0, W - 1);
        y = clampi(y, 0, H - 1);
        return INTERP == kBspline ? a.coeff[(size_t)y * W + x] : (float)a.target[(size_t)y * W + x];
    };
    if (INTERP == kLinear) {
        float top = at(ix, iy) * (1.0f - tx) + at(ix + 1, iy) * tx;
        float bottom = at(ix, iy + 1) * (1.0f - tx) + at(ix + 1, iy + 1) * tx;
        return top * (1.0f - ty) + bottom * ty;
    }
    float wx[4], wy[4];
    if (INTERP == kKeys) {
        keys_weights(tx, wx);
        keys_weights(ty, wy);
    } else {
        bspline_weights6(tx, wx);
        bspline_weights6(ty, wy);
    }
    float v = 0.0f;
#pragma unroll
    for (int yy = 0; yy < 4; ++yy) {
        float row = 0.0f;
#pragma unroll
        for (int xx = 0; xx < 4; ++xx) row += wx[xx] * at(ix + xx - 1, iy + yy - 1);
        v += wy[yy] * row;
    }
    return v;
}

// Warp state, wave-uniform.
struct Warp {
    float T[2];
    float L[4];
    float C[6];
};

template <int MODEL>
__device__ __forceinline__ void displacement(const Warp& q, float lx, float ly, float& dx, float& dy) {
    dx = q.T[0];
    dy = q.T[1];
    if (MODEL != kTranslation) {
        dx = dx + (q.L[0] * lx + q.L[1] * ly);
        dy = dy + (q.L[2] * lx + q.L[3] * ly);
        if (MODEL == kQuadratic) {
            dx += 0.5f * q.C[0] * lx * lx + q.C[1] * lx * ly + 0.5f * q.C[2] * ly * ly;
            dy += 0.5f * q.C[3] * lx * lx + q.C[4] * lx * ly + 0.5f * q.C[5] * ly * ly;
        }
    }
}

// Corner (or 3x3 for quadratic) bounds test, icgn.wgsl:648-685.
template <int MODEL>
// (bx0, by0)-(bx1, by1): local box whose corners are checked, the whole window
// (+-he) or a partial support's box (icgn.wgsl:638-660).
__device__ __forceinline__ bool warp_inside(const Warp& q, float cx, float cy, float he, float W, float H, float bx0,
                                            float by0, float bx1, float by1) {
    if (MODEL == kTranslation) {
        float minx = cx - 0.5f + bx0 + q.T[0], miny = cy - 0.5f + by0 + q.T[1];
        float maxx = cx - 0.5f + bx1 + q.T[0], maxy = cy - 0.5f + by1 + q.T[1];
        return minx >= 2.0f && miny >= 2.0f && maxx <= W - 3.0f && maxy <= H - 3.0f;
    }
    // Every point inside <=> the extremes inside and no coordinate NaN.  The
    // positions are the per-point ones; fminf/fmaxf skip a NaN operand, so a
    // NaN point is caught by the unordered test instead.  (Testing the points
    // one by one compiled to a 16-bit mask built lane by lane, twice the
    // instructions.)
    const int count = MODEL == kQuadratic ? 9 : 4;
    float x0 = 0.0f, x1 = 0.0f, y0 = 0.0f, y1 = 0.0f;
    bool nan = false;
#pragma unroll
    for (int p = 0; p < count; ++p) {
        float lx, ly;
        if (MODEL == kQuadratic) {
            lx = (float)(p % 3 - 1) * he;
            ly = (float)(p / 3 - 1) * he;
        } else {
            lx = (p == 1 || p == 3) ? bx1 : bx0;
            ly = p >= 2 ? by1 : by0;
        }
        float dx, dy;
        displacement<MODEL>(q, lx, ly, dx, dy);
        float px = cx - 0.5f + lx + dx, py = cy - 0.5f + ly + dy;
        nan = nan || __builtin_isunordered(px, py);
        x0 = p == 0 ? px : fminf(x0, px);
        x1 = p == 0 ? px : fmaxf(x1, px);
        y0 = p == 0 ? py : fminf(y0, py);
        y1 = p == 0 ? py : fmaxf(y1, py);
    }
    return !nan && x0 >= 2.0f && y0 >= 2.0f && x1 <= W - 3.0f && y1 <= H - 3.0f;
}

// icgn.wgsl:437-513; false when a pivot falls below 1e-8.  The Gauss-Jordan
// elimination of [M | I] runs column-parallel: lane e < 12 owns column e of the
// 6 x 12 augmented matrix (the 6 x 6 lifted warp operator M, then the
// identity), pivot columns are broadcast with readlane, and each lane applies
// the same row exchange, division and update to its own column.  Every value
// is formed by the same operations as in a one-lane elimination.
__device__ __forceinline__ bool compose_quadratic(Warp& q, const float d[12], const int lane) {
    const int e = lane < 12 ? lane : 11;
    float col[6];
    {
        float u = d[0], ux = d[1], uy = d[2], v = d[3], vx = d[4], vy = d[5];
        float uxx = d[6], uxy = d[7], uyy = d[8], vxx = d[9], vxy = d[10], vyy = d[11];
        float m[6][6] = {
            {1.0f + 2.0f * ux + ux * ux + u * uxx, 2.0f * u * uxy + 2.0f * (1.0f + ux) * uy, uy * uy + u * uyy,
             2.0f * u * (1.0f + ux), 2.0f * u * uy, u * u},
            {0.5f * (v * uxx + 2.0f * (1.0f + ux) * vx + u * vxx), 1.0f + uy * vx + ux * vy + v * uxy + u * vxy + vy + ux,
             0.5f * (v * uyy + 2.0f * uy * (1.0f + vy) + u * vyy), v + v * ux + u * vx, u + v * uy + u * vy, u * v},
            {vx * vx + v * vxx, 2.0f * v * vxy + 2.0f * vx * (1.0f + vy), 1.0f + 2.0f * vy + vy * vy + v * vyy,
             2.0f * v * vx, 2.0f * v * (1.0f + vy), v * v},
            {0.5f * uxx, uxy, 0.5f * uyy, 1.0f + ux, uy, u},
            {0.5f * vxx, vxy, 0.5f * vyy, vx, 1.0f + vy, v},
            {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f}};
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            float x = r == e - 6 ? 1.0f : 0.0f;
#pragma unroll
            for (int c = 0; c < 6; ++c) x = e == c ? m[r][c] : x;
            col[r] = x;
        }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        float pc[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) pc[r] = readlane_f32(col[r], c);
        // Partial pivoting on the magnitudes' bit patterns, which are wave-uniform, so
        // the search runs in scalar integer instructions: for non-NaN floats |a| > |b|
        // exactly when bits(|a|) > bits(|b|) as unsigned integers, and a NaN (bits
        // above those of infinity) neither wins nor lets a later candidate win over
        // it, as in the float comparison |pc[cand]| > |pc[pr]|.
        constexpr uint32_t kInfBits = 0x7f800000u;
        uint32_t pb = __float_as_uint(pc[c]) & 0x7fffffffu;
        int pr = c;
#pragma unroll
        for (int cand = c + 1; cand < 6; ++cand) {
            const uint32_t cb = __float_as_uint(pc[cand]) & 0x7fffffffu;
            if (cb <= kInfBits && pb <= kInfBits && cb > pb) {
                pr = cand;
                pb = cb;
            }
        }
        if (__uint_as_float(pb) < 1e-8f) return false;
        // The exchange (rarely needed: the lifted operator of an increment is close to
        // the identity) is skipped by a scalar branch.
        if (pr != c) {
#pragma unroll
            for (int r = c + 1; r < 6; ++r) {
                const bool sw = r == pr;
                float x = col[c], y = col[r];
                col[c] = sw ? y : x;
                col[r] = sw ? x : y;
                x = pc[c];
                y = pc[r];
                pc[c] = sw ? y : x;
                pc[r] = sw ? x : y;
            }
        }
        col[c] /= pc[c];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            if (r == c) continue;
            col[r] -= pc[r] * col[c];
        }
    }
    // Lane 6 + k holds column k of M^-1.
    const float cu[6] = {0.5f * q.C[0], q.C[1], 0.5f * q.C[2], 1.0f + q.L[0], q.L[1], q.T[0]};
    const float cv[6] = {0.5f * q.C[3], q.C[4], 0.5f * q.C[5], q.L[2], 1.0f + q.L[3], q.T[1]};
    float su = 0.0f, sv = 0.0f;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        su += cu[r] * col[r];
        sv += cv[r] * col[r];
    }
    float uu[6], uv[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        uu[k] = readlane_f32(su, 6 + k);
        uv[k] = readlane_f32(sv, 6 + k);
    }
    q.C[0] = 2.0f * uu[0];
    q.C[1] = uu[1];
    q.C[2] = 2.0f * uu[2];
    q.C[3] = 2.0f * uv[0];
    q.C[4] = uv[1];
    q.C[5] = 2.0f * uv[2];
    q.T[0] = uu[5];
    q.T[1] = uv[5];
    q.L[0] = uu[3] - 1.0f;
    q.L[1] = uu[4];
    q.L[2] = uv[3];
    q.L[3] = uv[4] - 1.0f;
    return true;
}

// icgn.wgsl:515-549
template <int MODEL>
__device__ __forceinline__ void compose_standard(Warp& q, const float d[6], float det) {
    if (MODEL == kTranslation) {
        q.T[0] = q.T[0] - d[0];
        q.T[1] = q.T[1] - d[3];
        return;
    }
    const float r = 1.0f / det;
    float i11 = (1.0f + d[5]) * r, i12 = -d[2] * r, i21 = -d[4] * r, i22 = (1.0f + d[1]) * r;
    float w11 = 1.0f + q.L[0], w12 = q.L[1], w21 = q.L[2], w22 = 1.0f + q.L[3];
    float itx = -(i11 * d[0] + i12 * d[3]);
    float ity = -(i21 * d[0] + i22 * d[3]);
    q.T[0] = q.T[0] + (w11 * itx + w12 * ity);
    q.T[1] = q.T[1] + (w21 * itx + w22 * ity);
    float u11 = w11 * i11 + w12 * i21, u12 = w11 * i12 + w12 * i22;
    float u21 = w21 * i11 + w22 * i21, u22 = w21 * i12 + w22 * i22;
    q.L[0] = u11 - 1.0f;
    q.L[1] = u12;
    q.L[2] = u21;
    q.L[3] = u22 - 1.0f;
}

__device__ __forceinline__ float hypot2(float a, float b) { return sqrtf(a * a + b * b); }

template <int MODEL>
__device__ __forceinline__ float increment_change(const float* d, float he) {
    if (MODEL == kTranslation) return hypot2(d[0], d[3]);
    float m = 0.0f;
    if (MODEL == kQuadratic) {
#pragma unroll
        for (int p = 0; p < 9; ++p) {
            float x = (float)(p % 3 - 1) * he, y = (float)(p / 3 - 1) * he;
            m = fmaxf(m, hypot2(d[0] + d[1] * x + d[2] * y + 0.5f * d[6] * x * x + d[7] * x * y + 0.5f * d[8] * y * y,
                                d[3] + d[4] * x + d[5] * y + 0.5f * d[9] * x * x + d[10] * x * y + 0.5f * d[11] * y * y));
        }
        return m;
    }
    // The largest squared length, then one square root: sqrtf is correctly
    // rounded and so monotone, and sqrtf(max) equals the max of the sqrtfs.  A
    // NaN length is skipped by fmaxf in either form.
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float x = (c == 1 || c == 3) ? he : -he, y = c >= 2 ? he : -he;
        const float a = d[0] + d[1] * x + d[2] * y, b = d[3] + d[4] * x + d[5] * y;
        m = fmaxf(m, a * a + b * b);
    }
    return sqrtf(m);
}

// Lane-parallel forms of the quadratic 3x3 tests: lane l evaluates grid point
// p = l % 9 at local (x, y) = ((p % 3 - 1) he, (p / 3 - 1) he) with the same
// expression as the serial loops above, instead of every lane evaluating all
// nine.  The vote and the maximum are exact, so the results are those of the
// serial forms.  Measured at 4K S25/D4: the solve fell from 6.96 to 5.8 ms.
__device__ __forceinline__ bool warp_inside_quadratic_lanes(const Warp& q, float cx, float cy, float x, float y,
                                                            float W, float H) {
    float dx, dy;
    displacement<kQuadratic>(q, x, y, dx, dy);
    const float px = cx - 0.5f + x + dx, py = cy - 0.5f + y + dy;
    const bool ok = px >= 2.0f && py >= 2.0f && px <= W - 3.0f && py <= H - 3.0f;
    return __builtin_amdgcn_ballot_w64(!ok) == 0;
}

__device__ __forceinline__ float increment_change_quadratic_lanes(const float* d, float x, float y) {
    float m = fmaxf(0.0f, hypot2(d[0] + d[1] * x + d[2] * y + 0.5f * d[6] * x * x + d[7] * x * y + 0.5f * d[8] * y * y,
                                 d[3] + d[4] * x + d[5] * y + 0.5f * d[9] * x * x + d[10] * x * y + 0.5f * d[11] * y * y));
    // Lanes 0-15 (row 0) hold all nine points; its maximum reaches every lane of row 0.
    m = fmaxf(m, dpp_f32<kQuadXor1>(m));
    m = fmaxf(m, dpp_f32<kQuadXor2>(m));
    m = fmaxf(m, dpp_f32<kRowHalfMirror>(m));
    m = fmaxf(m, dpp_f32<kRowMirror>(m));
    return readlane_f32(m, 0);
}

// Stage the target footprint [x0, x0+w) x [y0, y0+h) of the wave's LDS slice.
template <int INTERP, int TP>
__device__ __forceinline__ void stage_tile(const SolveArgs& a, float* tile, const Tile& t, int lane) {
    const int pitch = TP > 0 ? TP : t.w;
    auto load = [&](uint32_t g) -> float {
        return INTERP == kBspline ? load_global<float>(a.coeff, g * 4u) : (float)load_global<uint8_t>(a.target, g);
    };
    // A fixed pitch bounds the box width (t.w <= TP), so with TP <= 64 the
    // element-wise path below is never taken and is not compiled.
    if ((TP > 0 && TP <= 2 * 32) || t.w <= 2 * 32) {
        // Boxes up to 64 wide: half-wave h copies rows h, h + 2, ... (then the last
        // row in half-wave 0 when the height is odd); lane c copies columns c and
        // c + 32.  Lanes past the box width load its last column and skip the
        // store, so the loads of four row pairs are issued before the first wait
        // instead of one load and one wait per element.
        const int c = lane & 31, h = lane >> 5;
        const bool lo = c < t.w, hi = c + 32 < t.w;
        const uint32_t clo = (uint32_t)(lo ? c : t.w - 1), chi = (uint32_t)(hi ? c + 32 : t.w - 1);
        const uint32_t step = 2u * (uint32_t)a.width;
        uint32_t g = (uint32_t)(t.y0 + h) * (uint32_t)a.width + (uint32_t)t.x0;
        float* d = tile + (h * pitch + c);
        const int pairs = t.h >> 1;
        int k = 0;
        for (; k + 4 <= pairs; k += 4) {
            float v0[4], v1[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v0[u] = load(g + u * step + clo);
                v1[u] = load(g + u * step + chi);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (lo) d[u * 2 * pitch] = v0[u];
                if (hi) d[u * 2 * pitch + 32] = v1[u];
            }
            g += 4 * step;
            d += 8 * pitch;
        }
        for (; k < pairs; ++k) {
            const float x0 = load(g + clo), x1 = load(g + chi);
            if (lo) d[0] = x0;
            if (hi) d[32] = x1;
            g += step;
            d += 2 * pitch;
        }
        if ((t.h & 1) && h == 0) {
            const float x0 = load(g + clo), x1 = load(g + chi);
            if (lo) d[0] = x0;
            if (hi) d[32] = x1;
        }
        wave_lds_sync();
        return;
    }
    const int n = t.w * t.h;
    const float inv_w = 1.0f / (float)t.w;
    for (int i = lane; i < n; i += kWave) {
        int r = (int)(((float)i + 0.5f) * inv_w);
        int c = i - r * t.w;
        const uint32_t g = (uint32_t)(t.y0 + r) * (uint32_t)a.width + (uint32_t)(t.x0 + c);
        tile[r * pitch + c] = load(g);
    }
    wave_lds_sync();
}

// PPL > 0: the reference tile of each lane lives in registers (PPL pixels per
// lane, pixel p = lane + 64 k).  PPL == 0: any subset width, reference values
// are re-read from the plane every iteration.
// ordinal >= 0: a partial-support subset (entry `ordinal` of the partial list):
// only its support pixels enter the sums (solve_partial*, icgn.wgsl:1040-1085),
// its statistics carry validity -4, its own mean and weight sum, and its
// support box replaces the window in the bounds check.
template <int MODEL, int INTERP, int PPL, bool WEIGHTED, int TP>
__device__ __forceinline__ void solve_subset(const SolveArgs& a, const int subset, const int lane, float* tile,
                                             const int ordinal = -1) {
    constexpr int P = MODEL == kQuadratic ? 12 : 6;
    constexpr int M = MODEL == kQuadratic ? 28 : 16;
    constexpr int STRIDE = MODEL == kQuadratic ? 16 : 8;
    constexpr int NP = PPL > 0 ? PPL : 1;
    const float* meta = a.metadata + (size_t)subset * M;
    const float* inv = a.inverse + (size_t)subset * (P * P);

    const bool enabled = a.active ? a.active[subset] != 0u : true;
    if (!enabled && a.preserve_disabled) return;

    Warp q;
    q.T[0] = q.T[1] = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) q.L[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) q.C[k] = 0.0f;
    if (a.seed) {
        const float* sd = a.seed + (size_t)subset * STRIDE;
        q.T[0] = sd[0];
        q.T[1] = sd[3];
        if (MODEL != kTranslation) {
            q.L[0] = sd[1];
            q.L[1] = sd[2];
            q.L[2] = sd[4];
            q.L[3] = sd[5];
        }
        if (MODEL == kQuadratic)
#pragma unroll
            for (int k = 0; k < 6; ++k) q.C[k] = sd[6 + k];
    }
    const float meta1 = meta[1];
    // Partial-support passes run the weighted variants only, so the per-pixel
    // support test compiles away in the unweighted ones.
    const bool partial = WEIGHTED && ordinal >= 0;
    float status = partial ? ((enabled && meta1 == -4.0f) ? 0.0f : -1.0f)
                           : ((enabled && meta1 >= 0.5f) ? 0.0f : (a.roi_enabled ? fminf(-1.0f, meta1) : -1.0f));
    float iters = 0.0f;

    const int S = a.S, D = a.D, SS = S * S;
    const int gx = subset % a.grid_w, gy = subset / a.grid_w;
    const float he = 0.5f * ((float)S - 1.0f);
    const float cx = 0.5f * (float)S + (float)gx * (float)D;
    const float cy = 0.5f * (float)S + (float)gy * (float)D;
    const float ref_mean = (D != S || partial) ? meta[MODEL == kQuadratic ? 26 : 14] : 0.0f;
    const uint32_t* support = partial ? a.support + (size_t)ordinal * ((SS + 31) / 32) : nullptr;
    float bx0 = -he, by0 = -he, bx1 = he, by1 = he;
    if (partial) {
        const float4 box = a.support_boxes[ordinal];
        bx0 = box.x - he;
        by0 = box.y - he;
        bx1 = box.z - he;
        by1 = box.w - he;
    }
    const float wsum = WEIGHTED ? meta[MODEL == kQuadratic ? 27 : 15] : (float)SS;
    const float Wf = (float)a.width, Hf = (float)a.height;
    const int ox = gx * D, oy = gy * D;
    const float cx0 = cx - 0.5f, cy0 = cy - 0.5f;

    // Uncached sweeps: each lane's first pixel and the per-step increments.
    const int step_row = kWave / S, step_col = kWave - step_row * S;
    const int step_off = step_row * a.width + step_col;
    const int row0 = lane / S, col0 = lane - row0 * S;
    const int off0 = (oy + row0) * a.width + ox + col0;
    const float lx0 = (float)col0 - he, ly0 = (float)row0 - he;

    // Register-resident reference tile.
    float rf[NP], rgx[NP], rgy[NP], plx[NP], ply[NP], pw[WEIGHTED ? NP : 1];
    if (PPL > 0 && status == 0.0f) {
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            int p = lane + kWave * k;
            if (p < SS) {
                int row = p / S, col = p - row * S;
                float4 r = a.reference[(size_t)(oy + row) * a.width + ox + col];
                rf[k] = r.x - ref_mean;
                rgx[k] = r.y;
                rgy[k] = r.z;
                plx[k] = (float)col - he;
                ply[k] = (float)row - he;
                if (WEIGHTED) pw[k] = expf(-0.5f * (plx[k] * plx[k] + ply[k] * ply[k]) * a.inv_var);
                if (WEIGHTED && partial && !((support[p >> 5] >> (p & 31)) & 1u)) pw[k] = 0.0f;
            } else {
                rf[k] = rgx[k] = rgy[k] = plx[k] = ply[k] = 0.0f;
                if (WEIGHTED) pw[k] = 0.0f;
            }
        }
    }

    // Quadratic: lane r < 12 keeps row r of H^-1 for the update.
    float inv_row[MODEL == kQuadratic ? 12 : 1];
    if (MODEL == kQuadratic) {
        const float* row = inv + 12 * min(lane, 11);
#pragma unroll
        for (int k = 0; k < 12; ++k) inv_row[k] = row[k];
    }

    // Quadratic 3x3 bounds and change tests run lane-parallel: this lane's grid
    // point, formed where it is used so that it holds no registers across
    // iterations.  (The affine four-corner tests stay serial: measured at 4K
    // S31/D8, their lane-parallel form made the solve 7 % slower.)
    auto grid_point = [&](float& x, float& y) {
        const int p = lane % 9;
        x = (float)(p % 3 - 1) * he;
        y = (float)(p / 3 - 1) * he;
    };

    Tile t;
    t.valid = false;
    t.x0 = t.y0 = t.w = t.h = 0;
    float last_zncc = __builtin_nanf("");
    // Divergence patience (icgn.wgsl:103-111): an unconverged solve whose change
    // has not decreased for `patience` iterations after `min_iters` stops with
    // the internal status 3, reported as 0; a stopped solve skips the bounds test.
    float prev_change = 3.402823e38f;
    int stalled = 0;

    for (int it = 0; it < 30; ++it) {
        const bool stopped = a.patience > 0 && status == 3.0f;
        bool inside = stopped;
        if (!stopped) {
            if (MODEL == kQuadratic) {
                float x, y;
                grid_point(x, y);
                inside = warp_inside_quadratic_lanes(q, cx, cy, x, y, Wf, Hf);
            } else {
                inside = warp_inside<MODEL>(q, cx, cy, he, Wf, Hf, bx0, by0, bx1, by1);
            }
        }
        if (!inside && (!a.roi_enabled || status > -2.5f)) status = -1.0f;
        if (it >= a.iterations || fabsf(status) > 0.5f) break;

        // Conservative footprint of the warped subset (+1 px rounding slack).
        float ex = (fabsf(q.L[0]) + fabsf(q.L[1])) * he, ey = (fabsf(q.L[2]) + fabsf(q.L[3])) * he;
        if (MODEL == kQuadratic) {
            ex += (0.5f * fabsf(q.C[0]) + fabsf(q.C[1]) + 0.5f * fabsf(q.C[2])) * he * he;
            ey += (0.5f * fabsf(q.C[3]) + fabsf(q.C[4]) + 0.5f * fabsf(q.C[5])) * he * he;
        }
        const int lo = INTERP == kLinear ? 1 : 2, hi = INTERP == kLinear ? 2 : 3;
        int rx0 = (int)floorf(cx - 0.5f - he + q.T[0] - ex) - lo;
        int rx1 = (int)floorf(cx - 0.5f + he + q.T[0] + ex) + hi;
        int ry0 = (int)floorf(cy - 0.5f - he + q.T[1] - ey) - lo;
        int ry1 = (int)floorf(cy - 0.5f + he + q.T[1] + ey) + hi;
        bool in_image = rx0 >= 0 && ry0 >= 0 && rx1 <= a.width - 1 && ry1 <= a.height - 1;
        bool use_tile = false;
        if (in_image) {
            if (t.valid && rx0 >= t.x0 && ry0 >= t.y0 && rx1 < t.x0 + t.w && ry1 < t.y0 + t.h) {
                use_tile = true;
            } else {
                // A fixed pitch bounds the box width; its height by the slice size.
                auto fits = [&](const Tile& b) {
                    return TP > 0 ? (b.w <= TP && b.h * TP <= a.tile_cap) : (b.w * b.h <= a.tile_cap);
                };
                const int m = 2;
                Tile nt;
                nt.x0 = max(0, rx0 - m);
                nt.y0 = max(0, ry0 - m);
                nt.w = min(a.width - 1, rx1 + m) - nt.x0 + 1;
                nt.h = min(a.height - 1, ry1 + m) - nt.y0 + 1;
                if (!fits(nt)) {
                    nt.x0 = rx0;
                    nt.y0 = ry0;
                    nt.w = rx1 - rx0 + 1;
                    nt.h = ry1 - ry0 + 1;
                }
                if (fits(nt)) {
                    // Every lane computed the same box: keep it in scalar registers.
                    nt.x0 = __builtin_amdgcn_readfirstlane(nt.x0);
                    nt.y0 = __builtin_amdgcn_readfirstlane(nt.y0);
                    nt.w = __builtin_amdgcn_readfirstlane(nt.w);
                    nt.h = __builtin_amdgcn_readfirstlane(nt.h);
                    nt.valid = true;
                    t = nt;
                    stage_tile<INTERP, TP>(a, tile, t, lane);
                    use_tile = true;
                }
            }
        }

        float sc = 0.0f, sc2 = 0.0f, src = 0.0f;
        float sj[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        float sq[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        // Sample position (centre - 0.5 + l) + d(l): the first sum is an exact pixel
        // coordinate, so the position is rounded once, as in icgn.wgsl:741-743.
        // The local position plus displacement is formed first (|l| + |d| is small,
        // so it rounds finely), then the pixel coordinate of the centre is added
        // with one rounding per sample, as in icgn.wgsl:741-743.
        const float ux = 1.0f + q.L[0], vy = 1.0f + q.L[3];
        // Quadratic: the halves of the squared-term coefficients; the sums over
        // lx^2 and ly^2 run without the 1/2, which is applied (exactly) after the
        // reduction.
        const float hc0 = 0.5f * q.C[0], hc2 = 0.5f * q.C[2], hc3 = 0.5f * q.C[3], hc5 = 0.5f * q.C[5];
        auto accumulate = [&](auto sampler, float lx, float ly, float f, float gxv, float gyv, float iw,
                              float valid) {
            float ex, ey;
            float xx = 0.0f, xy = 0.0f, yy = 0.0f;
            if (MODEL == kQuadratic) {
                // l + d(l) from the three basis products the Jacobian sums use as
                // well, one rounding per term, the largest terms first.
                xx = lx * lx;
                xy = lx * ly;
                yy = ly * ly;
                ex = fmaf(hc0, xx, fmaf(q.C[1], xy, fmaf(hc2, yy, fmaf(q.L[1], ly, fmaf(ux, lx, q.T[0])))));
                ey = fmaf(hc3, xx, fmaf(q.C[4], xy, fmaf(hc5, yy, fmaf(vy, ly, fmaf(q.L[2], lx, q.T[1])))));
            } else if (MODEL == kTranslation) {
                ex = lx + q.T[0];
                ey = ly + q.T[1];
            } else {
                ex = fmaf(q.L[1], ly, fmaf(ux, lx, q.T[0]));
                ey = fmaf(q.L[2], lx, fmaf(vy, ly, q.T[1]));
            }
            float cur = sampler(cx0 + ex, cy0 + ey) * valid;
            float wc = iw * cur;
            float gxg = gxv * wc, gyg = gyv * wc;
            sc += wc;
            sc2 += wc * cur;
            src += iw * f * cur;
            sj[0] += gxg;
            sj[3] += gyg;
            if (MODEL != kTranslation) {
                sj[1] += gxg * lx;
                sj[2] += gxg * ly;
                sj[4] += gyg * lx;
                sj[5] += gyg * ly;
                if (MODEL == kQuadratic) {
                    sq[0] += gxg * xx;
                    sq[1] += gxg * xy;
                    sq[2] += gxg * yy;
                    sq[3] += gyg * xx;
                    sq[4] += gyg * xy;
                    sq[5] += gyg * yy;
                }
            }
        };
        auto sweep = [&](auto sampler) {
            if (PPL > 0) {
                // Slots below `full` hold a pixel in every lane; slot `full` only in
                // lanes below `rem`: its other lanes sample the subset centre and
                // contribute zero, so every slot runs with the whole wave and the
                // slot test is a scalar branch.
                const int full = SS / kWave, rem = SS - full * kWave;
                const float tail = lane < rem ? 1.0f : 0.0f;
#pragma unroll
                for (int k = 0; k < NP; ++k) {
                    if (k <= full)
                        accumulate(sampler, plx[k], ply[k], rf[k], rgx[k], rgy[k], WEIGHTED ? pw[k] : 1.0f,
                                   k == full ? tail : 1.0f);
                }
            } else {
                // Pixel p = lane + 64 j: the column, row, plane offset and local
                // coordinates advance by (64 mod S, 64 div S) per step, with one
                // carry, instead of a division per pixel.
                if (MODEL != kQuadratic && !WEIGHTED && S >= 30 && S <= 32) {
                    // One subset row per half-wave: lane 32h + c takes column c of
                    // rows h, h + 2, ... (columns c >= S idle).  The taps a half-wave
                    // reads then lie in one tile row and fall in distinct LDS banks;
                    // with rows packed across the wave, the pixels of the next row
                    // read one tile pitch further on and collide with a bank of
                    // the current row.  For S = 30..32 this costs at most 7 % more
                    // lane slots.
                    const int c = lane & 31;
                    const bool on = c < S;
                    int r = lane >> 5;
                    int off = (oy + r) * a.width + ox + (on ? c : 0);
                    const float lxc = (float)c - he;
                    float ly = (float)r - he;
                    for (; r < S; r += 2) {
                        if (on) {
                            const float4 rv = a.reference[off];
                            accumulate(sampler, lxc, ly, rv.x - ref_mean, rv.y, rv.z, 1.0f, 1.0f);
                        }
                        off += 2 * a.width;
                        ly += 2.0f;
                    }
                    return;
                }
                // The column carry stays a branch.  A select-based carry (and a
                // wave-uniform trip count) keeps its float increments in VGPRs,
                // and at the 128-VGPR budget of the quadratic kernels the compiler
                // then waited for each of the 16 LDS tap reads before issuing the
                // next in the persistent batched kernel (measured: its subset-41
                // launch 10.7 -> 15.1 ms).  The reference is addressed by a 32-bit
                // byte offset in the global address space.
                int col = col0, off = off0;
                float lx = lx0, ly = ly0;
                for (int p = lane; p < SS; p += kWave) {
                    const float3 r = load_global<float3>(a.reference, (uint32_t)off * 16u);
                    float iw = WEIGHTED ? expf(-0.5f * (lx * lx + ly * ly) * a.inv_var) : 1.0f;
                    if (partial && !((support[p >> 5] >> (p & 31)) & 1u)) iw = 0.0f;
                    accumulate(sampler, lx, ly, r.x - ref_mean, r.y, r.z, iw, 1.0f);
                    col += step_col;
                    off += step_off;
                    lx += (float)step_col;
                    ly += (float)step_row;
                    if (col >= S) {
                        col -= S;
                        off += a.width - S;
                        lx -= (float)S;
                        ly += 1.0f;
                    }
                }
            }
        };
        // One subset row per half-wave, B-spline taps from the LDS tile: the same
        // values and operations as the row branch of `sweep` with `tile_sample`,
        // scheduled for fewer instructions per pixel.
        // - The floor split uses v_fract_f32 and v_cvt_flr_i32_f32 (exact for the
        //   positive coordinates inside the tile).
        // - With a compile-time pitch the 16 taps are immediate offsets from one
        //   LDS address, formed by one 24-bit multiply-add from (ix, iy).
        // - The trip count is wave-uniform: rows 2k + h for k < S/2 in both
        //   half-waves, then row S - 1 in the first half-wave when S is odd.
        auto row_sweep_tile = [&]() {
            const int pitch = TP > 0 ? TP : t.w;
            const int c = lane & 31, h = lane >> 5;
            const bool on = c < S;
            const float lxc = (float)c - he;
            // Lane-constant parts of the warped position.
            const float axl = fmaf(ux, lxc, q.T[0]);
            const float ayl = fmaf(q.L[2], lxc, q.T[1]);
            // Tap (ix, iy) sits at iy * pitch + ix - (pitch (y0 + 1) + x0 + 1) in the slice.
            const float* origin = tile - (__mul24(t.y0 + 1, pitch) + t.x0 + 1);
            uint32_t boff = (uint32_t)((oy + h) * a.width + ox + (on ? c : 0)) * 16u;
            const uint32_t rstep = (uint32_t)(2 * a.width) * 16u;
            float ly = (float)h - he;
            auto pixel = [&]() {
                const float3 rv = load_global<float3>(a.reference, boff);
                const float px = cx0 + fmaf(q.L[1], ly, axl);
                const float py = cy0 + fmaf(vy, ly, ayl);
                int ix, iy;
                float tx, ty;
                floor_split(px, ix, tx);
                floor_split(py, iy, ty);
                float wx[4], wy[4];
                bspline_weights6(tx, wx);
                bspline_weights6(ty, wy);
                const float* b = tile_tap<TP>(origin, ix, iy);
                float cur;
#pragma unroll
                for (int yy = 0; yy < 4; ++yy) {
                    const float* r = b + yy * pitch;
                    float row = wx[0] * r[0];
#pragma unroll
                    for (int xx = 1; xx < 4; ++xx) row += wx[xx] * r[xx];
                    cur = yy == 0 ? wy[0] * row : cur + wy[yy] * row;
                }
                const float f = rv.x - ref_mean;
                const float gxg = rv.y * cur, gyg = rv.z * cur;
                sc += cur;
                sc2 += cur * cur;
                src += f * cur;
                sj[0] += gxg;
                sj[3] += gyg;
                if (MODEL != kTranslation) {
                    sj[2] += gxg * ly;
                    sj[5] += gyg * ly;
                }
            };
            for (int k = 0; k < (S >> 1); ++k) {
                if (on) pixel();
                boff += rstep;
                ly += 2.0f;
            }
            if ((S & 1) && on && h == 0) pixel();
            // The lane's column is constant: its sums of g gx lx and g gy lx are
            // lx times its sums of g gx and g gy.
            if (MODEL != kTranslation) {
                sj[1] = sj[0] * lxc;
                sj[4] = sj[3] * lxc;
            }
        };
        // The sampling path is chosen once per iteration (wave-uniform).
        if (TP > 0 && MODEL != kQuadratic && !WEIGHTED && INTERP == kBspline && S >= 30 && S <= 32 && use_tile)
            row_sweep_tile();
        else if (use_tile)
            sweep([&](float px, float py) { return tile_sample<INTERP, TP>(tile, t, px, py); });
        else
            sweep([&](float px, float py) { return global_sample<INTERP>(a, px, py); });

        {
            // All of the iteration's sums in one multi-value wave reduction.
            constexpr int NS = MODEL == kTranslation ? 5 : MODEL == kQuadratic ? 15 : 9;
            float r[NS];
            r[0] = sc;
            r[1] = sc2;
            r[2] = src;
            r[3] = sj[0];
            r[4] = sj[3];
            if constexpr (MODEL != kTranslation) {
                r[5] = sj[1];
                r[6] = sj[2];
                r[7] = sj[4];
                r[8] = sj[5];
            }
            if constexpr (MODEL == kQuadratic) {
#pragma unroll
                for (int k = 0; k < 6; ++k) r[9 + k] = sq[k];
            }
            wave_sums(r, lane);
            sc = r[0];
            sc2 = r[1];
            src = r[2];
            sj[0] = r[3];
            sj[3] = r[4];
            if constexpr (MODEL != kTranslation) {
                sj[1] = r[5];
                sj[2] = r[6];
                sj[4] = r[7];
                sj[5] = r[8];
            }
            if constexpr (MODEL == kQuadratic) {
#pragma unroll
                for (int k = 0; k < 6; ++k) sq[k] = r[9 + k];
            }
        }
        if (MODEL == kQuadratic) {
            // Sums of g lx^2 and g ly^2 to sums of g lx^2 / 2 and g ly^2 / 2.
            sq[0] *= 0.5f;
            sq[2] *= 0.5f;
            sq[3] *= 0.5f;
            sq[5] *= 0.5f;
        }

        // Hardware reciprocals (1 ulp) of per-subset constants replace two of the
        // iteration's divisions; the norm keeps a correctly rounded one.
        const float mean = sc * __builtin_amdgcn_rcpf(wsum);
        const float norm = sqrtf(fmaxf(0.0f, sc2 - sc * mean));
        const float ref_norm = meta[0];
        iters = (float)(it + 1);
        if (norm < 1e-9f * kSampleScale<INTERP> || ref_norm < 1e-9f) {
            status = -2.0f;
            continue;
        }
        const float inv_norm = 1.0f / norm;
        const float ratio = ref_norm * inv_norm;
        const float zncc = src * inv_norm * __builtin_amdgcn_rcpf(ref_norm);
        last_zncc = zncc;
        float change;
        if (MODEL == kQuadratic) {
            float rj[12] = {sj[0], sj[1], sj[2], sj[3], sj[4], sj[5], sq[0], sq[1], sq[2], sq[3], sq[4], sq[5]};
            float res[12], d[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) res[k] = meta[2 + k] - ratio * (rj[k] - mean * meta[14 + k]);
            // Lane r < 12 forms row r of H^-1 res; the rows are then broadcast.
            float v = 0.0f;
#pragma unroll
            for (int k = 0; k < 12; ++k) v += inv_row[k] * res[k];
#pragma unroll
            for (int r = 0; r < 12; ++r) d[r] = -readlane_f32(v, r);
            if (!compose_quadratic(q, d, lane)) {
                status = -1.0f;
                continue;
            }
            float x, y;
            grid_point(x, y);
            change = increment_change_quadratic_lanes(d, x, y);
        } else {
            float d[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            if (MODEL == kTranslation) {
                float rx = meta[2] - ratio * (sj[0] - mean * meta[8]);
                float ry = meta[5] - ratio * (sj[3] - mean * meta[11]);
                d[0] = -(inv[0] * rx + inv[3] * ry);
                d[3] = -(inv[18] * rx + inv[21] * ry);
            } else {
                float res[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) res[k] = meta[2 + k] - ratio * (sj[k] - mean * meta[8 + k]);
#pragma unroll
                for (int r = 0; r < 6; ++r) {
                    float v = 0.0f;
#pragma unroll
                    for (int k = 0; k < 6; ++k) v += inv[6 * r + k] * res[k];
                    d[r] = -v;
                }
            }
            float det = MODEL == kTranslation ? 1.0f : (1.0f + d[1]) * (1.0f + d[5]) - d[2] * d[4];
            if (fabsf(det) < 1e-8f) {
                status = -1.0f;
                continue;
            }
            compose_standard<MODEL>(q, d, det);
            change = increment_change<MODEL>(d, he);
        }
        status = change >= a.tolerance ? 0.0f : (zncc >= a.zncc_min ? 1.0f : -2.0f);
        if (a.patience > 0 && status == 0.0f) {
            if (it + 1 >= a.min_iters) stalled = change >= prev_change ? stalled + 1 : 0;
            prev_change = change;
            if (stalled >= a.patience) status = 3.0f;
        }
    }
    if (a.patience > 0 && status == 3.0f) status = 0.0f;

    if (lane == 0) {
        float4* out = reinterpret_cast<float4*>(a.state + (size_t)subset * STRIDE);
        out[0] = make_float4(q.T[0], q.L[0], q.L[1], q.T[1]);
        if (MODEL == kQuadratic) {
            out[1] = make_float4(q.L[2], q.L[3], q.C[0], q.C[1]);
            out[2] = make_float4(q.C[2], q.C[3], q.C[4], q.C[5]);
            out[3] = make_float4(status, iters, 0.0f, 0.0f);
        } else {
            out[1] = make_float4(q.L[2], q.L[3], status, iters);
        }
        if (a.zncc_out) a.zncc_out[subset] = last_zncc;
    }
}

// Register budget (waves per SIMD = 512 / VGPRs, allocated in steps of 8):
// - register-cached translation/affine variants (up to 16 pixels per lane):
//   3 waves, 168 VGPRs;
// - uncached translation/affine: 5 waves, 96 VGPRs, as many as the LDS slice
//   of a subset-31 footprint allows (98 VGPRs would leave 4);
// - uncached quadratic: 4 waves, 128 VGPRs, also in the persistent batched
//   kernel.
// None of these spill.  Keys and Gaussian-weighted variants keep the compiler's
// own allocation.
constexpr int solve_min_waves(int model, int interp, int ppl, bool weighted, bool batched = false) {
    if (interp == kKeys || weighted) return 1;
    if (ppl == 0) return model == kQuadratic ? 4 : 5;
    return (model != kQuadratic && ppl <= 16) ? 3 : 1;
}

template <int MODEL, int INTERP, int PPL, bool WEIGHTED, int TP>
__global__ __attribute__((amdgpu_flat_work_group_size(1, 256),
                          amdgpu_waves_per_eu(solve_min_waves(MODEL, INTERP, PPL, WEIGHTED)))) void k_solve(SolveArgs a) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = xcd_block(blockIdx.x, gridDim.x) * kSolveBlockWaves + wave;
    if (WEIGHTED && a.partial_list) {
        // Partial-support pass over the list of partial subsets.
        if (item >= a.partial_count) return;
        solve_subset<MODEL, INTERP, PPL, WEIGHTED, TP>(a, (int)a.partial_list[item], lane, lds + wave * a.tile_cap,
                                                       item);
        return;
    }
    if (item >= a.grid_w * a.grid_h) return;
    solve_subset<MODEL, INTERP, PPL, WEIGHTED, TP>(a, item, lane, lds + wave * a.tile_cap);
}

// Persistent solve over the subset grids of several image pairs (one variant
// of model, interpolant, register bucket and weighting).  Work item i is a
// block of (workgroup waves) subsets of pair p, block_start[p] <= i <
// block_start[p + 1]; each workgroup strides over the items, the XCD mapping
// giving every XCD contiguous runs of them.  Waves never synchronise with each
// other, so a wave moves to its next item as soon as its subset is done.
template <int MODEL, int INTERP, int PPL, bool WEIGHTED, int TP>
__global__ __attribute__((amdgpu_flat_work_group_size(1, 256),
                          amdgpu_waves_per_eu(solve_min_waves(MODEL, INTERP, PPL, WEIGHTED, true)))) void k_solve_batched(
    const SolveArgs* __restrict__ pairs,
                                                       const int* __restrict__ block_start, int npairs, int total,
                                                       int wave_stride) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int waves = __builtin_amdgcn_readfirstlane(blockDim.x >> 6);
    float* tile = lds + wave * wave_stride;
    for (int item = xcd_block(blockIdx.x, gridDim.x); item < total; item += gridDim.x) {
        int p = 0;
        while (p + 1 < npairs && block_start[p + 1] <= item) ++p;
        const SolveArgs a = pairs[p];
        const int subset = (item - block_start[p]) * waves + wave;
        if (subset < a.grid_w * a.grid_h) solve_subset<MODEL, INTERP, PPL, WEIGHTED, TP>(a, subset, lane, tile);
    }
}

// Register-cached reference pixels per lane (0: the reference is re-read from
// L1/L2 every iteration).  Measured on MI355X at 4K, the uncached sweep is the
// faster one from subset 21 on for translation and affine (22-32 % at 25-41) and
// at every size for the quadratic model (1.2-2.8x), where 12 more sums leave no
// room for a register cache at useful occupancy; only translation/affine
// subsets of at most 16x16 keep their 4 pixels per lane in registers.
int solver_pixels_per_lane(int S, int model) {
    if (model == kQuadratic) return 0;
    return S * S <= 4 * kWave ? 4 : 0;
}

// Tile pitch of each register-cached variant (0: pitch = staged width).
static int solver_pitch(int ppl) { return ppl == 4 ? 32 : 0; }

// Fixed row pitch of the uncached B-spline variants: translation/affine
// subsets 30-32 (unweighted, the row sweep) and quadratic subsets up to 32
// (weighted or not).  The zero-strain footprint with its re-staging margin is
// S + 9 wide and high, so 44 columns hold it up to subset 35.
constexpr int kRowSweepPitch = 44;

int solver_tile_pitch(int S, int model, int interp, int weighted) {
    const int ppl = solver_pixels_per_lane(S, model);
    if (ppl > 0) return solver_pitch(ppl);
    if (interp != kBspline) return 0;
    if (model == kQuadratic) return S <= 32 ? kRowSweepPitch : 0;
    return (!weighted && S >= 30 && S <= 32) ? kRowSweepPitch : 0;
}

size_t solver_lds_bytes(int S, int model, int* tile_cap) {
    // Footprint of a moderately strained subset plus the interpolation halo
    // and the re-staging margin; larger footprints fall back to L2 gathers.
    int pitch = solver_pitch(solver_pixels_per_lane(S, model));
    int cap;
    if (pitch > 0) {
        cap = pitch * (S + 17);
    } else {
        // Up to subset 32, room for strain beyond the margin; above it the
        // footprint with its margin (S + 9 square) only, so a 41 x 41 subset's
        // 10 KB slice leaves room for 4 waves per SIMD (16 per CU).
        // Quadratic subsets up to 32 keep S + 12 rows of the fixed 44-word pitch.
        cap = S <= 32 ? (S + 12) * (model == kQuadratic ? kRowSweepPitch : S + 12) : (S + 9) * (S + 9);
        if (cap > 4096) cap = 4096;  // 16 KB of LDS per wave
    }
    *tile_cap = cap;
    return (size_t)cap * sizeof(float) * kSolveBlockWaves;
}

template <int MODEL, int INTERP, bool WEIGHTED>
static hipError_t launch_ppl(int ppl, const SolveArgs& a, hipStream_t s) {
    int N = a.partial_list ? a.partial_count : a.grid_w * a.grid_h;
    if (N <= 0) return hipSuccess;
    if (a.partial_list && !WEIGHTED) return hipErrorInvalidValue;  // partial passes use the weighted variant
    int cap = 0;
    size_t lds = solver_lds_bytes(a.S, MODEL, &cap);
    if (cap != a.tile_cap) return hipErrorInvalidValue;
    dim3 grid((N + kSolveBlockWaves - 1) / kSolveBlockWaves), block(kWave * kSolveBlockWaves);
    if constexpr (INTERP == kBspline && (MODEL == kQuadratic || !WEIGHTED)) {
        // Fixed row pitch: immediate tap offsets (and the row sweep for
        // translation/affine subsets 30-32).
        if (ppl == 0 && solver_tile_pitch(a.S, MODEL, INTERP, WEIGHTED) == kRowSweepPitch) {
            k_solve<MODEL, INTERP, 0, WEIGHTED, kRowSweepPitch><<<grid, block, lds, s>>>(a);
            return hipGetLastError();
        }
    }
    switch (ppl) {
        case 4: k_solve<MODEL, INTERP, 4, WEIGHTED, 32><<<grid, block, lds, s>>>(a); break;
        default: k_solve<MODEL, INTERP, 0, WEIGHTED, 0><<<grid, block, lds, s>>>(a); break;
    }
    return hipGetLastError();
}

template <int MODEL, int INTERP>
static hipError_t launch_w(int weighted, const SolveArgs& a, hipStream_t s) {
    int ppl = solver_pixels_per_lane(a.S, MODEL);
    return weighted ? launch_ppl<MODEL, INTERP, true>(ppl, a, s) : launch_ppl<MODEL, INTERP, false>(ppl, a, s);
}

template <int MODEL>
static hipError_t launch_i(int interp, int weighted, const SolveArgs& a, hipStream_t s) {
    switch (interp) {
        case kBspline: return launch_w<MODEL, kBspline>(weighted, a, s);
        case kLinear: return launch_w<MODEL, kLinear>(weighted, a, s);
        case kKeys: return launch_w<MODEL, kKeys>(weighted, a, s);
    }
    return hipErrorInvalidValue;
}

template <int MODEL, int INTERP, bool WEIGHTED>
static hipError_t launch_batched_ppl(const BatchLaunch& b, hipStream_t s) {
    if (b.total <= 0) return hipSuccess;
    if (b.waves != 1 && b.waves != 2 && b.waves != 4) return hipErrorInvalidValue;
    dim3 grid(b.grid), block(kWave * b.waves);
    if constexpr (INTERP == kBspline && (MODEL == kQuadratic || !WEIGHTED)) {
        if (b.ppl == 0 && b.pitch == kRowSweepPitch) {
            k_solve_batched<MODEL, INTERP, 0, WEIGHTED, kRowSweepPitch><<<grid, block, b.lds, s>>>(
                b.pairs, b.block_start, b.npairs, b.total, b.wave_stride);
            return hipGetLastError();
        }
    }
    if (b.pitch != solver_pitch(b.ppl)) return hipErrorInvalidValue;
    switch (b.ppl) {
        case 4: k_solve_batched<MODEL, INTERP, 4, WEIGHTED, 32><<<grid, block, b.lds, s>>>(b.pairs, b.block_start, b.npairs, b.total, b.wave_stride); break;
        default: k_solve_batched<MODEL, INTERP, 0, WEIGHTED, 0><<<grid, block, b.lds, s>>>(b.pairs, b.block_start, b.npairs, b.total, b.wave_stride); break;
    }
    return hipGetLastError();
}

template <int MODEL, int INTERP>
static hipError_t launch_batched_w(const BatchLaunch& b, hipStream_t s) {
    return b.weighted ? launch_batched_ppl<MODEL, INTERP, true>(b, s) : launch_batched_ppl<MODEL, INTERP, false>(b, s);
}

template <int MODEL>
static hipError_t launch_batched_i(const BatchLaunch& b, hipStream_t s) {
    switch (b.interp) {
        case kBspline: return launch_batched_w<MODEL, kBspline>(b, s);
        case kLinear: return launch_batched_w<MODEL, kLinear>(b, s);
        case kKeys: return launch_batched_w<MODEL, kKeys>(b, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_solve_batched(const BatchLaunch& b, hipStream_t s) {
    switch (b.model) {
        case kAffine: return launch_batched_i<kAffine>(b, s);
        case kTranslation: return launch_batched_i<kTranslation>(b, s);
        case kQuadratic: return launch_batched_i<kQuadratic>(b, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_solve(int model, int interp, int weighted, const SolveArgs& a, hipStream_t s) {
    switch (model) {
        case kAffine: return launch_i<kAffine>(interp, weighted, a, s);
        case kTranslation: return launch_i<kTranslation>(interp, weighted, a, s);
        case kQuadratic: return launch_i<kQuadratic>(interp, weighted, a, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace ldic
