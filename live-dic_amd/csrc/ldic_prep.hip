// Per-subset reference statistics in float64 (wgpu_backend.py:639-823): the
// anchored mean, energy, J^T f, sum J, H = sum w J J^T and its inverse by
// Gauss-Jordan elimination in diagonal order.  One wave64 per subset:
// lanes stride over the subset pixels, DPP wave sums give every lane the
// totals, and the inversion runs column-parallel on the wave's LDS slice.
#include "ldic_device.h"
#include "ldic_kernels.h"

#include <float.h>

#include <type_traits>
#include <utility>

namespace ldic {

constexpr int kPrepWaves = 4;

// Row-major upper-triangle enumeration of a P x P symmetric matrix.
__host__ __device__ constexpr int tri_row(int e, int P) {
    int i = 0;
    while (e >= P - i) {
        e -= P - i;
        ++i;
    }
    return i;
}
__host__ __device__ constexpr int tri_col(int e, int P) {
    int i = 0;
    while (e >= P - i) {
        e -= P - i;
        ++i;
    }
    return i + e;
}

// Calls f(std::integral_constant<int, E>) for E = 0 .. N-1, so that indices
// derived from E (the triangle row and column above) are compile-time
// constants: registers instead of scratch arrays, no runtime index loops.
template <typename F, int... Es>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Es...>) {
    (f(std::integral_constant<int, Es>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

template <int P>
__device__ __forceinline__ void jacobian(double gx, double gy, double lx, double ly, double J[P]) {
#pragma clang fp contract(off)
    J[0] = gx;
    J[3] = gy;
    if (P > 2) {
        J[1] = gx * lx;
        J[2] = gx * ly;
        J[4] = gy * lx;
        J[5] = gy * ly;
    }
    if (P == 12) {
        double hx2 = 0.5 * lx * lx, xy = lx * ly, hy2 = 0.5 * ly * ly;
        J[6] = gx * hx2;
        J[7] = gx * xy;
        J[8] = gx * hy2;
        J[9] = gy * hx2;
        J[10] = gy * xy;
        J[11] = gy * hy2;
    }
}

struct PixelRef {
    double f, gx, gy, lx, ly, w;
};

__device__ __forceinline__ PixelRef pixel(const PrepArgs& a, int p, int ox, int oy, double half, double mean) {
    int row = p / a.S, col = p - row * a.S;
    size_t idx = (size_t)(oy + row) * a.width + ox + col;
    PixelRef r;
    r.f = a.intensity[idx] - mean;
    r.gx = a.grad_x[idx];
    r.gy = a.grad_y[idx];
    r.lx = (double)col - half;
    r.ly = (double)row - half;
    r.w = a.weighted ? a.weights[p] : 1.0;
    return r;
}

// Moments of one subset.  One pixel pass accumulates the Hessian entries
// [E0, E1) of the row-major upper triangle and, in the FIRST pass, also the
// energy, J^T f and sum J (and stores the tile-centred intensity when D == S).
// Totals are wave sums; the Hessian goes to the LDS matrix `h` with row pitch
// `pitch`, the other totals to `first`.
struct FirstMoments {
    double energy;
    double jtf[12];
    double sj[12];
};

template <int P, int E0, int E1, bool FIRST>
__device__ void moment_pass(const PrepArgs& a, int lane, int ox, int oy, double half, double mean, double* h,
                            int pitch, FirstMoments* first) {
#pragma clang fp contract(off)
    constexpr int NE = E1 - E0;
    constexpr int N = P == 2 ? 2 : P;          // size of the symmetric system
    constexpr int JP = P == 2 ? 6 : P;         // Jacobian components evaluated
    double acc[NE > 0 ? NE : 1];
#pragma unroll
    for (int e = 0; e < NE; ++e) acc[e] = 0.0;
    double energy = 0.0, jtf[JP], sj[JP];
#pragma unroll
    for (int k = 0; k < JP; ++k) jtf[k] = sj[k] = 0.0;
    const int SS = a.S * a.S;
    for (int p = lane; p < SS; p += kWave) {
        PixelRef r = pixel(a, p, ox, oy, half, mean);
        double J[12];
        jacobian<JP>(r.gx, r.gy, r.lx, r.ly, J);
        static_for<NE>([&](auto e) {
            constexpr int E = decltype(e)::value;
            constexpr int i = tri_row(E0 + E, N), j = tri_col(E0 + E, N);
            // Translation uses the (gx, gy) pair only: indices 0 and 3.
            constexpr int ii = P == 2 ? 3 * i : i, jj = P == 2 ? 3 * j : j;
            acc[E] += J[ii] * J[jj] * r.w;
        });
        if (FIRST) {
            energy += r.f * r.f * r.w;
#pragma unroll
            for (int k = 0; k < JP; ++k) {
                if (P == 2 && k != 0 && k != 3) continue;
                jtf[k] += J[k] * r.f * r.w;
                sj[k] += J[k] * r.w;
            }
            if (a.D == a.S) {
                int row = p / a.S, col = p - row * a.S;
                a.reference[(size_t)(oy + row) * a.width + ox + col].x = (float)r.f;
            }
        }
    }
    static_for<NE>([&](auto e) {
        constexpr int E = decltype(e)::value;
        constexpr int i = tri_row(E0 + E, N), j = tri_col(E0 + E, N);
        const double v = wave_sum(acc[E]);
        if (lane == 0) {
            h[i * pitch + j] = v;
            h[j * pitch + i] = v;
        }
    });
    if (FIRST) {
        first->energy = wave_sum(energy);
#pragma unroll
        for (int k = 0; k < 12; ++k) first->jtf[k] = first->sj[k] = 0.0;
#pragma unroll
        for (int k = 0; k < JP; ++k) {
            first->jtf[k] = wave_sum(jtf[k]);
            first->sj[k] = wave_sum(sj[k]);
        }
    }
}

// Quadratic moments in one pixel pass.  Every Jacobian component is a gradient
// times one of the monomials {1, x, y, x^2/2, xy, y^2/2}, so each of the 78
// Hessian entries is c * sum w g_s g_t x^p y^q with p + q <= 4, c in {1, 1/2,
// 1/4} and (s, t) in {xx, xy, yy}: 45 distinct sums.  J^T f and sum J are the
// degree <= 2 sums of w f g_s and w g_s (24).  The monomials of the integer
// subset offsets and the power-of-two factors are exact in float64.
__host__ __device__ constexpr int mono_index(int p, int q) { return (p + q) * (p + q + 1) / 2 + q; }
// Component k of the quadratic Jacobian: gradient s(k), monomial exponents
// (px(k), py(k)) and factor c(k).
__host__ __device__ constexpr int jac_grad(int k) { return (k / 3) % 2; }
__host__ __device__ constexpr int jac_mono(int k) { return k % 3 + (k >= 6 ? 3 : 0); }
__host__ __device__ constexpr int mono_px(int u) { return u == 1 ? 1 : u == 3 ? 2 : u == 4 ? 1 : 0; }
__host__ __device__ constexpr int mono_py(int u) { return u == 2 ? 1 : u == 4 ? 1 : u == 5 ? 2 : 0; }
__host__ __device__ constexpr double mono_c(int u) { return (u == 3 || u == 5) ? 0.5 : 1.0; }

__device__ void quadratic_moments(const PrepArgs& a, int lane, int ox, int oy, double half, double mean, double* h,
                                  int pitch, FirstMoments* first) {
#pragma clang fp contract(off)
    double M[3][15], F[2][6], G[2][6], energy = 0.0;
#pragma unroll
    for (int m = 0; m < 15; ++m) M[0][m] = M[1][m] = M[2][m] = 0.0;
#pragma unroll
    for (int m = 0; m < 6; ++m) F[0][m] = F[1][m] = G[0][m] = G[1][m] = 0.0;
    const int SS = a.S * a.S;
    for (int p = lane; p < SS; p += kWave) {
        PixelRef r = pixel(a, p, ox, oy, half, mean);
        const double x = r.lx, y = r.ly;
        const double x2 = x * x, xy = x * y, y2 = y * y;
        const double mono[15] = {1.0,    x,      y,      x2,      xy,      y2,      x2 * x, x2 * y,
                                 x * y2, y2 * y, x2 * x2, x2 * xy, x2 * y2, xy * y2, y2 * y2};
        const double gg[3] = {r.gx * r.gx * r.w, r.gx * r.gy * r.w, r.gy * r.gy * r.w};
        const double wf = r.f * r.w;
        const double gf[2] = {r.gx * wf, r.gy * wf}, gw[2] = {r.gx * r.w, r.gy * r.w};
        energy += r.f * wf;
#pragma unroll
        for (int st = 0; st < 3; ++st) {
            M[st][0] += gg[st];
#pragma unroll
            for (int m = 1; m < 15; ++m) M[st][m] += gg[st] * mono[m];
        }
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            F[g][0] += gf[g];
            G[g][0] += gw[g];
#pragma unroll
            for (int m = 1; m < 6; ++m) {
                F[g][m] += gf[g] * mono[m];
                G[g][m] += gw[g] * mono[m];
            }
        }
        if (a.D == a.S) {
            int row = p / a.S, col = p - row * a.S;
            a.reference[(size_t)(oy + row) * a.width + ox + col].x = (float)r.f;
        }
    }
#pragma unroll
    for (int m = 0; m < 15; ++m) M[0][m] = wave_sum(M[0][m]), M[1][m] = wave_sum(M[1][m]), M[2][m] = wave_sum(M[2][m]);
    first->energy = wave_sum(energy);
    static_for<12>([&](auto kc) {
        constexpr int k = decltype(kc)::value;
        constexpr int u = jac_mono(k), g = jac_grad(k);
        constexpr int m = mono_index(mono_px(u), mono_py(u));
        first->jtf[k] = mono_c(u) * wave_sum(F[g][m]);
        first->sj[k] = mono_c(u) * wave_sum(G[g][m]);
    });
    static_for<78>([&](auto e) {
        constexpr int E = decltype(e)::value;
        constexpr int i = tri_row(E, 12), j = tri_col(E, 12);
        constexpr int ui = jac_mono(i), uj = jac_mono(j);
        constexpr int st = jac_grad(i) + jac_grad(j);
        constexpr int m = mono_index(mono_px(ui) + mono_px(uj), mono_py(ui) + mono_py(uj));
        constexpr double c = mono_c(ui) * mono_c(uj);
        if (lane == 0) {
            const double v = c * M[st][m];
            h[i * pitch + j] = v;
            h[j * pitch + i] = v;
        }
    });
}

// Column-parallel Gauss-Jordan elimination of the P x 2P augmented matrix
// [H | I] whose left half is in LDS (row pitch 2P), in diagonal order without
// row exchanges.  H is a sum of outer products J J^T, symmetric positive
// semi-definite, so its pivots (the diagonals of the successive Schur
// complements) are non-negative and the elimination is stable without
// pivoting, as a Cholesky factorisation is; partial pivoting would pick the
// rows of the large quadratic terms first and exchange rows at most steps.
// Lane e < 2P keeps column e in registers; each pivot column is broadcast with
// readlane, and every lane applies the division by the pivot and the row
// updates to its own column.  The result is written back to LDS.  Returns
// false for an exactly singular matrix (a zero pivot), the failure of the
// reference's LU inversion; a nearly singular one yields entries beyond the
// float32 range, which the caller rejects.
template <int P>
__device__ bool invert(double* m, int lane) {
    constexpr int C = 2 * P;
    const int e = lane < C ? lane : C - 1;
    double col[P];
#pragma unroll
    for (int r = 0; r < P; ++r) col[r] = e < P ? m[r * C + e] : (r == e - P ? 1.0 : 0.0);
    bool singular = false;
#pragma unroll
    for (int c = 0; c < P; ++c) {
        double pc[P];
#pragma unroll
        for (int r = 0; r < P; ++r) pc[r] = readlane_f64(col[r], c);
        if (pc[c] == 0.0) {
            singular = true;
            break;
        }
        const double top = col[c] / pc[c];
        col[c] = top;
#pragma unroll
        for (int r = 0; r < P; ++r)
            if (r != c) col[r] -= pc[r] * top;
    }
    wave_lds_sync();
    if (lane < C)
#pragma unroll
        for (int r = 0; r < P; ++r) m[r * C + lane] = col[r];
    wave_lds_sync();
    return !singular;
}

// Validity, inversion and outputs of one subset whose Hessian sits in the
// wave's LDS matrix `mat` (P x 2P, left half).
template <int MODEL>
__device__ void finish_subset(const PrepArgs& a, const int subset, const int lane, double* mat, const double energy,
                              const double* jtf, const double* sj, const double mean) {
#pragma clang fp contract(off)
    constexpr int P = MODEL == kQuadratic ? 12 : 6;
    constexpr int M = MODEL == kQuadratic ? 28 : 16;
    constexpr int C = 2 * P;
    bool valid;
    double t00 = 0.0, t03 = 0.0, t33 = 0.0;
    if (MODEL == kTranslation) {
        double hxx = mat[0], hxy = mat[1], hyy = mat[C + 1];
        double det = hxx * hyy - hxy * hxy;
        double scale = fmax(fabs(hxx * hyy), 1.0);
        valid = isfinite(det) && fabs(det) > DBL_EPSILON * scale && energy > 1e-12;
        double idet = valid ? 1.0 / det : 0.0;
        t00 = hyy * idet;
        t03 = -hxy * idet;
        t33 = hxx * idet;
        valid = valid && fmax(fmax(fabs(t00), fabs(t03)), fabs(t33)) <= (double)FLT_MAX;
    } else {
        // Every entry of the inverse must be finite and representable in float32;
        // lanes check the P*P entries in parallel.
        bool ok = invert<P>(mat, lane);
        bool bad = false;
        for (int k = lane; k < P * P; k += kWave) {
            const double v = mat[(k / P) * C + P + k % P];
            bad = bad || !isfinite(v) || fabs(v) > (double)FLT_MAX;
        }
        valid = ok && !__any(bad) && energy > 1e-12;
    }

    float* meta = a.metadata + (size_t)subset * M;
    float* invp = a.inverse + (size_t)subset * P * P;
    if (lane == 0) {
        float status = valid ? 1.0f : 0.0f;
        if (a.valid) a.valid[subset] = valid ? 1 : 0;
        if (a.selected && a.selected[subset] == 0) status = -3.0f;
        meta[0] = (float)sqrt(energy);
        meta[1] = status;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            meta[2 + k] = (float)jtf[k];
            meta[2 + P + k] = (float)sj[k];
        }
        meta[2 + 2 * P] = a.D == a.S ? 0.0f : (float)mean;
        meta[3 + 2 * P] = (float)a.weight_sum;
    }
    for (int k = lane; k < P * P; k += kWave) {
        double v;
        if (MODEL == kTranslation)
            v = k == 0 ? t00 : (k == 3 || k == 18) ? t03 : k == 21 ? t33 : 0.0;
        else
            v = mat[(k / P) * C + P + k % P];
        invp[k] = valid ? (float)v : 0.0f;
    }
}

template <int MODEL>
__global__ __launch_bounds__(256) void k_subset_prep(PrepArgs a) {
#pragma clang fp contract(off)
    constexpr int P = MODEL == kQuadratic ? 12 : 6;
    __shared__ double mats[kPrepWaves][P * 2 * P];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int subset = xcd_block(blockIdx.x, gridDim.x) * kPrepWaves + wave;
    const int N = a.grid_w * a.grid_h;
    if (subset >= N) return;
    double* mat = mats[wave];
    const int gx = subset % a.grid_w, gy = subset / a.grid_w;
    const int ox = gx * a.D, oy = gy * a.D;
    const double half = 0.5 * (double)(a.S - 1);
    const int SS = a.S * a.S;

    // Pass 0: weighted mean of the raw (or Gaussian-filtered) intensity.
    double s0 = 0.0;
    for (int p = lane; p < SS; p += kWave) {
        int row = p / a.S, col = p - row * a.S;
        double w = a.weighted ? a.weights[p] : 1.0;
        s0 += a.intensity[(size_t)(oy + row) * a.width + ox + col] * w;
    }
    const double mean = wave_sum(s0) / a.weight_sum;

    // Pass 1: energy, J^T f, sum J and the Hessian, written straight into the
    // augmented P x 2P layout used by the inversion.
    constexpr int C = 2 * P;
    FirstMoments fm;
    if (MODEL == kQuadratic) {
        quadratic_moments(a, lane, ox, oy, half, mean, mat, C, &fm);
    } else if (MODEL == kAffine) {
        moment_pass<6, 0, 21, true>(a, lane, ox, oy, half, mean, mat, C, &fm);
    } else {
        moment_pass<2, 0, 3, true>(a, lane, ox, oy, half, mean, mat, C, &fm);
    }
    const double energy = fm.energy;
    const double* jtf = fm.jtf;
    const double* sj = fm.sj;
    wave_lds_sync();

    finish_subset<MODEL>(a, subset, lane, mat, energy, jtf, sj, mean);
}

// Quadratic preparation by separable sums, for uniform weights and step < S.
// A block takes one subset column and up to kStripRows subset rows.
//  1. One thread per image row of the strip forms the row-segment sums over
//     the S columns of the subset column (x = c - (S-1)/2):
//       q  0..14  sum g_s g_t x^p    (s t = xx, xy, yy; p = 0..4)
//       q 15..26  sum I g_s x^p, sum g_s x^p   (p = 0..2)
//       q 27, 28  sum I, sum I^2
//  2. One thread per (subset, q) sums a quantity down the subset's S rows
//     times y^e for the powers the subset needs (e <= 4 - p, or 2 - p).
//  3. One wave per subset turns the 71 moments into the Hessian, J^T f and
//     sum J (J^T f = sum I J - mean sum J, energy = sum I^2 - mean sum I),
//     then inverts as the direct kernel does.
// Overlapping subsets (step 4, size 25) share the row sums, so the pixel work
// per subset drops from S^2 products of 70 sums to about 5.3 S row products.
constexpr int kStripRows = 16;
constexpr int kStripQ = 29;
constexpr int kStripMoments = 71;
constexpr int kStripMaxRows = 160;  // image rows per strip (LDS bound)

__host__ __device__ constexpr int strip_vpow(int q) { return q < 15 ? 4 - q % 5 : q < 27 ? 2 - (q - 15) % 3 : 0; }
// Offset of quantity q's first moment (vertical power 0).
__host__ __device__ constexpr int strip_offset(int q) {
    return q < 15 ? (q / 5) * 15 + 5 * (q % 5) - (q % 5) * (q % 5 - 1) / 2
                  : q < 27 ? 45 + ((q - 15) / 3) * 6 + 3 * ((q - 15) % 3) - ((q - 15) % 3) * ((q - 15) % 3 - 1) / 2
                           : 69 + (q - 27);
}
static_assert(strip_offset(28) + 1 == kStripMoments, "moment count");

__host__ __device__ inline int strip_subset_rows(int S, int D) {
    const int fit = (kStripMaxRows - S) / D + 1;
    return fit < 1 ? 1 : (fit > kStripRows ? kStripRows : fit);
}

// Inverts, in each 16-lane row, the 12 x 12 matrix whose column e sits in
// col[] of the row's lane e (e < 12).  Gauss-Jordan elimination in diagonal
// order, in place: at pivot c, DPP row_newbcast gives every lane of the row the
// pivot column; lane c, whose column has become the unit vector, turns into
// column c of the inverse, formed exactly as the augmented identity column c
// of [M | I] is in invert() (1 / pivot, then 0 - p_r * top), and every other
// lane applies the same update as there.  So each entry goes through the same
// operations as in invert(), without the columns that hold only the unit
// vectors.  Returns false when a pivot is zero.
template <int C>
__device__ __forceinline__ double row_bcast_f64(double v) {
    int2 p = *reinterpret_cast<int2*>(&v);
    p.x = __builtin_amdgcn_update_dpp(0, p.x, 0x150 + C, 0xF, 0xF, false);  // row_newbcast:C
    p.y = __builtin_amdgcn_update_dpp(0, p.y, 0x150 + C, 0xF, 0xF, false);
    return *reinterpret_cast<double*>(&p);
}

__device__ __forceinline__ bool strip_invert(double (&col)[12], const int e) {
    constexpr int P = 12;
    bool singular = false;
    static_for<P>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        double pc[P];
#pragma unroll
        for (int r = 0; r < P; ++r) pc[r] = row_bcast_f64<c>(col[r]);
        singular = singular || pc[c] == 0.0;
        const bool pivot = e == c;
        const double top = (pivot ? 1.0 : col[c]) / pc[c];
#pragma unroll
        for (int r = 0; r < P; ++r)
            if (r != c) col[r] = (pivot ? 0.0 : col[r]) - pc[r] * top;
        col[c] = top;
    });
    return !singular;
}

// Row-segment sums of one image row over the subset column's S pixels
// (x = c - (S-1)/2), stored at out[q * pitch]: part 0 the 15 sums g_s g_t x^p
// (quantities 0-14), part 1 the other 14 (15-28).  Every sum accumulates the same
// products in the same order as when one thread forms all 29.
template <int PART>
__device__ __forceinline__ void strip_row_sums(const PrepArgs& a, const size_t row, const int S, const double half,
                                               double* out, const int pitch) {
    constexpr int Q0 = PART == 0 ? 0 : 15, NQ = PART == 0 ? 15 : kStripQ - 15;
    double t[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) t[q] = 0.0;
    auto add = [&](double I, double gx_, double gy_, int c) {
        const double x = (double)c - half, x2 = x * x;
        if constexpr (PART == 0) {
            const double xp[5] = {1.0, x, x2, x2 * x, x2 * x2};
            const double gg[3] = {gx_ * gx_, gx_ * gy_, gy_ * gy_};
#pragma unroll
            for (int st = 0; st < 3; ++st)
#pragma unroll
                for (int p = 0; p < 5; ++p) t[st * 5 + p] += gg[st] * xp[p];
        } else {
            const double xp[3] = {1.0, x, x2};
            const double gi[2] = {gx_ * I, gy_ * I}, g[2] = {gx_, gy_};
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int p = 0; p < 3; ++p) {
                    t[k * 3 + p] += gi[k] * xp[p];
                    t[6 + k * 3 + p] += g[k] * xp[p];
                }
            t[12] += I;
            t[13] += I * I;
        }
    };
    // Five columns' loads are issued together before their products.
    constexpr int B = 5;
    int c = 0;
    for (; c + B <= S; c += B) {
        double I[B], gxv[B], gyv[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            I[b] = PART == 1 ? a.intensity[row + c + b] : 0.0;
            gxv[b] = a.grad_x[row + c + b];
            gyv[b] = a.grad_y[row + c + b];
        }
#pragma unroll
        for (int b = 0; b < B; ++b) add(I[b], gxv[b], gyv[b], c + b);
    }
    for (; c < S; ++c) add(PART == 1 ? a.intensity[row + c] : 0.0, a.grad_x[row + c], a.grad_y[row + c], c);
#pragma unroll
    for (int q = 0; q < NQ; ++q) out[(Q0 + q) * pitch] = t[q];
}

__global__ __launch_bounds__(256) void k_subset_prep_strips(PrepArgs a, int rows_per_block, int pitch) {
    constexpr int P = 12;
    extern __shared__ double T[];  // kStripQ x pitch row sums
    __shared__ double mom[kStripRows * kStripMoments];
    const int strips_y = (a.grid_h + rows_per_block - 1) / rows_per_block;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int gx = blk / strips_y, gy0 = (blk - gx * strips_y) * rows_per_block;
    if (gx >= a.grid_w) return;
    const int rows = min(rows_per_block, a.grid_h - gy0);
    const int S = a.S, D = a.D;
    const int R = (rows - 1) * D + S;
    const int ox = gx * D, oy0 = gy0 * D;
    const double half = 0.5 * (double)(S - 1);

    // 1. Row-segment sums, each image row by two threads: waves 0-1 form the 15
    //    sums g_s g_t x^p, waves 2-3 the other 14 (the split is wave-uniform).
    {
        const int part = threadIdx.x >> 7;
        for (int r = threadIdx.x & 127; r < R; r += 128) {
            const size_t row = (size_t)(oy0 + r) * a.width + ox;
            if (part == 0)
                strip_row_sums<0>(a, row, S, half, T + r, pitch);
            else
                strip_row_sums<1>(a, row, S, half, T + r, pitch);
        }
    }
    __syncthreads();

    // 2. Column sums per subset and quantity.
    for (int i = threadIdx.x; i < rows * kStripQ; i += blockDim.x) {
        const int j = i / kStripQ, q = i - j * kStripQ;
        const double* col = T + q * pitch + j * D;
        double m[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 5
        for (int r = 0; r < S; ++r) {
            const double v = col[r], y = (double)r - half, y2 = y * y;
            m[0] += v;
            m[1] += v * y;
            m[2] += v * y2;
            m[3] += v * (y2 * y);
            m[4] += v * (y2 * y2);
        }
        const int E = strip_vpow(q), o = j * kStripMoments + strip_offset(q);
#pragma unroll
        for (int e = 0; e < 5; ++e)
            if (e <= E) mom[o + e] = m[e];
    }
    __syncthreads();

    // 3. Four subsets per wave, one per 16-lane row: lane e < 12 of row g holds
    //    column e of subset j0 + g's Hessian in registers, built from the
    //    moments, and the rows invert their matrices side by side (strip_invert).
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int g = lane >> 4, e = lane & 15;
    const int ec = e < P ? e : P - 1;  // lanes 12-15 shadow column 11
    for (int j0 = wave * 4; j0 < rows; j0 += kPrepWaves * 4) {
        const int j = min(j0 + g, rows - 1);
        const double* m = mom + j * kStripMoments;
        const double sum_i = m[strip_offset(27)];
        const double mean = sum_i / a.weight_sum;
        const double energy = fmax(0.0, m[strip_offset(28)] - mean * sum_i);
        const int ue = jac_mono(ec), ge = jac_grad(ec);
        const double sg = m[strip_offset(21 + ge * 3 + mono_px(ue)) + mono_py(ue)];
        const double si = m[strip_offset(15 + ge * 3 + mono_px(ue)) + mono_py(ue)];
        const double jtf = mono_c(ue) * (si - mean * sg), sj = mono_c(ue) * sg;
        double col[P];
        static_for<P>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            constexpr int ur = jac_mono(r), gr = jac_grad(r);
            const int st = gr + ge, px = mono_px(ur) + mono_px(ue), py = mono_py(ur) + mono_py(ue);
            col[r] = mono_c(ur) * mono_c(ue) * m[strip_offset(st * 5 + px) + py];
        });
        const bool ok = strip_invert(col, e);
        bool bad = false;
#pragma unroll
        for (int r = 0; r < P; ++r) bad = bad || !isfinite(col[r]) || fabs(col[r]) > (double)FLT_MAX;
        const uint64_t bad_lanes = __ballot(bad && e < P);
        const bool valid = ok && ((bad_lanes >> (16 * g)) & 0xFFFull) == 0 && energy > 1e-12;
        if (j0 + g >= rows || e >= P) continue;
        const int subset = (gy0 + j) * a.grid_w + gx;
        float* meta = a.metadata + (size_t)subset * 28;
        float* invp = a.inverse + (size_t)subset * P * P;
        if (e == 0) {
            float status = valid ? 1.0f : 0.0f;
            if (a.valid) a.valid[subset] = valid ? 1 : 0;
            if (a.selected && a.selected[subset] == 0) status = -3.0f;
            meta[0] = (float)sqrt(energy);
            meta[1] = status;
            meta[2 + 2 * P] = a.D == a.S ? 0.0f : (float)mean;
            meta[3 + 2 * P] = (float)a.weight_sum;
        }
        meta[2 + e] = (float)jtf;
        meta[2 + P + e] = (float)sj;
#pragma unroll
        for (int r = 0; r < P; ++r) invp[r * P + e] = valid ? (float)col[r] : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// Partial-support statistics (roi_policy="partial"; the reference computes them on
// the host, live_dic/partial.py:26-130).  One wave per partial subset, float64, from
// the prepared float32 plane (f, gx, gy) over the support pixels only:
//   pass 1: weight sum and weighted mean of f;
//   pass 2: with fc = f - mean, the energy sum w fc^2, J^T fc, sum w J and the
//           Hessian sum w J J^T (J = [gx, gx x, gx y, gy, gy x, gy y]);
// then, on one lane, the admission gates: full rank of the Hessian scaled by
// (1, 1/S, 1/S, 1, 1/S, 1/S) (eigenvalues by cyclic Jacobi rotations; rank counts
// those above eps * P * the largest), its condition <= 1e8, weighted RMS contrast
// >= 1 and more support pixels than parameters; an admitted subset gets its inverse
// (closed form for translation, partially pivoted Gauss-Jordan for affine), which
// must be finite and representable in float32, and validity -4.
// ---------------------------------------------------------------------------
constexpr double kPartialMaxCondition = 1e8;
constexpr double kPartialMinContrast = 1.0;

__device__ void jacobi_eigenvalues(double (&a)[6][6], int n, double* eig) {
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += a[i][i] * a[i][i];
            for (int j = i + 1; j < n; ++j) off += a[i][j] * a[i][j];
        }
        if (off <= 1e-40 * diag || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                if (a[p][q] == 0.0) continue;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - sn * akq;
                    a[k][q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - sn * aqk;
                    a[q][k] = sn * apk + c * aqk;
                }
            }
    }
    for (int i = 0; i < n; ++i) eig[i] = a[i][i];
    for (int i = 1; i < n; ++i)  // ascending
        for (int j = i; j > 0 && eig[j] < eig[j - 1]; --j) {
            const double t = eig[j];
            eig[j] = eig[j - 1];
            eig[j - 1] = t;
        }
}

// In-place inverse of an n x n matrix by Gauss-Jordan with partial pivoting; false if singular.
__device__ bool gauss_jordan_inverse(double (&a)[6][6], int n, double (&inv)[6][6]) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) inv[i][j] = i == j ? 1.0 : 0.0;
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (fabs(a[r][col]) > fabs(a[piv][col])) piv = r;
        if (a[piv][col] == 0.0) return false;
        if (piv != col)
            for (int k = 0; k < n; ++k) {
                double t = a[col][k]; a[col][k] = a[piv][k]; a[piv][k] = t;
                t = inv[col][k]; inv[col][k] = inv[piv][k]; inv[piv][k] = t;
            }
        const double d = a[col][col];
        for (int k = 0; k < n; ++k) {
            a[col][k] /= d;
            inv[col][k] /= d;
        }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double f = a[r][col];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                a[r][k] -= f * a[col][k];
                inv[r][k] -= f * inv[col][k];
            }
        }
    }
    return true;
}

__global__ void __launch_bounds__(64) k_partial_stats(PartialStatsArgs a) {
    const int j = blockIdx.x, lane = threadIdx.x;
    if (j >= a.count) return;
    const int S = a.S, P = a.model == kTranslation ? 2 : 6;
    const uint32_t index = a.indices[j];
    const int x0 = (int)(index % a.grid_w) * a.D, y0 = (int)(index / a.grid_w) * a.D;
    const uint32_t* words = a.support + (size_t)j * ((S * S + 31) / 32);
    const double half = 0.5 * (S - 1);
    double w_sum = 0.0, wf_sum = 0.0, npix = 0.0;
    for (int p = lane; p < S * S; p += 64) {
        if (!((words[p >> 5] >> (p & 31)) & 1u)) continue;
        const double w = a.weights ? a.weights[p] : 1.0;
        const float4 r = a.reference[(size_t)(y0 + p / S) * a.width + x0 + p % S];
        w_sum += w;
        wf_sum += w * (double)r.x;
        npix += 1.0;
    }
    w_sum = wave_sum(w_sum);
    wf_sum = wave_sum(wf_sum);
    npix = wave_sum(npix);
    const double mean = wf_sum / w_sum;
    double energy = 0.0, jtf[6] = {0, 0, 0, 0, 0, 0}, sj[6] = {0, 0, 0, 0, 0, 0}, h[21];
    for (int k = 0; k < 21; ++k) h[k] = 0.0;
    for (int p = lane; p < S * S; p += 64) {
        if (!((words[p >> 5] >> (p & 31)) & 1u)) continue;
        const double w = a.weights ? a.weights[p] : 1.0;
        const float4 r = a.reference[(size_t)(y0 + p / S) * a.width + x0 + p % S];
        const double x = (p % S) - half, y = (p / S) - half;
        const double fc = (double)r.x - mean, gx = r.y, gy = r.z;
        const double jac[6] = {gx, gx * x, gx * y, gy, gy * x, gy * y};
        energy += w * fc * fc;
        int t = 0;
        for (int u = 0; u < 6; ++u) {
            jtf[u] += w * fc * jac[u];
            sj[u] += w * jac[u];
            for (int v = u; v < 6; ++v) h[t++] += w * jac[u] * jac[v];
        }
    }
    energy = wave_sum(energy);
    for (int u = 0; u < 6; ++u) {
        jtf[u] = wave_sum(jtf[u]);
        sj[u] = wave_sum(sj[u]);
    }
    for (int k = 0; k < 21; ++k) h[k] = wave_sum(h[k]);
    if (lane != 0) return;
    const int order_t[2] = {0, 3}, order_a[6] = {0, 1, 2, 3, 4, 5};
    const int* keep = P == 2 ? order_t : order_a;
    auto hess = [&](int u, int v) {
        if (u > v) { const int t = u; u = v; v = t; }
        return h[u * 6 - u * (u - 1) / 2 + (v - u)];
    };
    const double scale[6] = {1.0, 1.0 / S, 1.0 / S, 1.0, 1.0 / S, 1.0 / S};
    double m[6][6], eig[6];
    for (int r = 0; r < P; ++r)
        for (int c = 0; c < P; ++c) m[r][c] = hess(keep[r], keep[c]) * scale[keep[r]] * scale[keep[c]];
    jacobi_eigenvalues(m, P, eig);
    const double top = eig[P - 1];
    int rank = 0;
    for (int i = 0; i < P; ++i) rank += eig[i] > DBL_EPSILON * P * fmax(top, 0.0) ? 1 : 0;
    double condition = eig[0] > 0.0 ? top / eig[0] : 0.0;
    if (!isfinite(condition)) condition = 0.0;
    const double contrast = sqrt(energy / w_sum);
    bool ok = rank == P && condition <= kPartialMaxCondition && contrast >= kPartialMinContrast && npix > P;
    double inv[6][6];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) inv[r][c] = 0.0;
    if (ok) {
        if (P == 2) {
            const double h00 = hess(0, 0), h01 = hess(0, 3), h11 = hess(3, 3);
            const double det = h00 * h11 - h01 * h01;
            inv[0][0] = h11 / det;
            inv[0][3] = -h01 / det;
            inv[3][0] = -h01 / det;
            inv[3][3] = h00 / det;
        } else {
            double hm[6][6], hi[6][6];
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) hm[r][c] = hess(r, c);
            ok = gauss_jordan_inverse(hm, 6, hi);
            for (int r = 0; r < 6 && ok; ++r)
                for (int c = 0; c < 6; ++c) inv[r][c] = hi[r][c];
        }
        for (int r = 0; r < 6 && ok; ++r)
            for (int c = 0; c < 6; ++c)
                if (!isfinite(inv[r][c]) || fabs(inv[r][c]) > (double)FLT_MAX) ok = false;
        if (!ok)
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) inv[r][c] = 0.0;
    }
    float* row = a.metadata + (size_t)index * 16;
    row[0] = (float)sqrt(energy);
    row[1] = ok ? -4.0f : 0.0f;
    for (int u = 0; u < 6; ++u) {
        row[2 + u] = (float)jtf[u];
        row[8 + u] = (float)sj[u];
    }
    row[14] = (float)mean;
    row[15] = (float)w_sum;
    float* out = a.inverse + (size_t)index * 36;
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) out[r * 6 + c] = (float)inv[r][c];
    float* diag = a.diagnostics + (size_t)j * 4;
    diag[0] = ok ? 1.0f : 0.0f;
    diag[1] = (float)rank;
    diag[2] = (float)condition;
    diag[3] = (float)contrast;
}

hipError_t launch_partial_stats(const PartialStatsArgs& a, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    k_partial_stats<<<a.count, 64, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_subset_prep(const PrepArgs& a, hipStream_t s) {
    int N = a.grid_w * a.grid_h;
    if (N <= 0) return hipSuccess;
    if (a.model == kQuadratic && !a.weighted && a.D < a.S && a.S <= kStripMaxRows) {
        const int rows = strip_subset_rows(a.S, a.D);
        const int strips = a.grid_w * ((a.grid_h + rows - 1) / rows);
        const int pitch = (rows - 1) * a.D + a.S;
        k_subset_prep_strips<<<strips, 256, sizeof(double) * kStripQ * pitch, s>>>(a, rows, pitch);
        return hipGetLastError();
    }
    dim3 grid((N + kPrepWaves - 1) / kPrepWaves);
    switch (a.model) {
        case kAffine: k_subset_prep<kAffine><<<grid, 256, 0, s>>>(a); break;
        case kTranslation: k_subset_prep<kTranslation><<<grid, 256, 0, s>>>(a); break;
        case kQuadratic: k_subset_prep<kQuadratic><<<grid, 256, 0, s>>>(a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace ldic
