// Device-side helpers shared by the live-dic gfx950 kernels.
//
// Wave64 reductions use DPP row operations inside each 16-lane row and
// v_readlane across the four rows, so a subset's sums end up wave-uniform
// without LDS traffic or barriers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace ldic {

constexpr int kWave = 64;

// DPP controls (GCN/CDNA encoding).
constexpr int kQuadXor1 = 0xB1;      // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;      // quad_perm [2,3,0,1]
constexpr int kRowHalfMirror = 0x141;
constexpr int kRowMirror = 0x140;
constexpr int kRowBcast15 = 0x142;   // lane 15 of each row to the next row
constexpr int kRowBcast31 = 0x143;   // lane 31 to rows 2 and 3

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// DPP move restricted to the rows in ROWS; the other rows read 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_rows_f32(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    int2 p = *reinterpret_cast<int2*>(&v);
    p.x = __builtin_amdgcn_update_dpp(0, p.x, CTRL, 0xF, 0xF, false);
    p.y = __builtin_amdgcn_update_dpp(0, p.y, CTRL, 0xF, 0xF, false);
    return *reinterpret_cast<double*>(&p);
}

__device__ __forceinline__ float readlane_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    int2 p = *reinterpret_cast<int2*>(&v);
    p.x = __builtin_amdgcn_readlane(p.x, lane);
    p.y = __builtin_amdgcn_readlane(p.y, lane);
    return *reinterpret_cast<double*>(&p);
}

// Sum over the 64 lanes; every lane receives the same bits.  Within a row the
// mirrored partners hold identical partial sums, so the order is symmetric.
__device__ __forceinline__ float wave_sum(float v) {
    v += dpp_f32<kQuadXor1>(v);
    v += dpp_f32<kQuadXor2>(v);
    v += dpp_f32<kRowHalfMirror>(v);
    v += dpp_f32<kRowMirror>(v);
    return (readlane_f32(v, 0) + readlane_f32(v, 16)) + (readlane_f32(v, 32) + readlane_f32(v, 48));
}

// The same sum, with the cross-row step done by row broadcasts: rows 1 and 3
// add rows 0 and 2, then rows 2 and 3 add lane 31, and lane 63 holds
// (r3 + r2) + (r1 + r0), bit-identical to wave_sum because float addition is
// commutative.  It compiles to ten VALU operations and one lane read (each
// broadcast is a zeroed DPP move and an add); wave_sums below forms the same
// sums for several values at once.
__device__ __forceinline__ float wave_sum_bcast(float v) {
    v += dpp_f32<kQuadXor1>(v);
    v += dpp_f32<kQuadXor2>(v);
    v += dpp_f32<kRowHalfMirror>(v);
    v += dpp_f32<kRowMirror>(v);
    v += dpp_rows_f32<kRowBcast15, 0xA>(v);
    v += dpp_rows_f32<kRowBcast31, 0xC>(v);
    return readlane_f32(v, 63);
}

// Sums of several values at once (wave_sums below).  One step of the row
// reduction for two values a and b: a lane keeps a when h is false, b when h
// is true, and adds its DPP partner's copy of the same value, which the partner
// does not keep (h differs between partners).  Both values are thus summed
// over the pair of lanes with one DPP add, and a lane carries one of them on.
template <int CTRL>
__device__ __forceinline__ float dpp_pair_sum(float a, float b, bool h) {
    const float keep = h ? b : a, give = h ? a : b;
    return keep + dpp_f32<CTRL>(give);
}

// Row pairs (0, 1) and (2, 3): rows 0 and 2 of the result hold x's two-row
// sums, rows 1 and 3 y's (v_permlane16_swap exchanges x's odd rows with y's
// even rows).
__device__ __forceinline__ float rows16_pair_sum(float x, float y) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Rows (0 + 2, 1 + 3): rows 0 and 1 of the result hold x's sums, rows 2 and 3
// y's (v_permlane32_swap exchanges x's rows 2-3 with y's rows 0-1).
__device__ __forceinline__ float rows32_pair_sum(float x, float y) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// Row reduction of a group of 2^K values (K <= 3).  The lane selectors are
// h1 = b0 ^ b2, h2 = b1 ^ b2, h3 = b2 ^ b3 of the lane index: each flips
// between the partners of its own step (lane ^ 1, ^ 2, ^ 7) and is unchanged
// between the partners of every later step (^ 2, ^ 7, ^ 15, and across rows),
// so partners always hold the same value.  A lane ends with value
// 4 h3 + 2 h2 + h1 of the group, summed over its row.
template <int K>
__device__ __forceinline__ float row_group_sum(const float* v, bool h1, bool h2, bool h3) {
    float z;
    if constexpr (K == 0) {
        z = v[0];
        z += dpp_f32<kQuadXor1>(z);
        z += dpp_f32<kQuadXor2>(z);
    } else if constexpr (K == 1) {
        z = dpp_pair_sum<kQuadXor1>(v[0], v[1], h1);
        z += dpp_f32<kQuadXor2>(z);
    } else {
        float w[1 << (K - 1)];
#pragma unroll
        for (int a = 0; a < (1 << (K - 1)); ++a) w[a] = dpp_pair_sum<kQuadXor1>(v[2 * a], v[2 * a + 1], h1);
        if constexpr (K == 2) {
            z = dpp_pair_sum<kQuadXor2>(w[0], w[1], h2);
        } else {
            const float x0 = dpp_pair_sum<kQuadXor2>(w[0], w[1], h2), x1 = dpp_pair_sum<kQuadXor2>(w[2], w[3], h2);
            z = dpp_pair_sum<kRowHalfMirror>(x0, x1, h3);
            z += dpp_f32<kRowMirror>(z);
            return z;
        }
    }
    z += dpp_f32<kRowHalfMirror>(z);
    z += dpp_f32<kRowMirror>(z);
    return z;
}

// Lane of row 0 holding value j of a group of 2^K after row_group_sum: the
// lane with b2 = 0 and (h1, h2, h3) = the bits of j.
template <int K>
__host__ __device__ constexpr int row_group_lane(int j) {
    return K == 0 ? 0 : K == 1 ? (j & 1) : K == 2 ? (j & 3) : ((j & 3) | ((j & 4) << 1));
}

// The N (< 16) values v[] summed over the wave, in place (wave-uniform).  Each
// sum is bit-identical to wave_sum_bcast of that value: every addition joins
// the same two partial sums (lanes ^ 1, ^ 2, mirrored 8 and 16 lanes, then
// rows 0 + 1 and 2 + 3, then the two halves), since float addition is
// commutative.  Values go in groups of 8, 4, 2 and 1 (the binary digits of N),
// each reduced within rows by row_group_sum, which halves the number of values
// per step; the groups' results are then joined across rows by the permlane
// swaps, and group g ends in row g.  Nine values take 31 VALU operations and 9
// lane reads instead of 90 and 9.  Lane is the lane index (0-63).
template <int N>
__device__ __forceinline__ void wave_sums(float (&v)[N], int lane) {
    static_assert(N >= 1 && N < 16, "1 to 15 values");
    const bool h1 = ((lane ^ (lane >> 2)) & 1) != 0;
    const bool h2 = (((lane >> 1) ^ (lane >> 2)) & 1) != 0;
    const bool h3 = (((lane >> 2) ^ (lane >> 3)) & 1) != 0;
    constexpr int o4 = N & 8, o2 = o4 + (N & 4), o1 = o2 + (N & 2);
    constexpr int G = ((N >> 3) & 1) + ((N >> 2) & 1) + ((N >> 1) & 1) + (N & 1);
    float z[4];
    int g = 0;
    if constexpr ((N & 8) != 0) z[g++] = row_group_sum<3>(v, h1, h2, h3);
    if constexpr ((N & 4) != 0) z[g++] = row_group_sum<2>(v + o4, h1, h2, h3);
    if constexpr ((N & 2) != 0) z[g++] = row_group_sum<1>(v + o2, h1, h2, h3);
    if constexpr ((N & 1) != 0) z[g++] = row_group_sum<0>(v + o1, h1, h2, h3);
    float f;
    if constexpr (G == 1) {
        const float r = rows16_pair_sum(z[0], z[0]);
        f = rows32_pair_sum(r, r);
    } else if constexpr (G == 2) {
        const float r = rows16_pair_sum(z[0], z[1]);
        f = rows32_pair_sum(r, r);
    } else if constexpr (G == 3) {
        f = rows32_pair_sum(rows16_pair_sum(z[0], z[1]), rows16_pair_sum(z[2], z[2]));
    } else {
        f = rows32_pair_sum(rows16_pair_sum(z[0], z[1]), rows16_pair_sum(z[2], z[3]));
    }
    // Group g (in the order 8, 4, 2, 1) holds its totals in row g.
    g = 0;
    if constexpr ((N & 8) != 0) {
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = readlane_f32(f, 16 * g + row_group_lane<3>(j));
        ++g;
    }
    if constexpr ((N & 4) != 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) v[o4 + j] = readlane_f32(f, 16 * g + row_group_lane<2>(j));
        ++g;
    }
    if constexpr ((N & 2) != 0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) v[o2 + j] = readlane_f32(f, 16 * g + row_group_lane<1>(j));
        ++g;
    }
    if constexpr ((N & 1) != 0) v[o1] = readlane_f32(f, 16 * g);
}

__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_f64<kQuadXor1>(v);
    v += dpp_f64<kQuadXor2>(v);
    v += dpp_f64<kRowHalfMirror>(v);
    v += dpp_f64<kRowMirror>(v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

__device__ __forceinline__ float wave_min(float v) {
    v = fminf(v, dpp_f32<kQuadXor1>(v));
    v = fminf(v, dpp_f32<kQuadXor2>(v));
    v = fminf(v, dpp_f32<kRowHalfMirror>(v));
    v = fminf(v, dpp_f32<kRowMirror>(v));
    return fminf(fminf(readlane_f32(v, 0), readlane_f32(v, 16)), fminf(readlane_f32(v, 32), readlane_f32(v, 48)));
}

__device__ __forceinline__ float wave_max(float v) {
    v = fmaxf(v, dpp_f32<kQuadXor1>(v));
    v = fmaxf(v, dpp_f32<kQuadXor2>(v));
    v = fmaxf(v, dpp_f32<kRowHalfMirror>(v));
    v = fmaxf(v, dpp_f32<kRowMirror>(v));
    return fmaxf(fmaxf(readlane_f32(v, 0), readlane_f32(v, 16)), fmaxf(readlane_f32(v, 32), readlane_f32(v, 48)));
}

// Orders LDS writes of one wave before its later LDS reads.  Waves of a block
// work on different subsets, so no block barrier is involved.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroups are dealt to the 8 XCDs round-robin.  Give each XCD one
// contiguous run of logical blocks so neighbouring subsets, which overlap in
// the reference and target planes, share that XCD's L2.
__device__ __forceinline__ int xcd_block(int block, int count) {
    const int xcds = 8;
    int xcd = block % xcds;
    int local = block / xcds;
    int per = count / xcds;
    int extra = count % xcds;
    return xcd * per + (xcd < extra ? xcd : extra) + local;
}

// Single reflection without edge repeat, then the halo clamp
// (shaders/icgn.wgsl:128-143, shaders/full_pyramid.wgsl:150-158).
__device__ __forceinline__ int mirror_clamp(int c, int size) {
    int r = c < 0 ? -c : (c >= size ? 2 * size - 2 - c : c);
    return min(max(r, 0), size - 1);
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// Loads from device memory by a 32-bit byte offset from a wave-uniform base,
// in the global address space: a pointer read from a kernel-argument array is
// generic to the compiler, and a flat load also counts against the LDS wait
// counter, so every wait for an LDS read would wait for it too.  The base
// stays in scalar registers and the offset is the only per-lane address.
typedef __attribute__((address_space(1))) const char global_byte;

template <typename T>
__device__ __forceinline__ T load_global(const void* base, uint32_t byte_offset) {
    const global_byte* g = (const global_byte*)base + byte_offset;
    if constexpr (std::is_same<T, float3>::value) {
        // float3 is a class type: its bits are read through the builtin
        // three-element vector, one dwordx3 load (the first three floats of a
        // float4 element, whose fourth is padding).
        typedef float v3f __attribute__((ext_vector_type(3)));
        const v3f v = *(__attribute__((address_space(1))) const v3f*)g;
        return make_float3(v.x, v.y, v.z);
    } else {
        return *(__attribute__((address_space(1))) const T*)g;
    }
}

}  // namespace ldic
