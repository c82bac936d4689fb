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
// commutative.  Seven VALU operations instead of eleven.
__device__ __forceinline__ float wave_sum_bcast(float v) {
    v += dpp_f32<kQuadXor1>(v);
    v += dpp_f32<kQuadXor2>(v);
    v += dpp_f32<kRowHalfMirror>(v);
    v += dpp_f32<kRowMirror>(v);
    v += dpp_rows_f32<kRowBcast15, 0xA>(v);
    v += dpp_rows_f32<kRowBcast31, 0xC>(v);
    return readlane_f32(v, 63);
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
