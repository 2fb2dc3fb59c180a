// etraj_device.hpp -- wavefront helpers shared by the gfx950 kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "etraj_engine.hpp"

namespace etraj {

__device__ inline int lane_id() { return (int)(threadIdx.x & 63); }
// this workgroup's copy of the per-job scalar sums and counters (kAccSlots)
__device__ inline int acc_slot() { return (int)(blockIdx.x & (kAccSlots - 1)); }

__device__ inline double readlane_f64(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ inline float readlane_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ inline double uniform_f64(double v) {
    int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
// A per-segment row of the field table as a buffer resource: a load through
// it takes a wave-uniform byte offset in a scalar register and returns the
// values in vector registers, with no vector address arithmetic (a plain load
// of a uniform address would land in scalar registers, and every VALU operand
// would then need a v_mov). Reads at offsets of `bytes` or more return zeros.
// Word 3 = 0x00020000: the gfx9 untyped-buffer setting (32-bit data format).
__device__ inline __amdgpu_buffer_rsrc_t row_rsrc(const void* base, int64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ inline float4 row_load_f4(__amdgpu_buffer_rsrc_t rs, int byte_off) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rs, 0, byte_off, 0));
}
__device__ inline double2 row_load_d2(__amdgpu_buffer_rsrc_t rs, int byte_off) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rs, 0, byte_off, 0));
}

// The segment holding item (0-based over all electrons of all segments):
// pre[s] <= item < pre[s + 1] (the last such s where segments are empty).
__device__ inline int64_t item_segment(int64_t item, const int64_t* pre, int64_t n_seg) {
    int64_t lo = 0, hi = n_seg;
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (pre[mid] <= item) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ inline int64_t uniform_i64(int64_t v) {
    uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ inline double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline int64_t wave_sum_i64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Adds a fixed-point value to the two words slot[0] (lo), slot[1] (hi): the
// carry out of the low word is known from the value the atomic returns, so
// the 128-bit sum is exact whatever the order of the additions.
__device__ inline void fix_atomic_add(uint64_t* slot, Fix128 v) {
    if ((v.lo | v.hi) == 0) return;
    const uint64_t old = atomicAdd((unsigned long long*)slot, (unsigned long long)v.lo);
    const uint64_t hi = v.hi + ((old + v.lo < old) ? 1 : 0);
    if (hi) atomicAdd((unsigned long long*)(slot + 1), (unsigned long long)hi);
}

// 128-bit integer sum over the wavefront (exact)
__device__ inline Fix128 fix_wave_sum(Fix128 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t lo = (uint64_t)__shfl_xor((long long)v.lo, o, 64);
        const uint64_t hi = (uint64_t)__shfl_xor((long long)v.hi, o, 64);
        const uint64_t s = v.lo + lo;
        v.hi = v.hi + hi + (s < v.lo ? 1 : 0);
        v.lo = s;
    }
    return v;
}

}  // namespace etraj
