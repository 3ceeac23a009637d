// Cross-lane and cross-wave sums for the permanent kernels.
//
// Within a 64-lane wavefront: DPP quad permutes for lane distances 1 and 2,
// ds_swizzle (xor mode, inside 32-lane halves) for 4, 8 and 16, and
// v_readlane for the two halves.  Across the waves of a workgroup: one LDS
// slot per wave.  No atomics anywhere; the per-block results are combined by
// a separate tree pass.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace phip {

__device__ __forceinline__ uint32_t xor_lanes_u32(uint32_t v, int dist) {
    switch (dist) {
        case 1:  // quad_perm [1,0,3,2]
            return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
        case 2:  // quad_perm [2,3,0,1]
            return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
        case 4:
            return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (4 << 10) | 0x1F);
        case 8:
            return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (8 << 10) | 0x1F);
        default:  // 16
            return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (16 << 10) | 0x1F);
    }
}

__device__ __forceinline__ uint64_t xor_lanes_u64(uint64_t v, int dist) {
    uint32_t lo = xor_lanes_u32((uint32_t)v, dist);
    uint32_t hi = xor_lanes_u32((uint32_t)(v >> 32), dist);
    return ((uint64_t)hi << 32) | lo;
}

// Sum over the wavefront; the result is uniform (returned to every lane).
__device__ __forceinline__ uint64_t wave_reduce_sum_u64(uint64_t v) {
    v += xor_lanes_u64(v, 1);
    v += xor_lanes_u64(v, 2);
    v += xor_lanes_u64(v, 4);
    v += xor_lanes_u64(v, 8);
    v += xor_lanes_u64(v, 16);
    uint32_t lo0 = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 0);
    uint32_t hi0 = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 0);
    uint32_t lo1 = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 32);
    uint32_t hi1 = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 32);
    return (((uint64_t)hi0 << 32) | lo0) + (((uint64_t)hi1 << 32) | lo1);
}

// Sum over a workgroup of BLOCK threads (a multiple of 64); the result is
// valid in thread 0.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_reduce_sum_u64(uint64_t v) {
    static_assert(BLOCK % 64 == 0, "block must be whole wavefronts");
    constexpr int W = BLOCK / 64;
    __shared__ uint64_t slots[W];
    const uint64_t w = wave_reduce_sum_u64(v);
    if constexpr (W == 1) {
        return w;
    } else {
        if ((threadIdx.x & 63) == 0) slots[threadIdx.x >> 6] = w;
        __syncthreads();
        uint64_t s = 0;
        if (threadIdx.x == 0) {
#pragma unroll
            for (int i = 0; i < W; ++i) s += slots[i];
        }
        return s;
    }
}

}  // namespace phip
