// Wave64 helpers for gfx950 (CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace culip {

constexpr int WAVE = 64;

// Orders LDS traffic between the lanes of one wave: the wave's DS operations
// execute in program order, so a compiler fence plus a wave barrier is enough
// for write-then-read across lanes of the same wave.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int lane_id() { return (int) (threadIdx.x & (WAVE - 1)); }

// number of set lanes below this one in mask
__device__ __forceinline__ int lanes_below(uint64_t mask)
{
    return (int) __builtin_amdgcn_mbcnt_hi((uint32_t) (mask >> 32),
        __builtin_amdgcn_mbcnt_lo((uint32_t) mask, 0u));
}

__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

template<typename T>
__device__ __forceinline__ T wave_bcast(T v, int src) { return __shfl(v, src, WAVE); }

template<typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
    for(int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

template<typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
    for(int o = WAVE / 2; o > 0; o >>= 1) { T u = __shfl_xor(v, o, WAVE); v = v > u ? v : u; }
    return v;
}

}  // namespace culip
