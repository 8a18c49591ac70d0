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

// Ordered scans over the threads of a workgroup for an associative (not
// necessarily commutative) operation on a value type V with members that
// __shfl can move. Op: V op(const V& earlier, const V& later), id: identity.
// Returns the exclusive prefix of this thread; *total receives the
// combination over all threads. NW = waves per workgroup; sh holds NW values.
template<typename V, typename Op, int NW>
__device__ __forceinline__ V block_exclusive_scan(V v, Op op, const V& id, V* sh, V* total)
{
    const int lane = lane_id();
    const int wv = (int) (threadIdx.x / WAVE);
    V inc = v;
#pragma unroll
    for(int o = 1; o < WAVE; o <<= 1) {
        V u = V::shfl_up(inc, o);
        if(lane >= o) inc = op(u, inc);
    }
    if(lane == WAVE - 1) sh[wv] = inc;
    __syncthreads();
    V pre = id;
    for(int q = 0; q < wv; ++q) pre = op(pre, sh[q]);
    V all = id;
#pragma unroll
    for(int q = 0; q < NW; ++q) all = op(all, sh[q]);
    V ex = V::shfl_up(inc, 1);
    if(lane == 0) ex = id;
    __syncthreads();
    if(total) *total = all;
    return op(pre, ex);
}

}  // namespace culip
