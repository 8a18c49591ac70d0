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

// DPP lane moves (GFX9 data-parallel primitives, one VALU op per 32 bits):
// lane i receives src of the lane selected by CTRL within the rows enabled in
// ROW_MASK; lanes without a source, or in disabled rows, receive old.
// CTRL: 0x111 + k - 1 = row_shr:k (from lane i - k of the same 16-lane row),
// 0x142 = row_bcast:15 (lane 15 of the row before), 0x143 = row_bcast:31
// (lane 31), 0x130 = wave_shl:1 (from lane i + 1).
template<int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_move(int old, int src)
{
    return __builtin_amdgcn_update_dpp(old, src, CTRL, ROW_MASK, 0xf, false);
}

template<int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double old, double src)
{
    const int hi = dpp_move<CTRL, ROW_MASK>(__double2hiint(old), __double2hiint(src));
    const int lo = dpp_move<CTRL, ROW_MASK>(__double2loint(old), __double2loint(src));
    return __hiloint2double(hi, lo);
}

template<int CTRL, int ROW_MASK>
__device__ __forceinline__ void seg_scan_step(double& v, int sg)
{
    const double u = dpp_move<CTRL, ROW_MASK>(0.0, v);
    const int su = dpp_move<CTRL, ROW_MASK>(-1, sg);
    v = (su == sg) ? v + u : v;
}

// Segmented inclusive sum over the 64 lanes of a wave, all lanes active:
// lane i returns the sum of v over the lanes j <= i with sg_j == sg_i. sg
// must be non-decreasing over the lanes (segments are runs) and >= 0. The
// sums run in a fixed tree: within 16-lane rows by distances 1, 2, 4, 8,
// then across rows through lanes 15 and 31.
__device__ __forceinline__ double seg_scan_runs(double v, int sg)
{
    seg_scan_step<0x111, 0xf>(v, sg);
    seg_scan_step<0x112, 0xf>(v, sg);
    seg_scan_step<0x114, 0xf>(v, sg);
    seg_scan_step<0x118, 0xf>(v, sg);
    seg_scan_step<0x142, 0xa>(v, sg);
    seg_scan_step<0x143, 0xc>(v, sg);
    return v;
}

// seg_scan_runs(a * b, sg) with the first step's sum written out as
// fma(a, b, u), u the neighbour's rounded product: each lane's own product
// enters its first partial sum unrounded. Every kernel that sums products
// through this rounds them identically, independently of what the compiler
// would contract.
__device__ __forceinline__ double seg_scan_runs_fma(double a, double b, int sg)
{
    const double v0 = a * b;
    const double u = dpp_move<0x111, 0xf>(0.0, v0);
    const int su = dpp_move<0x111, 0xf>(-1, sg);
    double v = (su == sg) ? fma(a, b, u) : v0;
    seg_scan_step<0x112, 0xf>(v, sg);
    seg_scan_step<0x114, 0xf>(v, sg);
    seg_scan_step<0x118, 0xf>(v, sg);
    seg_scan_step<0x142, 0xa>(v, sg);
    seg_scan_step<0x143, 0xc>(v, sg);
    return v;
}

// v of lane i + 1 (fill on lane 63), all lanes active
__device__ __forceinline__ int next_lane(int v, int fill) { return dpp_move<0x130, 0xf>(fill, v); }

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
