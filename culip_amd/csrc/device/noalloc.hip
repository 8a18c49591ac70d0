// Device memory of the allocation-free entry points: one static array per
// GPU, so that a call allocates nothing (callers under graph capture, in
// tight serving loops, or that may not allocate).
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.hpp"

namespace culip {

// Work vectors of culip_bildlt_solve_noalloc: the right-hand side and the
// two iterates of block-Jacobi sweeps, NOALLOC_SCRATCH_VECTORS vectors of
// NOALLOC_MAX_ROWS doubles (24 MiB), held by every GPU the library runs on
// until the process ends. Host code must not pass this name to a kernel or
// a copy: its device address comes from noalloc_scratch().
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_VECTORS * NOALLOC_MAX_ROWS];

double* noalloc_scratch()
{
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static double* addr[MAX_DEVICES] = {};
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if(!addr[dev]) {
        // looked up once per device, with that device current: each GPU
        // holds its own copy of the array at its own address
        void* p = nullptr;
        if(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        addr[dev] = static_cast<double*>(p);
    }
    return addr[dev];
}

}  // namespace culip
