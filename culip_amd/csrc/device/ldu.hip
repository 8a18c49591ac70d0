// Triangular substitution over an imported sparse factor
// (ImportLDUPreconditioner, import_ldu_preconditioner.impl.cuh:197-330:
// csrsv with a non-unit diagonal, or its approximate Jacobi form) and the
// permutation / scaling steps of solve_PS (:414-484).
//
// Exact solves: one launch per dependency level; eight lanes per row read
// the row's entries coalesced, sum in FP64 with a fixed xor tree, and the
// first lane divides by the diagonal. Rows of a level depend only on rows
// of earlier levels, so x is updated in place.
#include "ldu.hpp"

#include <hip/hip_runtime.h>

namespace culip {

namespace {

constexpr int LPR = 8;   // lanes per row
constexpr int BLK = 256;

template<typename VT>
__global__ __launch_bounds__(BLK) void k_tri_level(int32_t count, const int32_t* __restrict__ rows,
    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci, const VT* __restrict__ cv,
    const int32_t* __restrict__ dpos, int lower, double* __restrict__ x)
{
    const int64_t tid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t slot = tid / LPR;
    const bool active = slot < count;
    const int32_t i = active ? rows[slot] : 0;
    double acc = 0.0;
    if(active) {
        const int32_t s = rp[i], e = rp[i + 1];
        for(int32_t q = s + sub; q < e; q += LPR) {
            const int32_t j = ci[q];
            if(lower ? (j < i) : (j > i)) acc = fma((double) cv[q], x[j], acc);
        }
    }
#pragma unroll
    for(int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
    if(active && sub == 0) {
        const int32_t p = dpos[i];
        const double d = p >= 0 ? (double) cv[p] : 1.0;
        x[i] = (x[i] - acc) / d;
    }
}

// x = dinv .* b
__global__ __launch_bounds__(BLK) void k_diag_scale(int32_t n, const double* __restrict__ dinv,
    const double* __restrict__ b, double* __restrict__ x)
{
    for(int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n;
        i += (int64_t) gridDim.x * blockDim.x)
        x[i] = dinv[i] * b[i];
}

// xo = (xi + dinv .* b) - dinv .* (M xi), M over the operand's triangle
// including its diagonal
template<typename VT>
__global__ __launch_bounds__(BLK) void k_tri_jacobi(int32_t n, const int32_t* __restrict__ rp,
    const int32_t* __restrict__ ci, const VT* __restrict__ cv, int lower,
    const double* __restrict__ dinv, const double* __restrict__ b, const double* __restrict__ xi,
    double* __restrict__ xo)
{
    const int64_t tid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t row0 = tid / LPR;
    const int64_t stride = (int64_t) gridDim.x * blockDim.x / LPR;
    for(int64_t row = row0; row < n; row += stride) {
        const int32_t i = (int32_t) row;
        double acc = 0.0;
        const int32_t s = rp[i], e = rp[i + 1];
        for(int32_t q = s + sub; q < e; q += LPR) {
            const int32_t j = ci[q];
            if(lower ? (j <= i) : (j >= i)) acc = fma((double) cv[q], xi[j], acc);
        }
#pragma unroll
        for(int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
        if(sub == 0) xo[i] = (xi[i] + dinv[i] * b[i]) - dinv[i] * acc;
    }
}

__global__ __launch_bounds__(BLK) void k_permute_scale(int32_t n, const int32_t* __restrict__ perm,
    const double* __restrict__ scale, int push, const double* __restrict__ b,
    double* __restrict__ x)
{
    for(int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n;
        t += (int64_t) gridDim.x * blockDim.x) {
        const int32_t i = (int32_t) t;
        const int32_t pi = perm ? perm[i] : i;
        if(push) {
            // x = s .* (P b): b[i] moves to pi, then row pi is scaled
            x[pi] = scale ? scale[pi] * b[i] : b[i];
        } else {
            // x = P' (s .* b): row i pulls the scaled entry pi
            x[i] = scale ? scale[pi] * b[pi] : b[pi];
        }
    }
}

int grid_for(int64_t threads)
{
    int64_t g = (threads + BLK - 1) / BLK;
    if(g < 1) g = 1;
    if(g > 65535) g = 65535;
    return (int) g;
}

}  // namespace

template<typename VT>
void tri_solve_levels(const TriOperand<VT>& M, const int32_t* lvl_off, int32_t levels, double* x,
    void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    for(int32_t l = 0; l < levels; ++l) {
        const int32_t count = lvl_off[l + 1] - lvl_off[l];
        if(count <= 0) continue;
        const int grid = (int) (((int64_t) count * LPR + BLK - 1) / BLK);
        k_tri_level<VT><<<grid, BLK, 0, st>>>(count, M.order + lvl_off[l], M.row_ptr, M.col,
            M.val, M.diag_pos, M.lower ? 1 : 0, x);
    }
}

template<typename VT>
void tri_solve_jacobi(const TriOperand<VT>& M, const double* b, double* x, double* tmp,
    int32_t sweeps, void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    // start in the buffer that makes the last sweep land in x
    double* cur = (sweeps & 1) ? tmp : x;
    double* nxt = (sweeps & 1) ? x : tmp;
    k_diag_scale<<<grid_for(M.n), BLK, 0, st>>>(M.n, M.inv_diag, b, cur);
    for(int32_t s = 0; s < sweeps; ++s) {
        k_tri_jacobi<VT><<<grid_for((int64_t) M.n * LPR), BLK, 0, st>>>(M.n, M.row_ptr, M.col,
            M.val, M.lower ? 1 : 0, M.inv_diag, b, cur, nxt);
        double* t = cur;
        cur = nxt;
        nxt = t;
    }
}

void permute_scale(int32_t n, const int32_t* perm, const double* scale, bool push,
    const double* b, double* x, void* stream)
{
    k_permute_scale<<<grid_for(n), BLK, 0, (hipStream_t) stream>>>(n, perm, scale, push ? 1 : 0,
        b, x);
}

template void tri_solve_levels<double>(const TriOperand<double>&, const int32_t*, int32_t,
    double*, void*);
template void tri_solve_levels<float>(const TriOperand<float>&, const int32_t*, int32_t, double*,
    void*);
template void tri_solve_jacobi<double>(const TriOperand<double>&, const double*, double*,
    double*, int32_t, void*);
template void tri_solve_jacobi<float>(const TriOperand<float>&, const double*, double*, double*,
    int32_t, void*);

}  // namespace culip
