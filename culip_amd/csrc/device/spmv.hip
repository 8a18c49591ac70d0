// CSR SpMV (cuSPARSE csrmv semantics, sparse_la_instances.cu:130-232), the
// matrix-free KKT operator (AugmentedMatrixSPMV, spmv.impl.cuh:362-427) and
// the normal-equations operator (NormalMatrixSPMV, spmv.impl.cuh:122-265).
// y = alpha * A x + beta * y. Eight lanes per row, coalesced entry loads,
// fixed-order reduction (xor tree within the eight lanes).
#include "spmv.hpp"
#include "batch.hpp"
#include "wave.hpp"

#include <hip/hip_runtime.h>

namespace culip {

namespace {

constexpr int LPR = 8;   // lanes per row
constexpr int SPMV_U = 3; // entries per lane and round of a CSR row
// the same for the batched product, whose lane groups hold four rows: two
// entries per row and round keep it at six waves per SIMD (three: four)
constexpr int BATCH_SPMV_U = 2;

// the sum of one CSR row over the eight lanes sub = 0..7 of a group (every
// lane returns it)
__device__ __forceinline__ double csr_row_sum(const int32_t* __restrict__ rp,
    const int32_t* __restrict__ ci, const double* __restrict__ cv, const double* __restrict__ x,
    int64_t row, int sub)
{
    double acc = 0.0;
    const int32_t s = rp[row], e = rp[row + 1];
    // the lane's entries q0, q0 + LPR, q0 + 2 LPR are loaded together
    // (clamped in range), then summed in that order
    for(int32_t q0 = s + sub; q0 < e; q0 += LPR * SPMV_U) {
        int32_t c[SPMV_U];
        double v[SPMV_U], xv[SPMV_U];
#pragma unroll
        for(int u = 0; u < SPMV_U; ++u) {
            const int32_t q = min(q0 + u * LPR, e - 1);
            c[u] = ci[q];
            v[u] = cv[q];
        }
#pragma unroll
        for(int u = 0; u < SPMV_U; ++u) xv[u] = x[c[u]];
#pragma unroll
        for(int u = 0; u < SPMV_U; ++u)
            if(q0 + u * LPR < e) acc = fma(v[u], xv[u], acc);
    }
#pragma unroll
    for(int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
    return acc;
}

__global__ __launch_bounds__(256) void k_csr_spmv(int32_t m, const int32_t* __restrict__ rp,
    const int32_t* __restrict__ ci, const double* __restrict__ cv, const double* __restrict__ x,
    double* __restrict__ y, double alpha, double beta)
{
    const int64_t tid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t row0 = tid / LPR;
    const int64_t stride = (int64_t) gridDim.x * blockDim.x / LPR;
    for(int64_t row = row0; row < m; row += stride) {
        const double acc = csr_row_sum(rp, ci, cv, x, row, sub);
        if(sub == 0) {
            const double r = alpha * acc;
            y[row] = (beta == 0.0) ? r : fma(beta, y[row], r);
        }
    }
}

// y = K x of every listed system: item {s, c} is rows BATCH_SPMV_ROWS c ..
// of system s. Each group of eight lanes sums BATCH_SPMV_ROWS / 32 rows,
// rows 32 k + group, with the loads of all its rows issued together; every
// row is summed exactly as csr_row_sum sums it (the lane's entries in
// order, then the xor tree over the eight lanes), alpha 1, beta 0.
__global__ __launch_bounds__(256) void bk_csr_spmv(const int2* items, const BatchSys* sys,
    int32_t xid, int32_t yid)
{
    constexpr int RB = BATCH_SPMV_ROWS / 32;
    const int2 it = items[blockIdx.x];
    if(it.x < 0) return;   // a hole of an XCD-aware list (the whole workgroup)
    const BatchSys& S = sys[it.x];
    const int sub = threadIdx.x & (LPR - 1);
    const int grp = threadIdx.x / LPR;
    const int32_t* __restrict__ rp = S.k_rp;
    const int32_t* __restrict__ ci = S.k_col;
    const double* __restrict__ cv = S.k_val;
    const double* __restrict__ x = batch_vec(S, xid);
    double* __restrict__ y = batch_vec(S, yid);
    const int32_t n = S.n;
    int32_t row[RB], s[RB], e[RB];
    double acc[RB];
#pragma unroll
    for(int j = 0; j < RB; ++j) {
        row[j] = it.y * BATCH_SPMV_ROWS + 32 * j + grp;
        const int32_t r = min(row[j], n - 1);
        s[j] = rp[r];
        e[j] = rp[r + 1];
        if(row[j] >= n) e[j] = s[j];
        acc[j] = 0.0;
    }
    for(int32_t off = sub;; off += LPR * BATCH_SPMV_U) {
        int32_t c[RB][BATCH_SPMV_U];
        double v[RB][BATCH_SPMV_U], xv[RB][BATCH_SPMV_U];
#pragma unroll
        for(int j = 0; j < RB; ++j)
#pragma unroll
            for(int u = 0; u < BATCH_SPMV_U; ++u) {
                const int32_t q = max(min(s[j] + off + u * LPR, e[j] - 1), 0);
                c[j][u] = ci[q];
                v[j][u] = cv[q];
            }
#pragma unroll
        for(int j = 0; j < RB; ++j)
#pragma unroll
            for(int u = 0; u < BATCH_SPMV_U; ++u) xv[j][u] = x[c[j][u]];
        bool more = false;
#pragma unroll
        for(int j = 0; j < RB; ++j) {
#pragma unroll
            for(int u = 0; u < BATCH_SPMV_U; ++u)
                if(s[j] + off + u * LPR < e[j]) acc[j] = fma(v[j][u], xv[j][u], acc[j]);
            more = more || s[j] + off + LPR * BATCH_SPMV_U < e[j];
        }
        if(wave_ballot(more) == 0) break;
    }
#pragma unroll
    for(int j = 0; j < RB; ++j) {
        double a = acc[j];
#pragma unroll
        for(int o = LPR / 2; o > 0; o >>= 1) a += __shfl_xor(a, o, LPR);
        if(sub == 0 && row[j] < n) y[row[j]] = 1.0 * a;
    }
}

// KKT: b1 = H x1 + A' x2 (H = diag(d), or a CSR matrix hp/hi/hv) ;
// b2 = delta * x2(ix) + A x1 ; y = alpha b + beta y. A row of H and the
// row of A' are summed by the same eight lanes, H's entries first.
__global__ __launch_bounds__(256) void k_kkt_primal(int32_t n, const double* d,
    const int32_t* hp, const int32_t* hi, const double* hv, const int32_t* atp,
    const int32_t* ati, const double* atv, const double* x, double* y, double alpha, double beta)
{
    const int64_t tid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t row = tid / LPR;
    double acc = 0.0;
    if(row < n) {
        if(hp)
            for(int32_t q = hp[row] + sub; q < hp[row + 1]; q += LPR) acc = fma(hv[q], x[hi[q]], acc);
        const double* x2 = x + n;
        for(int32_t q = atp[row] + sub; q < atp[row + 1]; q += LPR) acc = fma(atv[q], x2[ati[q]], acc);
    }
#pragma unroll
    for(int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
    if(row < n && sub == 0) {
        const double b = hp ? acc : fma(d[row], x[row], acc);
        y[row] = (beta == 0.0) ? alpha * b : fma(beta, y[row], alpha * b);
    }
}

__global__ __launch_bounds__(256) void k_kkt_dual(int32_t n, int32_t m, const double* dix,
    const int32_t* ap, const int32_t* ai, const double* av, const double* x, double* y,
    double alpha, double beta)
{
    const int64_t tid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t row = tid / LPR;
    double acc = 0.0;
    if(row < m) {
        for(int32_t q = ap[row] + sub; q < ap[row + 1]; q += LPR) acc = fma(av[q], x[ai[q]], acc);
    }
#pragma unroll
    for(int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
    if(row < m && sub == 0) {
        const double b = fma(dix[row], x[n + row], acc);
        y[n + row] = (beta == 0.0) ? alpha * b : fma(beta, y[n + row], alpha * b);
    }
}

// normal equations: t = (A' x) ./ d, then y = alpha (A t + dix .* x) + beta y
__global__ __launch_bounds__(256) void k_normal_inner(int32_t n, const int32_t* atp,
    const int32_t* ati, const double* atv, const double* d, const double* x, double* t)
{
    const int64_t tid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t row = tid / LPR;
    double acc = 0.0;
    if(row < n)
        for(int32_t q = atp[row] + sub; q < atp[row + 1]; q += LPR) acc = fma(atv[q], x[ati[q]], acc);
#pragma unroll
    for(int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
    if(row < n && sub == 0) t[row] = acc / d[row];
}

__global__ __launch_bounds__(256) void k_normal_outer(int32_t m, const int32_t* ap,
    const int32_t* ai, const double* av, const double* dix, const double* t, const double* x,
    double* y, double alpha, double beta)
{
    const int64_t tid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t row = tid / LPR;
    double acc = 0.0;
    if(row < m)
        for(int32_t q = ap[row] + sub; q < ap[row + 1]; q += LPR) acc = fma(av[q], t[ai[q]], acc);
#pragma unroll
    for(int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
    if(row < m && sub == 0) {
        const double b = fma(dix[row], x[row], acc);
        y[row] = (beta == 0.0) ? alpha * b : fma(beta, y[row], alpha * b);
    }
}

}  // namespace

void normal_spmv(const DevNormal& N, const double* x, double* y, double alpha, double beta,
    void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    if(N.n > 0)
        k_normal_inner<<<(unsigned) (((int64_t) N.n * LPR + 255) / 256), 256, 0, st>>>(N.n,
            N.At.row_ptr, N.At.col, N.At.val, N.d, x, N.t);
    if(N.m > 0)
        k_normal_outer<<<(unsigned) (((int64_t) N.m * LPR + 255) / 256), 256, 0, st>>>(N.m,
            N.A.row_ptr, N.A.col, N.A.val, N.dix, N.t, x, y, alpha, beta);
}

void csr_spmv(const DevCsr& A, const double* x, double* y, double alpha, double beta,
    void* stream)
{
    if(A.m == 0) return;
    const int64_t threads = (int64_t) A.m * LPR;
    int64_t grid = (threads + 255) / 256;
    if(grid > 65536) grid = 65536;
    k_csr_spmv<<<(unsigned) grid, 256, 0, (hipStream_t) stream>>>(A.m, A.row_ptr, A.col, A.val,
        x, y, alpha, beta);
}

void kkt_spmv(const DevKkt& K, const double* x, double* y, double alpha, double beta,
    void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    if(K.n > 0)
        k_kkt_primal<<<(unsigned) (((int64_t) K.n * LPR + 255) / 256), 256, 0, st>>>(K.n, K.d,
            K.H.row_ptr, K.H.col, K.H.val, K.At.row_ptr, K.At.col, K.At.val, x, y, alpha, beta);
    if(K.m > 0)
        k_kkt_dual<<<(unsigned) (((int64_t) K.m * LPR + 255) / 256), 256, 0, st>>>(K.n, K.m,
            K.dix, K.A.row_ptr, K.A.col, K.A.val, x, y, alpha, beta);
}


void batch_csr_spmv(const BatchLists& L, const BatchSys* sys, int32_t x, int32_t y,
    hipStream_t st)
{
    if(L.spmv.count > 0)
        bk_csr_spmv<<<(unsigned) L.spmv.count, 256, 0, st>>>(L.spmv.items, sys, x, y);
}

}  // namespace culip
