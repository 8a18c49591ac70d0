// Numeric BILDLt factor for gfx950, one level at a time.
//
// Reference: libs/la/preconditioner/block_ildlt.impl.cuh:1519-1671 (schedule)
// and block_ildlt_kernels.impl.cuh:1146-2074 (kernels). Per level:
//   k_pivots_pairs
//               one wave per pivot block: A_ii -= sum_k U_ki' D_k U_ki over
//               the block column, 1x1/2x2 LDL' of the tile (static or
//               numeric pivots), row sums; in the same launch, one lane per
//               small off-diagonal block: its Schur updates
//   k_offdiag_tiny_scale
//               those blocks through U_ii' and D, row partial sums, compact
//               entries
//   k_offdiag   one wave per off-diagonal block: A_ij -= sum_k U_ki' D_k U_kj
//               over the block's symbolic Schur pairs, U_ij = D_i \ (U_ii' \
//               A_ij), per-row partial sums of squares; the result leaves as
//               a dense tile or as a compact (index, value) list of the
//               entries with |v| > 1e-10
//   k_rownorm   one wave per block row: row sums of squares in a fixed order
//   k_level_scan
//               chunked scan in level-list order: entries surviving the
//               threshold, leftover-space pool, granted extra space, output
//               offsets and the factored metadata
//   k_store     one wave per block: compacted final storage
//
// The dependency blocks of a tile are walked with their metadata fetched
// lane-parallel up front. In the pivot blocks a block whose nonzero entries share
// one column (one entry, on sparse KKT patterns) has its whole update -- one
// diagonal element -- formed by its own lane, so the serial walk over the
// block column costs a few scalar instructions per block; other blocks are
// staged as tiles. The update U' D U is applied over the nonzero rows
// of the right factor only, with D applied on the fly: every term the
// reference adds is added, in the same order with the same fused
// multiply-adds as the CPU oracle's restatement; the skipped terms are exact
// zeros. FP64 factors therefore agree with the oracle bit for bit.
#include "factor.hpp"
#include "wave.hpp"

#include <hip/hip_runtime.h>

namespace culip {

namespace {

constexpr double NZ_EPS = 1e-10;

// first Schur pair of an off-diagonal task: low and high 32 bits in the
// third int4's x and w
__device__ __forceinline__ int64_t pair_first(int4 t2)
{
    return (int64_t) (((uint64_t) (uint32_t) t2.w << 32) | (uint32_t) t2.x);
}

template<typename VT>
__device__ __forceinline__ void tile_zero(VT* T)
{
    const int lane = lane_id();
#pragma unroll
    for(int t = 0; t < 16; ++t) T[64 * t + lane] = (VT) 0;
}

template<typename VT>
__device__ __forceinline__ void tile_copy(VT* dst, const VT* src)
{
    const int lane = lane_id();
#pragma unroll
    for(int t = 0; t < 16; ++t) dst[64 * t + lane] = src[64 * t + lane];
}

// d_load_block (kernels :22-81) for an input block: zero, then fill
template<typename VT>
__device__ void tile_load(VT* T, const BlockMeta& m, const int32_t* bs, const int32_t* ixs,
    const VT* vals)
{
    const int lane = lane_id();
    const int h = bs[m.block_row + 1] - bs[m.block_row];
    const int w = bs[m.block_col + 1] - bs[m.block_col];
    tile_zero(T);
    wave_sync();
    if(m.format == FMT_DENSE) {
#pragma unroll
        for(int t = 0; t < 16; ++t) {
            const int idx = 64 * t + lane, r = idx >> 5, c = idx & 31;
            if(r < h && c < w) T[idx] = vals[(int64_t) m.val_ptr + r * w + c];
        }
    } else {
        for(int e = lane; e < m.nnz; e += 64) {
            const int ix = ixs[(int64_t) m.col_ptr + e];
            const int r = ix / w, c = ix - r * w;
            T[32 * r + c] = vals[(int64_t) m.val_ptr + e];
        }
    }
    wave_sync();
}

// metadata of a factored dependency block, one per lane
struct Dep {
    int fmt, nnz, cp, vp, rb, rk;
};

__device__ __forceinline__ Dep load_dep(const BlockMeta* bout, const int32_t* bs, int blk)
{
    const BlockMeta m = bout[blk];
    Dep d;
    d.fmt = m.format;
    d.nnz = m.nnz;
    d.cp = m.col_ptr;
    d.vp = m.val_ptr;
    d.rb = bs[m.block_row];
    d.rk = bs[m.block_row + 1] - d.rb;
    return d;
}

// src is uniform across the wave: the result lives in scalar registers
__device__ __forceinline__ Dep bcast(const Dep& d, int src)
{
    Dep o;
    o.fmt = __builtin_amdgcn_readlane(d.fmt, src);
    o.nnz = __builtin_amdgcn_readlane(d.nnz, src);
    o.cp = __builtin_amdgcn_readlane(d.cp, src);
    o.vp = __builtin_amdgcn_readlane(d.vp, src);
    o.rb = __builtin_amdgcn_readlane(d.rb, src);
    o.rk = __builtin_amdgcn_readlane(d.rk, src);
    return o;
}

// registers holding one dependency block: its (up to 64) sparse entries and
// the tridiagonal D of its rows
template<typename VT>
struct Pre {
    int ix;
    VT v;
    VT sup, dia, sub;
};

template<typename VT>
__device__ __forceinline__ bool small_sparse(const Dep& d)
{
    return d.fmt == FMT_SPARSE && d.nnz <= 64;
}

template<typename VT>
__device__ __forceinline__ Pre<VT> prefetch(const Dep& d, const int32_t* ixs, const VT* vals,
    const double* tri, bool want_d)
{
    const int lane = lane_id();
    Pre<VT> p;
    p.ix = 0;
    p.v = 0;
    p.sup = p.dia = p.sub = 0;
    if(small_sparse<VT>(d) && lane < d.nnz) {
        p.ix = ixs[(int64_t) d.cp + lane];
        p.v = vals[(int64_t) d.vp + lane];
    }
    const int l = lane & 31;
    if(want_d && l < d.rk) {
        const double* t = tri + 3 * (int64_t) (d.rb + l);
        p.sup = (VT) t[0];
        p.dia = (VT) t[1];
        p.sub = (VT) t[2];
    }
    return p;
}

// Writes a dependency block (width w) into tile X and, when mask is set,
// the nonzero rows of each column into mask[c] (LDS, integer atomics).
template<typename VT>
__device__ void stage(VT* X, uint32_t* mask, const Dep& d, const Pre<VT>& p, int w,
    const int32_t* ixs, const VT* vals)
{
    const int lane = lane_id();
    if(small_sparse<VT>(d)) {
        if(lane < d.nnz) {
            const int r = p.ix / w, c = p.ix - r * w;
            X[32 * r + c] = p.v;
            if(mask && p.v != (VT) 0) atomicOr(&mask[c], 1u << r);
        }
    } else if(d.fmt == FMT_DENSE) {
#pragma unroll
        for(int t = 0; t < 16; ++t) {
            const int idx = 64 * t + lane, r = idx >> 5, c = idx & 31;
            if(r < d.rk && c < w) {
                const VT v = vals[(int64_t) d.vp + r * w + c];
                X[idx] = v;
                if(mask && v != (VT) 0) atomicOr(&mask[c], 1u << r);
            }
        }
    } else {
        for(int e = lane; e < d.nnz; e += 64) {
            const int ix = ixs[(int64_t) d.cp + e];
            const VT v = vals[(int64_t) d.vp + e];
            const int r = ix / w, c = ix - r * w;
            X[32 * r + c] = v;
            if(mask && v != (VT) 0) atomicOr(&mask[c], 1u << r);
        }
    }
}

template<typename VT>
__device__ void unstage(VT* X, uint32_t* mask, const Dep& d, const Pre<VT>& p, int w)
{
    const int lane = lane_id();
    if(small_sparse<VT>(d)) {
        if(lane < d.nnz) {
            const int r = p.ix / w, c = p.ix - r * w;
            X[32 * r + c] = (VT) 0;
        }
    } else {
        tile_zero(X);
    }
    if(mask && lane < 32) mask[lane] = 0;
}

template<typename VT>
__device__ __forceinline__ void stage_d(VT* cs, VT* cd, VT* cb, const Pre<VT>& p)
{
    const int lane = lane_id();
    if(lane < 32) {
        cs[lane] = p.sup;
        cd[lane] = p.dia;
        cb[lane] = p.sub;
    }
}

// T[c][j] -= sum_l R[l][j] * Lt[c][l] over the nonzero rows l of R's column
// j, ascending, with Lt[c][l] = (A' D)[c][l] formed from A[l-1..l+1][c] as
// d_multiply_block_by_block_diagonal does (kernels :372-456)
template<typename VT>
__device__ void product(VT* T, const VT* A, const VT* R, const uint32_t* mask, const VT* cs,
    const VT* cd, const VT* cb, int lh, int rh)
{
    const int lane = lane_id();
    const int j = lane & 31;
    const int c0 = lane >> 5;
    uint32_t mk = (j < rh) ? mask[j] : 0u;
    if(mk) {
        VT acc[16];
#pragma unroll
        for(int t = 0; t < 16; ++t) acc[t] = 0;
        while(mk) {
            const int l = __ffs(mk) - 1;
            mk &= mk - 1;
            const VT r = R[32 * l + j];
            const VT sup = cs[l], dia = cd[l], sub = cb[l];
#pragma unroll
            for(int t = 0; t < 16; ++t) {
                const int c = c0 + 2 * t;
                const VT left = (l == 0) ? (VT) 0 : A[32 * (l - 1) + c];
                const VT mid = A[32 * l + c];
                const VT right = (l == 31) ? (VT) 0 : A[32 * (l + 1) + c];
                const VT lt = fma(sub, right, fma(dia, mid, sup * left));
                acc[t] = fma(r, lt, acc[t]);
            }
        }
#pragma unroll
        for(int t = 0; t < 16; ++t) {
            const int c = c0 + 2 * t;
            if(c < lh) T[32 * c + j] -= acc[t];
        }
    }
}

// The Givens-form solve of d_left_solve_with_D_Ut (kernels :1533-1566) for
// row l of a block of hgt rows, D pointing at row l's tridiagonal triple:
// x1 for row l, x2 for row l + 1 when a 2x2 pivot starts at l (p2). The
// arithmetic runs in AT: the factor's type as the reference computes it,
// or FP64 for an FP32 factor whose 2x2 pivots are inverted in FP64. A 1x1
// pivot (d12 = 0) gives x1 = rl / d11 exactly either way.
template<typename VT, typename AT>
__device__ __forceinline__ void givens_d_solve(const double* D, int l, int hgt, VT rl_, VT right_,
    bool p2, VT& x1_, VT& x2_)
{
    const AT rl = (AT) rl_, right = (AT) right_;
    const AT d11 = (AT) D[1];
    const AT d12 = (AT) D[2];
    AT d22 = (l + 1 < hgt) ? (AT) D[4] : (AT) 0;
    if(l == 31) d22 = d11;
    const AT r = sqrt(fma(d11, d11, d12 * d12));
    const AT c = d11 / r, sn = d12 / r;
    AT x1 = fma(c, rl, sn * right);
    AT x2 = fma(-sn, rl, c * right);
    const AT t11 = fma(c, d11, sn * d12);
    const AT t12 = fma(c, d12, sn * d22);
    const AT t22 = fma(-sn, d12, c * d22);
    x2 = p2 ? (x2 / t22) : (AT) 0;
    x1 = fma(-t12, x2, x1) / t11;
    x1_ = (VT) x1;
    x2_ = (VT) x2;
}

template<typename VT>
__device__ __forceinline__ void givens_d_solve(bool f64, const double* D, int l, int hgt, VT rl,
    VT right, bool p2, VT& x1, VT& x2)
{
    if(f64) givens_d_solve<VT, double>(D, l, hgt, rl, right, p2, x1, x2);
    else givens_d_solve<VT, VT>(D, l, hgt, rl, right, p2, x1, x2);
}

// d_ldlt_1x1_step (kernels :538-585)
template<typename VT>
__device__ void step_1x1(VT* T, double* tri, int p, int len)
{
    const int lane = lane_id();
    const int j = lane & 31, half = lane >> 5;
    VT d = T[32 * p + p];
    bool mod = false;
    if(fabs((double) d) < 1e-8) { d = (VT) 1e-6; mod = true; }
    const VT a = (j < p) ? (VT) 0 : T[32 * p + j];
    const VT u = a / d;
    const VT ru = u * d;
    wave_sync();
    if(half == 0) T[32 * p + j] = u;
    // trailing rows i > p, two per pass: only rows with a nonzero multiplier
    // ru_i change, and only if the pivot row has a nonzero u_j right of the
    // diagonal; the other updates would add exact zeros
    const uint64_t ucols = wave_ballot(half == 0 && j > p && j < len && u != (VT) 0);
    uint64_t rows = ucols ? wave_ballot(half == 0 && j > p && j < len && ru != (VT) 0) : 0ull;
    while(rows) {
        const int ia = __ffsll((long long) rows) - 1;
        rows &= rows - 1;
        const int ib = rows ? __ffsll((long long) rows) - 1 : -1;
        if(rows) rows &= rows - 1;
        const VT ra = wave_bcast(ru, ia);
        const VT rb = wave_bcast(ru, ib < 0 ? ia : ib);
        const int i = half ? ib : ia;
        const VT rui = half ? rb : ra;
        if(i >= 0 && j < len) T[32 * i + j] = fma(-rui, u, T[32 * i + j]);
    }
    wave_sync();
    if(mod && lane == 0) T[32 * p + p] = 1;
    if(lane < 3) tri[3 * p + lane] = (lane == 1) ? (double) d : 0.0;
    wave_sync();
}

// d_ldlt_2x2_step (kernels :589-675): Givens rotation of the 2x2 pivot.
// The arithmetic runs in AT (see givens_d_solve); the multipliers are
// stored in the factor's type and the trailing update subtracts their
// products with the FP64-exact row factors r1, r2. A replaced near-singular
// pivot is (VT) 1e-6 on the diagonal in both forms.
template<typename VT, typename AT>
__device__ void step_2x2(VT* T, double* tri, int p, int len)
{
    const int lane = lane_id();
    const int j = lane & 31, half = lane >> 5;
    bool mod = false;
    AT d11 = (AT) T[32 * p + p];
    const AT d12 = (AT) T[32 * p + p + 1];
    AT d22 = (AT) T[32 * (p + 1) + p + 1];
    const AT det = fma(d11, d22, -(d12 * d12));
    if(fabs((double) det) < 1e-8) { d11 = (AT) (VT) 1e-6; d22 = (AT) (VT) 1e-6; mod = true; }
    const AT r = sqrt(fma(d11, d11, d12 * d12));
    const AT c = d11 / r, s = d12 / r;
    const AT t11 = fma(c, d11, s * d12);
    const AT t12 = fma(c, d12, s * d22);
    const AT t22 = fma(-s, d12, c * d22);
    const AT a1 = (j < p) ? (AT) 0 : (AT) T[32 * p + j];
    const AT a2 = (j < p) ? (AT) 0 : ((j == p) ? d12 : (AT) T[32 * (p + 1) + j]);
    AT x1 = fma(c, a1, s * a2);
    AT x2 = fma(-s, a1, c * a2);
    x2 = x2 / t22;
    x1 = fma(-t12, x2, x1) / t11;
    const VT u1 = (VT) x1, u2 = (VT) x2;
    const AT r1 = fma(x1, d11, x2 * d12);
    const AT r2 = fma(x1, d12, x2 * d22);
    wave_sync();
    if(half == 0) {
        T[32 * p + j] = (j < p) ? (VT) 0 : u1;
        T[32 * (p + 1) + j] = (j < p) ? (VT) 0 : u2;
    }
    // trailing rows i > p + 1 as in step_1x1: rows with a nonzero (r1_i,
    // r2_i), when the pivot rows have a nonzero (u1_j, u2_j) beyond them
    const uint64_t ucols = wave_ballot(half == 0 && j > p + 1 && j < len &&
        (u1 != (VT) 0 || u2 != (VT) 0));
    uint64_t rows = ucols ? wave_ballot(half == 0 && j > p + 1 && j < len &&
        (r1 != (AT) 0 || r2 != (AT) 0)) : 0ull;
    while(rows) {
        const int ia = __ffsll((long long) rows) - 1;
        rows &= rows - 1;
        const int ib = rows ? __ffsll((long long) rows) - 1 : -1;
        if(rows) rows &= rows - 1;
        const AT r1a = wave_bcast(r1, ia), r2a = wave_bcast(r2, ia);
        const AT r1b = wave_bcast(r1, ib < 0 ? ia : ib), r2b = wave_bcast(r2, ib < 0 ? ia : ib);
        const int i = half ? ib : ia;
        const AT r1i = half ? r1b : r1a, r2i = half ? r2b : r2a;
        if(i >= 0 && j < len)
            T[32 * i + j] = (VT) ((AT) T[32 * i + j] - fma(r1i, (AT) u1, r2i * (AT) u2));
    }
    wave_sync();
    if(mod && lane == 0) {
        T[32 * p + p] = 1;
        T[32 * p + p + 1] = 0;
        T[32 * (p + 1) + p + 1] = 1;
    }
    if(lane < 3) {
        tri[3 * p + lane] = (lane == 0) ? 0.0 : ((lane == 1) ? (double) d11 : (double) d12);
        tri[3 * (p + 1) + lane] = (lane == 0) ? (double) d12 : ((lane == 1) ? (double) d22 : 0.0);
    }
    wave_sync();
}

template<typename VT>
__device__ VT row_ssq(const VT* T, int r, int w)
{
    VT s = 0;
    for(int j = 0; j < w; ++j) s = fma(T[32 * r + j], T[32 * r + j], s);
    return s;
}

// dependency blocks staged per round of the pivot update; the first tile
// holds the pivot tile for the factorization afterwards. Two keep the pivot
// blocks' LDS (17.8 KB in FP64) below that of the one-lane off-diagonal
// form they share launches with, so eight workgroups fit a CU.
constexpr int PIV_ROUND = 2;

template<typename VT>
struct alignas(16) PivLds {
    VT X[PIV_ROUND][1024];
    VT cs[PIV_ROUND][32], cd[PIV_ROUND][32], cb[PIV_ROUND][32];
    uint32_t mask[PIV_ROUND][32];
};

// T (registers, rows c0 + 2t of column j) -= A' D A for a staged tile A
// whose nonzero rows per column are in mask; the arithmetic and order of
// product() with R = A, one output row at a time
template<typename VT>
__device__ void product_regs(VT (&T)[16], const VT* A, const uint32_t* mask, const VT* cs,
    const VT* cd, const VT* cb, int len)
{
    const int lane = lane_id();
    const int j = lane & 31;
    const int c0 = lane >> 5;
    const uint32_t mk = (j < len) ? mask[j] : 0u;
    if(mk) {
        // one pass over the nonzero rows with sixteen independent
        // accumulators (the loads of a row feed all of them)
        VT acc[16];
#pragma unroll
        for(int t = 0; t < 16; ++t) acc[t] = 0;
        for(uint32_t bits = mk; bits; bits &= bits - 1) {
            const int l = __ffs(bits) - 1;
            const VT r = A[32 * l + j];
            const VT sup = cs[l], dia = cd[l], sub = cb[l];
#pragma unroll
            for(int t = 0; t < 16; ++t) {
                const int c = c0 + 2 * t;
                const VT left = (l == 0) ? (VT) 0 : A[32 * (l - 1) + c];
                const VT mid = A[32 * l + c];
                const VT right = (l == 31) ? (VT) 0 : A[32 * (l + 1) + c];
                const VT lt = fma(sub, right, fma(dia, mid, sup * left));
                acc[t] = fma(r, lt, acc[t]);
            }
        }
#pragma unroll
        for(int t = 0; t < 16; ++t)
            if(c0 + 2 * t < len) T[t] -= acc[t];
    }
}

// the tile element T[c][c] held by this lane (rows cq + 2t of column j)
template<typename VT>
__device__ __forceinline__ VT diag_get(const VT (&T)[16], bool own, int t0)
{
    VT v = (VT) 0;
#pragma unroll
    for(int t = 0; t < 16; ++t)
        if(own && t == t0) v = T[t];
    return v;
}

template<typename VT>
__device__ __forceinline__ void diag_put(VT (&T)[16], bool own, int t0, VT v)
{
#pragma unroll
    for(int t = 0; t < 16; ++t)
        if(own && t == t0) T[t] = v;
}

__device__ __forceinline__ double lane_value(double v, int src)
{
    const unsigned long long b = (unsigned long long) __double_as_longlong(v);
    const unsigned lo = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) b, src);
    const unsigned hi = (unsigned) __builtin_amdgcn_readlane((int) (unsigned) (b >> 32), src);
    return __longlong_as_double((long long) (((unsigned long long) hi << 32) | lo));
}

__device__ __forceinline__ float lane_value(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// nonzero entries of a dependency block taken by the one-column form
constexpr int COL_E = 4;

// A dependency block of width w whose nonzero entries all lie in one column
// c updates only T[c][c] in product_regs; this forms that update in one
// lane with the same operations in the same order (rows ascending, the
// three-term D product, fma accumulation). Returns false for any other
// block; c = -1 when the block has no nonzero entry.
template<typename VT>
__device__ __forceinline__ bool column_update(const Dep& d, int w, const int32_t* ixs,
    const VT* vals, const double* tri, int& c_out, VT& v_out)
{
    c_out = -1;
    v_out = (VT) 0;
    if(d.fmt != FMT_SPARSE || d.nnz > COL_E) return false;
    constexpr int NONE = 1 << 20;
    int r[COL_E];
    VT v[COL_E];
    int col = -1;
    bool one_col = true;
#pragma unroll
    for(int e = 0; e < COL_E; ++e) {
        r[e] = NONE;
        v[e] = (VT) 0;
        if(e < d.nnz) {
            const int ix = ixs[(int64_t) d.cp + e];
            const VT x = vals[(int64_t) d.vp + e];
            if(x != (VT) 0) {
                const int rr = ix / w, cc = ix - rr * w;
                one_col = one_col && (col < 0 || cc == col);
                col = cc;
                r[e] = rr;
                v[e] = x;
            }
        }
    }
    if(!one_col) return false;
    if(col < 0) return true;
    // rows ascending (sorting network on four)
    auto cswap = [&](int p, int q) {
        if(r[q] < r[p]) {
            const int tr = r[p]; r[p] = r[q]; r[q] = tr;
            const VT tv = v[p]; v[p] = v[q]; v[q] = tv;
        }
    };
    cswap(0, 1); cswap(2, 3); cswap(0, 2); cswap(1, 3); cswap(1, 2);
    VT acc = (VT) 0;
#pragma unroll
    for(int k = 0; k < COL_E; ++k) {
        if(r[k] != NONE) {
            const int l = r[k];
            const double* t = tri + 3 * (int64_t) (d.rb + l);
            const VT sup = (VT) t[0], dia = (VT) t[1], sub = (VT) t[2];
            const VT left = (k > 0 && l > 0 && r[k - 1] == l - 1) ? v[k - 1] : (VT) 0;
            const VT right = (k + 1 < COL_E && l < 31 && r[k + 1] == l + 1) ? v[k + 1] : (VT) 0;
            const VT lt = fma(sub, right, fma(dia, v[k], sup * left));
            acc = fma(v[k], lt, acc);
        }
    }
    c_out = col;
    v_out = acc;
    return true;
}

template<typename VT>
struct alignas(16) OffLds {
    VT T[1024];
    VT A[1024];
    VT R[1024];
    VT cs[32], cd[32], cb[32];
    uint32_t mask[32];
    uint32_t umask[32];
};

// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Numeric pivoting inside a pivot tile T (stride 32, the matrix in its upper
// triangle), one wave. Lane l < 32 stands for row / column l, as a warp lane
// of the reference does.

// largest value over the wave and the lowest lane holding it among the
// lanes with `in` set (the reference's shuffle reduction keeps the lower
// lane on ties)
template<typename VT>
__device__ __forceinline__ int wave_argmax_first(VT v, bool in, VT& best)
{
    best = wave_max(in ? v : (VT) 0);
    const uint64_t at = wave_ballot(in && v == best);
    return at ? __ffsll((long long) at) - 1 : -1;
}

template<typename VT>
__device__ __forceinline__ VT sym_abs(const VT* T, int a, int b)
{
    return (a < b) ? fabs(T[32 * a + b]) : fabs(T[32 * b + a]);
}

// d_select_pivot_bunch_kaufman (kernels :743-834); alpha = (1 + sqrtf(17))
// / 8 evaluated in float as the reference writes it
template<typename VT>
__device__ void select_bk(const VT* T, int len, int p, int& order, int& p0, int& p1)
{
    const int l = lane_id();
    const VT alpha = (VT) ((1.0f + sqrtf(17.0f)) / 8.0f);
    order = 1;
    p0 = p;
    p1 = p + 1;
    VT w1;
    const VT v = (l > p && l < len) ? fabs(T[32 * p + min(l, 31)]) : (VT) 0;
    const int r = wave_argmax_first(v, l > p && l < len, w1);
    if(!(w1 > (VT) 0)) return;
    const VT app = fabs(T[32 * p + p]);
    if(app >= alpha * w1) return;
    const bool in = l >= p && l < len && l != r;
    const VT wr = wave_max(in ? sym_abs(T, min(l, 31), r) : (VT) 0);
    if(app * wr >= alpha * w1 * w1) return;
    if(fabs(T[32 * r + r]) >= alpha * wr) { p0 = r; return; }
    order = 2;
    p1 = r;
}

// d_select_pivot_rook (kernels :838-948)
template<typename VT>
__device__ void select_rook(const VT* T, int len, int p, int& order, int& p0, int& p1)
{
    const int l = lane_id();
    const VT alpha = (VT) ((1.0f + sqrtf(17.0f)) / 8.0f);
    order = 1;
    p0 = p;
    p1 = p + 1;
    VT wi;
    const VT v = (l > p && l < len) ? fabs(T[32 * p + min(l, 31)]) : (VT) 0;
    int r = wave_argmax_first(v, l > p && l < len, wi);
    if(!(wi > (VT) 0)) return;
    int i = p;
    if(fabs(T[32 * p + p]) >= alpha * wi) return;
    for(;;) {
        VT wr;
        const bool in = l >= p && l < len && l != r;
        int q = wave_argmax_first(in ? sym_abs(T, min(l, 31), r) : (VT) 0, in, wr);
        // no nonzero candidate: the reference's reduction leaves lane 0's
        // start index, max(0, p)
        if(!(wr > (VT) 0)) q = p;
        if(fabs(T[32 * r + r]) >= alpha * wr) { p0 = r; return; }
        if(wi == wr) { order = 2; p0 = i; p1 = r; return; }
        i = r;
        r = q;
        wi = wr;
    }
}

// d_dense_pivot_op (kernels :679-739): symmetric interchange of rows and
// columns a and b of the tile (earlier, factored rows included, their
// entries in the two columns move too) and of perm[a], perm[b]
template<typename VT>
__device__ void dense_pivot_op(VT* T, int32_t* perm, int len, int pa, int pb)
{
    const int l = lane_id();
    wave_sync();
    if(l < len) {
        int x, y;
        if(l < pa) { x = 32 * l + pa; y = 32 * l + pb; }
        else if(l == pa) { x = 32 * pa + pa; y = 32 * pb + pb; }
        else if(l < pb) { x = 32 * pa + l; y = 32 * l + pb; }
        else if(l == pb) { x = y = 32 * pa + pb; }
        else { x = 32 * pa + l; y = 32 * pb + l; }
        const VT tx = T[x], ty = T[y];
        T[x] = ty;
        T[y] = tx;
    }
    if(l == 0) {
        const int32_t t = perm[pa];
        perm[pa] = perm[pb];
        perm[pb] = t;
    }
    wave_sync();
}

// k_factor_pivot_blocks (kernels :1146-1283) for the pivot block at level
// position q, one wave
template<typename VT>
__device__ void pivots_one(const FactorArgs<VT>& a, const LevelRange& lr, const int q,
    PivLds<VT>& s)
{
    const int lane = lane_id();
    const int b = a.lvlb_idx[lr.lo + q];
    const BlockMeta m = a.bin[b];
    const int i = m.block_row;
    const int fr = a.bs[i];
    const int len = a.bs[i + 1] - fr;
    const int cb0 = a.csc_off[i];
    const int nd = a.csc_off[i + 1] - 1 - cb0;
    const int j = lane & 31, cq = lane >> 5;
    // metadata of the first 64 dependency blocks, in flight during the tile
    // load; later chunks are fetched while the previous one is walked
    Dep nxt{};
    if(lane < nd) nxt = load_dep(a.bout, a.bs, a.csc_idx[cb0 + lane]);
    // the pivot tile in registers (rows cq + 2t of column j)
    VT* T0 = s.X[0];
    tile_load(T0, m, a.bs, a.ix_in, a.val_in);
    VT T[16];
#pragma unroll
    for(int t = 0; t < 16; ++t) T[t] = T0[32 * (cq + 2 * t) + j];
    wave_sync();
#pragma unroll
    for(int u = 0; u < PIV_ROUND; ++u) tile_zero(s.X[u]);
    if(lane < 32) {
#pragma unroll
        for(int u = 0; u < PIV_ROUND; ++u) s.mask[u][lane] = 0;
    }
    wave_sync();
    // updates from the blocks (k, i) above the diagonal, in CSC order, 64
    // blocks per chunk. A block whose nonzero entries lie in one column c
    // (nearly all of them on sparse KKT patterns) changes only T[c][c]; its
    // lane forms that value and the walk subtracts it from the diagonal
    // register of the owning lane. Runs of other blocks are staged as tiles,
    // PIV_ROUND per round, and go through product_regs.
    const bool owns_diag = ((j & 1) == cq) && j < len;
    const int tdiag = (j - cq) >> 1;
    for(int c0 = 0; c0 < nd; c0 += 64) {
        const Dep mine = nxt;
        const int cnt = min(64, nd - c0);
        int uc = -1;
        VT uv = (VT) 0;
        const bool simple = lane < cnt && column_update(mine, len, a.ix_out, a.val_out, a.tri,
            uc, uv);
        const uint64_t smask = wave_ballot(simple);
        nxt = Dep{};
        if(c0 + 64 + lane < nd) nxt = load_dep(a.bout, a.bs, a.csc_idx[cb0 + c0 + 64 + lane]);
        VT dg = diag_get(T, owns_diag, tdiag);
        int d = 0;
        while(d < cnt) {
            if((smask >> d) & 1ull) {
                const int c = __builtin_amdgcn_readlane(uc, d);
                if(c >= 0) {
                    const VT v = lane_value(uv, d);
                    if(owns_diag && j == c) dg -= v;
                }
                ++d;
                continue;
            }
            // a run of staged blocks [d, d1)
            const uint64_t later = smask >> d;
            const int d1 = later ? min(cnt, d + __ffsll((long long) later) - 1) : cnt;
            diag_put(T, owns_diag, tdiag, dg);
            for(int r0 = d; r0 < d1; r0 += PIV_ROUND) {
                const int nr = min(PIV_ROUND, d1 - r0);
                Dep cur[PIV_ROUND];
                Pre<VT> pc[PIV_ROUND];
#pragma unroll
                for(int u = 0; u < PIV_ROUND; ++u) {
                    cur[u] = bcast(mine, u < nr ? r0 + u : r0);
                    if(u < nr) pc[u] = prefetch(cur[u], a.ix_out, a.val_out, a.tri, true);
                }
#pragma unroll
                for(int u = 0; u < PIV_ROUND; ++u) {
                    if(u < nr) {
                        stage(s.X[u], s.mask[u], cur[u], pc[u], len, a.ix_out, a.val_out);
                        stage_d(s.cs[u], s.cd[u], s.cb[u], pc[u]);
                    }
                }
                wave_sync();
#pragma unroll
                for(int u = 0; u < PIV_ROUND; ++u)
                    if(u < nr) product_regs(T, s.X[u], s.mask[u], s.cs[u], s.cd[u], s.cb[u], len);
                wave_sync();
#pragma unroll
                for(int u = 0; u < PIV_ROUND; ++u)
                    if(u < nr) unstage(s.X[u], s.mask[u], cur[u], pc[u], len);
                wave_sync();
            }
            dg = diag_get(T, owns_diag, tdiag);
            d = d1;
        }
        diag_put(T, owns_diag, tdiag, dg);
    }
    // the updated tile back to LDS (tile 0 is zero again after unstaging)
#pragma unroll
    for(int t = 0; t < 16; ++t) T0[32 * (cq + 2 * t) + j] = T[t];
    wave_sync();
    // d_factor_pivot_factorize (kernels :956-1083): static pivots, or a
    // numeric choice per step with symmetric interchanges (the pivot starts
    // and the block's rows in perm are rewritten as pivots are chosen)
    double* tri = a.tri + 3 * (int64_t) fr;
    int32_t* fpiv = a.is_piv + fr;
    int32_t* bperm = a.perm + fr;
    int p = 0;
    while(p < len) {
        int order, p0 = p, p1 = p + 1;
        if(a.pivoting == 1 && p < len - 1) select_bk(T0, len, p, order, p0, p1);
        else if(a.pivoting == 2 && p < len - 1) select_rook(T0, len, p, order, p0, p1);
        else order = (p < len - 1 && !fpiv[p + 1]) ? 2 : 1;
        if(order == 1) {
            if(a.pivoting && p0 != p) dense_pivot_op(T0, bperm, len, p, p0);
            step_1x1(T0, tri, p, len);
            if(a.pivoting && lane == 0) fpiv[p] = 1;
        } else {
            if(a.pivoting) {
                if(p0 != p) dense_pivot_op(T0, bperm, len, p, p0);
                dense_pivot_op(T0, bperm, len, p + 1, p1);
            }
            if(a.piv64) step_2x2<VT, double>(T0, tri, p, len);
            else step_2x2<VT, VT>(T0, tri, p, len);
            if(a.pivoting && lane == 0) {
                fpiv[p] = 1;
                fpiv[p + 1] = 0;
            }
        }
        wave_sync();
        p += order;
    }
    uint32_t um = 0;
    if(lane < len) {
        a.rn[fr + lane] = row_ssq(T0, lane, len);
        // nonzero columns right of the diagonal, for the U_ii' solves
        for(int c = lane + 1; c < len; ++c)
            if(T0[32 * lane + c] != (VT) 0) um |= 1u << c;
        a.piv_umask[(int64_t) q * 32 + lane] = um;
    } else if(lane < 32) {
        a.piv_umask[(int64_t) q * 32 + lane] = 0u;   // all 32 words defined
    }
    // and whether there is any such column at all (one word per tile)
    const bool um_some = wave_ballot(um != 0u) != 0ull;
    if(lane == 0) a.piv_any[q] = um_some ? 1u : 0u;
    VT* out = a.if_val + (int64_t) q * 1024;
    for(int e = lane; e < len * len; e += 64) {
        const int r = e / len, c = e - r * len;
        out[e] = T0[32 * r + c];
    }
    if(lane == 0) {
        a.if_cnt[q] = len * len;
        a.piv_loc[i] = q;
    }
}

// k_update_scale_offdiagonal_blocks (kernels :1628-1803) with
// d_left_solve_with_D_Ut (:1489-1570): general form, all dependency blocks
// staged as tiles; takes the positions k_offdiag leaves to it.
template<typename VT>
__device__ void offdiag_general_one(const FactorArgs<VT>& a, const LevelRange& lr, int q,
    OffLds<VT>& s);

template<typename VT>
__device__ __forceinline__ void offdiag_general_body(const FactorArgs<VT>& a, const LevelRange& lr,
    int bid_, int nblk_)
{
    __shared__ OffLds<VT> s;
    const int ndef = a.defer_cnt[lr.level];
    for(int t = bid_; t < ndef; t += nblk_) {
        offdiag_general_one(a, lr, a.defer_list[t], s);
        __syncthreads();
    }
}

template<typename VT>
__device__ void offdiag_general_one(const FactorArgs<VT>& a, const LevelRange& lr, int q,
    OffLds<VT>& s)
{
    const int lane = lane_id();
    const int b = a.lvlb_idx[lr.lo + q];
    const BlockMeta m = a.bin[b];
    const int i = m.block_row, jb = m.block_col;
    const int fr = a.bs[i];
    const int hgt = a.bs[i + 1] - fr;
    const int wid = a.bs[jb + 1] - a.bs[jb];
    const int64_t p0 = a.upd_off[b];
    const int np = (int) (a.upd_off[b + 1] - p0);
    const int pq = a.piv_loc[i];
    const bool all1x1 = (wave_ballot(lane < hgt && !a.is_piv[fr + (lane & 31)]) == 0);
    tile_zero(s.A);
    if(lane < 32) { s.mask[lane] = 0; s.umask[lane] = 0; }
    if(np > 0) tile_zero(s.R);
    tile_load(s.T, m, a.bs, a.ix_in, a.val_in);
    for(int c0 = 0; c0 < np; c0 += 64) {
        Dep mi{}, mj{};
        if(c0 + lane < np) {
            mi = load_dep(a.bout, a.bs, a.upd_ki[p0 + c0 + lane]);
            mj = load_dep(a.bout, a.bs, a.upd_kj[p0 + c0 + lane]);
        }
        const int cnt = min(64, np - c0);
        Dep ci = bcast(mi, 0), cj = bcast(mj, 0);
        Pre<VT> pi = prefetch(ci, a.ix_out, a.val_out, a.tri, true);
        Pre<VT> pj = prefetch(cj, a.ix_out, a.val_out, a.tri, false);
        for(int d = 0; d < cnt; ++d) {
            Dep ni = ci, nj = cj;
            Pre<VT> qi = pi, qj = pj;
            if(d + 1 < cnt) {
                ni = bcast(mi, d + 1);
                nj = bcast(mj, d + 1);
                qi = prefetch(ni, a.ix_out, a.val_out, a.tri, true);
                qj = prefetch(nj, a.ix_out, a.val_out, a.tri, false);
            }
            stage(s.A, (uint32_t*) nullptr, ci, pi, hgt, a.ix_out, a.val_out);
            stage(s.R, s.mask, cj, pj, wid, a.ix_out, a.val_out);
            stage_d(s.cs, s.cd, s.cb, pi);
            wave_sync();
            product(s.T, s.A, s.R, s.mask, s.cs, s.cd, s.cb, hgt, wid);
            wave_sync();
            unstage(s.A, (uint32_t*) nullptr, ci, pi, hgt);
            unstage(s.R, s.mask, cj, pj, wid);
            wave_sync();
            ci = ni; cj = nj; pi = qi; pj = qj;
        }
    }

    // rows in the block row's pivoting order (kernels :1733-1751)
    if(a.pivoting) {
        tile_copy(s.A, s.T);
        wave_sync();
        for(int e = lane; e < hgt * 32; e += 64) {
            const int r = e >> 5, c = e & 31;
            s.T[e] = s.A[32 * (a.perm[fr + r] - fr) + c];
        }
        wave_sync();
        tile_zero(s.A);
        wave_sync();
    }
    // pivot tile U_ii (row-major) into A; umask[r] = nonzero columns > r
    const VT* piv = a.if_val + (int64_t) pq * 1024;
    for(int e = lane; e < hgt * hgt; e += 64) {
        const int r = e / hgt, c = e - r * hgt;
        const VT v = piv[e];
        s.A[32 * r + c] = v;
        if(c > r && v != (VT) 0) atomicOr(&s.umask[r], 1u << c);
    }
    wave_sync();

    // column j per lane (lanes >= 32 idle): forward substitution with U_ii'
    // (terms with a zero multiplier are exact no-ops and are skipped)
    const int j = lane;
    const VT* U = s.A;
    VT rhs[32];
#pragma unroll
    for(int l = 0; l < 32; ++l) rhs[l] = (l < hgt && j < wid) ? s.T[32 * l + j] : (VT) 0;
#pragma unroll
    for(int ii = 0; ii < 32; ++ii) {
        if(ii < hgt) {
            const uint32_t um = s.umask[ii];
            if(rhs[ii] != (VT) 0 && um) {
                const VT pv = rhs[ii] / U[32 * ii + ii];
#pragma unroll
                for(int l = ii + 1; l < 32; ++l)
                    if((um >> l) & 1u) rhs[l] = fma(-U[32 * ii + l], pv, rhs[l]);
            }
        }
    }
    const double* D = a.tri + 3 * (int64_t) fr;
    const int32_t* ip = a.is_piv + fr;
    wave_sync();
    if(all1x1) {
        // 1x1 pivots: the Givens-form solve reduces exactly to rhs / d
#pragma unroll
        for(int l = 0; l < 32; ++l)
            if(l < hgt && j < wid) s.T[32 * l + j] = (VT) ((double) rhs[l] / D[3 * l + 1]);
    } else {
#pragma unroll
        for(int l = 0; l < 32; ++l) {
            if(l < hgt) {
                const bool p1 = ip[l] != 0;
                const bool p2 = (l + 1 < hgt) && !ip[l + 1];
                const VT rl = rhs[l];
                const VT right = (l < 31) ? rhs[l + 1] : rhs[31];
                VT x1, x2;
                givens_d_solve(a.piv64 != 0, D + 3 * l, l, hgt, rl, right, p2, x1, x2);
                if(j < wid) {
                    if(p1) s.T[32 * l + j] = x1;
                    if(p2) s.T[32 * (l + 1) + j] = x2;
                }
            }
        }
    }
    wave_sync();
    if(lane < hgt) a.rn_part[(int64_t) q * 32 + lane] = row_ssq(s.T, lane, wid);

    VT* oval = a.if_val + (int64_t) q * 1024;
    if(m.format == FMT_DENSE) {
        for(int e = lane; e < hgt * wid; e += 64) {
            const int r = e / wid, c = e - r * wid;
            oval[e] = s.T[32 * r + c];
        }
        if(lane == 0) a.if_cnt[q] = hgt * wid;
    } else {
        uint16_t* oix = a.if_idx + (int64_t) q * 1024;
        int base = 0;
#pragma unroll
        for(int t = 0; t < 16; ++t) {
            const int idx = 64 * t + lane, r = idx >> 5, c = idx & 31;
            const VT v = s.T[idx];
            const bool pred = r < hgt && c < wid && fabs((double) v) > NZ_EPS;
            const uint64_t msk = wave_ballot(pred);
            if(pred) {
                const int pos = base + lanes_below(msk);
                oix[pos] = (uint16_t) idx;
                oval[pos] = v;
            }
            base += __popcll(msk);
        }
        if(lane == 0) a.if_cnt[q] = base;
    }
}

// ---------------------------------------------------------------------------
// Off-diagonal blocks whose right factors U_kj are sparse with at most 64
// entries (every block at level-0 fill of the KKT workloads): the output
// tile lives in registers (lane j + 32 c0 holds rows c0, c0 + 2, ... of
// column j), U_ki is staged in the one LDS tile, and U_kj becomes a list of
// its nonzero entries ordered by (column, row), so a wave needs about 10 KB
// of LDS instead of 25 KB. Arithmetic and order are those of the general
// form. A block with a larger right factor is handed to k_offdiag_general.

template<typename VT>
struct alignas(16) OffLdsSmall {
    VT X[1024];
    VT cs[32], cd[32], cb[32];
    VT rv[64];
    int32_t cst[33];
    uint8_t rl[64];
};

__device__ __forceinline__ Dep dep_meta(const BlockMeta* bout, int blk, int rb, int rk)
{
    const BlockMeta m = bout[blk];
    return Dep{m.format, m.nnz, m.col_ptr, m.val_ptr, rb, rk};
}

// nonzero entries of a small sparse block (width w, one per lane) sorted by
// (column, row) into rl / rv; cst[c] = entries in columns < c, c = 0..32
template<typename VT>
__device__ void list_build(VT* rv, uint8_t* rl, int32_t* cst, const Dep& d, const Pre<VT>& p,
    int w)
{
    const int lane = lane_id();
    constexpr int NONE = 0x7fffffff;
    int key = NONE, r = 0;
    if(lane < d.nnz && p.v != (VT) 0) {
        r = p.ix / w;
        key = (p.ix - r * w) * 32 + r;
    }
    int rank = 0, below = 0;
    for(int f = 0; f < d.nnz; ++f) {
        const int kf = wave_bcast(key, f);
        rank += (kf < key);
        below += (kf != NONE) && ((kf >> 5) < lane);
    }
    if(key != NONE) {
        rl[rank] = (uint8_t) r;
        rv[rank] = p.v;
    }
    if(lane <= 32) cst[lane] = below;
}

// T[c][j] -= sum_l R[l][j] * Lt[c][l] over the listed rows l of column j,
// ascending, Lt as in product()
template<typename VT>
__device__ void product_list(VT (&T)[16], const VT* A, const VT* rv, const uint8_t* rl,
    const int32_t* cst, const VT* cs, const VT* cd, const VT* cb, int lh, int rh)
{
    const int lane = lane_id();
    const int j = lane & 31;
    const int c0 = lane >> 5;
    const int k0 = (j < rh) ? cst[j] : 0;
    const int k1 = (j < rh) ? cst[j + 1] : 0;
    if(k0 < k1) {
        // one output row at a time keeps a single accumulator live
#pragma unroll
        for(int t = 0; t < 16; ++t) {
            const int c = c0 + 2 * t;
            if(c < lh) {
                VT acc = 0;
                for(int k = k0; k < k1; ++k) {
                    const int l = rl[k];
                    const VT r = rv[k];
                    const VT left = (l == 0) ? (VT) 0 : A[32 * (l - 1) + c];
                    const VT mid = A[32 * l + c];
                    const VT right = (l == 31) ? (VT) 0 : A[32 * (l + 1) + c];
                    const VT lt = fma(cb[l], right, fma(cd[l], mid, cs[l] * left));
                    acc = fma(r, lt, acc);
                }
                T[t] -= acc;
            }
        }
    }
}

template<typename VT>
__device__ void offdiag_small_one(const FactorArgs<VT>& a, const LevelRange& lr, int q,
    OffLdsSmall<VT>& s)
{
    const int lane = lane_id();
    const int64_t p = (int64_t) lr.lo + q;
    const int4 t0 = ((const int4*) a.off_task)[3 * p];
    const int4 t1 = ((const int4*) a.off_task)[3 * p + 1];
    const int4 t2 = ((const int4*) a.off_task)[3 * p + 2];
    const int fr = t0.y;
    const int hgt = t0.z & 255, wid = (t0.z >> 8) & 255;
    const bool all1x1 = ((t0.z >> 16) & 1) != 0;
    const int pq = t0.w;
    const int np = t2.y;
    const int4* pairs = (const int4*) a.upd4 + pair_first(t2);
    const int j = lane & 31, c0 = lane >> 5;

    // input block A_ij into the register tile
    tile_zero(s.X);
    wave_sync();
    if(t1.x == FMT_DENSE) {
#pragma unroll
        for(int t = 0; t < 16; ++t) {
            const int idx = 64 * t + lane, r = idx >> 5, c = idx & 31;
            if(r < hgt && c < wid) s.X[idx] = a.val_in[(int64_t) t1.w + r * wid + c];
        }
    } else {
        for(int e = lane; e < t1.y; e += 64) {
            const int ix = a.ix_in[(int64_t) t1.z + e];
            const int r = ix / wid, c = ix - r * wid;
            s.X[32 * r + c] = a.val_in[(int64_t) t1.w + e];
        }
    }
    wave_sync();
    VT T[16];
#pragma unroll
    for(int t = 0; t < 16; ++t) T[t] = s.X[32 * (c0 + 2 * t) + j];
    wave_sync();
    tile_zero(s.X);
    wave_sync();

    // A_ij -= U_ki' D_k U_kj over the Schur pairs, in order
    for(int e0 = 0; e0 < np; e0 += 64) {
        Dep mi{}, mj{};
        if(e0 + lane < np) {
            const int4 pr = pairs[e0 + lane];
            mi = dep_meta(a.bout, pr.x, pr.z, pr.w);
            mj = dep_meta(a.bout, pr.y, pr.z, pr.w);
        }
        if(wave_ballot(e0 + lane < np && !small_sparse<VT>(mj))) {
            if(lane == 0) a.defer_list[atomicAdd(a.defer_cnt + lr.level, 1)] = q;
            return;
        }
        const int cnt = min(64, np - e0);
        Dep ci = bcast(mi, 0), cj = bcast(mj, 0);
        Pre<VT> pi = prefetch(ci, a.ix_out, a.val_out, a.tri, true);
        Pre<VT> pj = prefetch(cj, a.ix_out, a.val_out, a.tri, false);
        for(int d = 0; d < cnt; ++d) {
            Dep ni = ci, nj = cj;
            Pre<VT> qi = pi, qj = pj;
            if(d + 1 < cnt) {
                ni = bcast(mi, d + 1);
                nj = bcast(mj, d + 1);
                qi = prefetch(ni, a.ix_out, a.val_out, a.tri, true);
                qj = prefetch(nj, a.ix_out, a.val_out, a.tri, false);
            }
            stage(s.X, (uint32_t*) nullptr, ci, pi, hgt, a.ix_out, a.val_out);
            list_build(s.rv, s.rl, s.cst, cj, pj, wid);
            stage_d(s.cs, s.cd, s.cb, pi);
            wave_sync();
            product_list(T, s.X, s.rv, s.rl, s.cst, s.cs, s.cd, s.cb, hgt, wid);
            wave_sync();
            unstage(s.X, (uint32_t*) nullptr, ci, pi, hgt);
            wave_sync();
            ci = ni; cj = nj; pi = qi; pj = qj;
        }
    }

    // the updated tile back to LDS, row c, column j
#pragma unroll
    for(int t = 0; t < 16; ++t) s.X[32 * (c0 + 2 * t) + j] = T[t];
    wave_sync();

    // U_ii' forward substitution in place, column jc per lane (lanes >= 32
    // idle), over the nonzero columns of the pivot rows (exact no-ops skipped)
    const VT* piv = a.if_val + (int64_t) pq * 1024;
    const uint32_t um_row = (lane < hgt) ? a.piv_umask[(int64_t) pq * 32 + lane] : 0u;
    const int jc = lane;
    VT* col = s.X + jc;
    for(int ii = 0; ii < hgt; ++ii) {
        const uint32_t um = (uint32_t) __builtin_amdgcn_readlane((int) um_row, ii);
        if(um && jc < wid) {
            const VT rv = col[32 * ii];
            if(rv != (VT) 0) {
                const VT pv = rv / piv[ii * hgt + ii];
                for(uint32_t bits = um; bits; bits &= bits - 1) {
                    const int l = __ffs(bits) - 1;
                    col[32 * l] = fma(-piv[ii * hgt + l], pv, col[32 * l]);
                }
            }
        }
    }
    // D solve in place (d_left_solve_with_D_Ut, kernels :1533-1566); a row
    // is written only after both rows it reads have been read
    const double* D = a.tri + 3 * (int64_t) fr;
    const int32_t* ip = a.is_piv + fr;
    if(jc < wid) {
        if(all1x1) {
            for(int l = 0; l < hgt; ++l) col[32 * l] = (VT) ((double) col[32 * l] / D[3 * l + 1]);
        } else {
            for(int l = 0; l < hgt; ++l) {
                const bool p1 = ip[l] != 0;
                const bool p2 = (l + 1 < hgt) && !ip[l + 1];
                const VT rl = col[32 * l];
                const VT right = (l < 31) ? col[32 * (l + 1)] : rl;
                VT x1, x2;
                givens_d_solve(a.piv64 != 0, D + 3 * l, l, hgt, rl, right, p2, x1, x2);
                if(p1) col[32 * l] = x1;
                if(p2) col[32 * (l + 1)] = x2;
            }
        }
    }
    wave_sync();
    if(lane < hgt) a.rn_part[(int64_t) q * 32 + lane] = row_ssq(s.X, lane, wid);

    VT* oval = a.if_val + (int64_t) q * 1024;
    if(t1.x == FMT_DENSE) {
        for(int e = lane; e < hgt * wid; e += 64) {
            const int r = e / wid, c = e - r * wid;
            oval[e] = s.X[32 * r + c];
        }
        if(lane == 0) a.if_cnt[q] = hgt * wid;
    } else {
        uint16_t* oix = a.if_idx + (int64_t) q * 1024;
        int base = 0;
#pragma unroll
        for(int t = 0; t < 16; ++t) {
            const int idx = 64 * t + lane, r = idx >> 5, c = idx & 31;
            const VT v = s.X[idx];
            const bool pred = r < hgt && c < wid && fabs((double) v) > NZ_EPS;
            const uint64_t msk = wave_ballot(pred);
            if(pred) {
                const int pos = base + lanes_below(msk);
                oix[pos] = (uint16_t) idx;
                oval[pos] = v;
            }
            base += __popcll(msk);
        }
        if(lane == 0) a.if_cnt[q] = base;
    }
}

// the positions k_offdiag_tiny leaves to the wave form, grid-stride
template<typename VT>
__device__ __forceinline__ void offdiag_body(const FactorArgs<VT>& a, const LevelRange& lr,
    int bid_, int nblk_)
{
    __shared__ OffLdsSmall<VT> s;
    const int nw = a.wave_cnt[lr.level];
    for(int t = bid_; t < nw; t += nblk_) {
        offdiag_small_one(a, lr, a.wave_list[t], s);
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// One lane per off-diagonal block for small blocks (at level-0 fill of the
// KKT workloads: 1-2 entries each; a wave then works on 64 blocks instead of
// one). Static eligibility (task bit, set by the host): sparse input with at
// most TINY_E entries, at most 8 Schur pairs (16 at a level with 8192 or
// more off-diagonal blocks), 1x1 pivots in the block row.
// At run time every dependency must be sparse with at most TINY_E entries
// and the tile must stay within TINY_T nonzeros; otherwise the position goes
// to the wave form before anything is written. The tile is a list of (row * 32 + col, value) in LDS,
// one column of the list arrays per lane; every value is formed by the same
// operations in the same order as in the wave form (terms that are exact
// zeros in the wave form are skipped).
constexpr int TINY_E = 8;
constexpr int TINY_T = 32;

template<typename VT>
struct TinyLds {
    VT v[TINY_T][TINY_TPB];
    int16_t ix[TINY_T][TINY_TPB];
};

template<typename VT>
__device__ __forceinline__ void to_wave_form(const FactorArgs<VT>& a, const LevelRange& lr, int q)
{
    a.wave_list[atomicAdd(a.wave_cnt + lr.level, 1)] = q;
}

// metadata of the two factored blocks of one Schur pair, as far as the
// one-lane form needs it; ok = both are sparse with at most TINY_E entries
struct TinyMeta {
    bool ok;
    int nnz_i, cp_i, vp_i, nnz_j, cp_j, vp_j;
};

__device__ __forceinline__ TinyMeta tiny_meta(const BlockMeta* bout, int4 pp, bool valid)
{
    TinyMeta t{};
    t.ok = true;
    if(!valid) return t;
    const BlockMeta mi = bout[pp.x], mj = bout[pp.y];
    t.ok = mi.format == FMT_SPARSE && mi.nnz <= TINY_E && mj.format == FMT_SPARSE &&
        mj.nnz <= TINY_E;
    t.nnz_i = mi.nnz;
    t.cp_i = mi.col_ptr;
    t.vp_i = mi.val_ptr;
    t.nnz_j = mj.nnz;
    t.cp_j = mj.col_ptr;
    t.vp_j = mj.val_ptr;
    return t;
}

// the stored entries of both blocks (nothing for a pair outside the form)
template<typename VT>
struct TinyRaw {
    int ixi[TINY_E], ixj[TINY_E];
    VT vi[TINY_E], vj[TINY_E];
};

template<typename VT>
__device__ __forceinline__ void tiny_raw(const int32_t* ix_out, const VT* val_out,
    const TinyMeta& m, TinyRaw<VT>& r)
{
#pragma unroll
    for(int e = 0; e < TINY_E; ++e) {
        r.ixi[e] = 0; r.vi[e] = 0; r.ixj[e] = 0; r.vj[e] = 0;
        if(m.ok && e < m.nnz_i) {
            r.ixi[e] = ix_out[(int64_t) m.cp_i + e];
            r.vi[e] = val_out[(int64_t) m.vp_i + e];
        }
        if(m.ok && e < m.nnz_j) {
            r.ixj[e] = ix_out[(int64_t) m.cp_j + e];
            r.vj[e] = val_out[(int64_t) m.vp_j + e];
        }
    }
}

// One-lane form, first phase: the Schur updates of the off-diagonal block at
// level position q (thread tid of its workgroup), which read only blocks of
// earlier levels and so run beside the level's pivot blocks. The tile list
// is parked in the position's compact store (if_idx / if_val, count in
// if_cnt); if_cnt[q] = -1 marks a position left to the other kernels.
template<typename VT>
__device__ void tiny_pairs_one(const FactorArgs<VT>& a, const LevelRange& lr, const int q,
    const int tid, TinyLds<VT>& s)
{
    if(q >= lr.size) return;
    a.if_cnt[q] = -1;
    const int64_t p = (int64_t) lr.lo + q;
    const int4 t0 = ((const int4*) a.off_task)[3 * p];
    const int4 t1 = ((const int4*) a.off_task)[3 * p + 1];
    const int4 t2 = ((const int4*) a.off_task)[3 * p + 2];
    if(a.pivoting) {
        // rows in pivoting order: every block takes the general form
        a.defer_list[atomicAdd(a.defer_cnt + lr.level, 1)] = q;
        return;
    }
    if(!(t2.z & 1)) { to_wave_form(a, lr, q); return; }
    const int hgt = t0.z & 255, wid = (t0.z >> 8) & 255;
    const int np = t2.y;
    const int4* pairs = (const int4*) a.upd4 + pair_first(t2);
    VT* tv = &s.v[0][tid];
    int16_t* ti = &s.ix[0][tid];
    auto V = [&](int k) -> VT& { return tv[k * TINY_TPB]; };
    auto I = [&](int k) -> int16_t& { return ti[k * TINY_TPB]; };
    int n = 0;
    auto find = [&](int idx) {
        for(int k = 0; k < n; ++k)
            if(I(k) == idx) return k;
        return -1;
    };
    // input entries
    for(int e = 0; e < t1.y; ++e) {
        const int ix = a.ix_in[(int64_t) t1.z + e];
        const int r = ix / wid, c = ix - r * wid;
        I(n) = (int16_t) (32 * r + c);
        V(n) = a.val_in[(int64_t) t1.w + e];
        ++n;
    }
    // A_ij -= U_ki' D_k U_kj over the Schur pairs, in order. Loads run
    // ahead of the arithmetic: while pair pr is computed, the entries of
    // pair pr + 1, the metadata of pair pr + 2 and the pair record of pair
    // pr + 3 are in flight.
    auto rec = [&](int k) { return k < np ? pairs[k] : make_int4(0, 0, 0, 0); };
    int4 pc = rec(0), p1 = rec(1), p2 = rec(2);
    TinyMeta mc = tiny_meta(a.bout, pc, np > 0), m1 = tiny_meta(a.bout, p1, np > 1);
    TinyRaw<VT> rc;
    tiny_raw(a.ix_out, a.val_out, mc, rc);
    for(int pr = 0; pr < np; ++pr) {
        if(!mc.ok) {
            to_wave_form(a, lr, q);
            return;
        }
        // next entries, metadata and record
        TinyRaw<VT> rn;
        tiny_raw(a.ix_out, a.val_out, m1, rn);
        const TinyMeta m2 = tiny_meta(a.bout, p2, pr + 2 < np);
        const int4 p3 = rec(pr + 3);
        const int rb = pc.z;
        // U_ki: rows l of block k, columns c (rows of block i), width hgt
        int al[TINY_E], ac[TINY_E];
        VT av[TINY_E];
#pragma unroll
        for(int e = 0; e < TINY_E; ++e) {
            al[e] = -4; ac[e] = -1; av[e] = 0;
            if(e < mc.nnz_i) {
                const int ix = rc.ixi[e];
                al[e] = ix / hgt;
                ac[e] = ix - al[e] * hgt;
                av[e] = rc.vi[e];
            }
        }
        // U_kj: rows l, columns j, width wid; nonzero entries by (column, row)
        int rl[TINY_E], rj[TINY_E];
        VT rv[TINY_E];
#pragma unroll
        for(int e = 0; e < TINY_E; ++e) {
            rl[e] = 0; rj[e] = 64; rv[e] = 0;
            if(e < mc.nnz_j) {
                const int ix = rc.ixj[e];
                const VT v = rc.vj[e];
                if(v != (VT) 0) {
                    rl[e] = ix / wid;
                    rj[e] = ix - rl[e] * wid;
                    rv[e] = v;
                }
            }
        }
        const bool one_one = mc.nnz_i == 1 && mc.nnz_j == 1;
        rc = rn;
        mc = m1;
        m1 = m2;
        pc = p1;
        p1 = p2;
        p2 = p3;
        if(one_one) {
            // one entry on each side (nearly every pair at level-0 fill):
            // the general form below reduces to one product for the single
            // element (column of U_ki's entry, column of U_kj's entry), with
            // the same operations
            if(rj[0] < 64) {
                const int l = rl[0];
                const double* td = a.tri + 3 * (int64_t) (rb + l);
                const VT sup = (VT) td[0], dia = (VT) td[1], sub = (VT) td[2];
                const VT left = (al[0] == l - 1) ? av[0] : (VT) 0;
                const VT mid = (al[0] == l) ? av[0] : (VT) 0;
                const VT right = (al[0] == l + 1) ? av[0] : (VT) 0;
                const VT lt = fma(sub, right, fma(dia, mid, sup * left));
                const VT acc = fma(rv[0], lt, (VT) 0);
                const int idx = 32 * ac[0] + rj[0];
                int k = find(idx);
                if(k < 0) {
                    if(acc == (VT) 0) continue;
                    if(n == TINY_T) { to_wave_form(a, lr, q); return; }
                    k = n++;
                    I(k) = (int16_t) idx;
                    V(k) = 0;
                }
                V(k) -= acc;
            }
            continue;
        }
        int key[TINY_E], rank[TINY_E];
#pragma unroll
        for(int e = 0; e < TINY_E; ++e) key[e] = rj[e] * 32 + rl[e];
#pragma unroll
        for(int e = 0; e < TINY_E; ++e) {
            int rk = 0;
#pragma unroll
            for(int f = 0; f < TINY_E; ++f) rk += (key[f] < key[e]) || (key[f] == key[e] && f < e);
            rank[e] = rk;
        }
        // distinct columns of U_ki
        int dc[TINY_E];
#pragma unroll
        for(int e = 0; e < TINY_E; ++e) {
            bool seen = ac[e] < 0;
#pragma unroll
            for(int f = 0; f < e; ++f) seen = seen || (ac[f] == ac[e]);
            dc[e] = -1;
            if(!seen) dc[e] = ac[e];
        }
        // columns j in ascending order; within a column, rows ascending
        VT acc[TINY_E];
        int curj = -1;
        for(int rk = 0; rk <= TINY_E; ++rk) {
            int e = -1;
#pragma unroll
            for(int f = 0; f < TINY_E; ++f)
                if(rank[f] == rk) e = f;
            const int j = (rk < TINY_E && e >= 0) ? rj[e] : 64;
            if(j != curj && curj >= 0 && curj < 64) {
                // column curj done: T[c][curj] -= acc[c]
#pragma unroll
                for(int f = 0; f < TINY_E; ++f) {
                    if(dc[f] < 0) continue;
                    const int idx = 32 * dc[f] + curj;
                    int k = find(idx);
                    if(k < 0) {
                        if(acc[f] == (VT) 0) continue;
                        if(n == TINY_T) { to_wave_form(a, lr, q); return; }
                        k = n++;
                        I(k) = (int16_t) idx;
                        V(k) = 0;
                    }
                    V(k) -= acc[f];
                }
            }
            if(j >= 64) break;
            if(j != curj) {
                curj = j;
#pragma unroll
                for(int f = 0; f < TINY_E; ++f) acc[f] = 0;
            }
            const int l = rl[e];
            const VT r = rv[e];
            const double* td = a.tri + 3 * (int64_t) (rb + l);
            const VT sup = (VT) td[0], dia = (VT) td[1], sub = (VT) td[2];
#pragma unroll
            for(int f = 0; f < TINY_E; ++f) {
                if(dc[f] < 0) continue;
                const int c = dc[f];
                VT left = 0, mid = 0, right = 0;
#pragma unroll
                for(int g = 0; g < TINY_E; ++g) {
                    if(ac[g] != c) continue;
                    if(al[g] == l - 1) left = av[g];
                    if(al[g] == l) mid = av[g];
                    if(al[g] == l + 1) right = av[g];
                }
                const VT lt = fma(sub, right, fma(dia, mid, sup * left));
                acc[f] = fma(r, lt, acc[f]);
            }
        }
    }
    // the tile list for the second phase
    VT* lv = a.if_val + (int64_t) q * 1024;
    uint16_t* li = a.if_idx + (int64_t) q * 1024;
    for(int k = 0; k < n; ++k) {
        li[k] = (uint16_t) I(k);
        lv[k] = V(k);
    }
    a.if_cnt[q] = n;
}

// One-lane form, second phase (after the level's pivot blocks): the tile
// list of the first phase through U_ii' and D, then the row partial sums and
// the compact entries.
template<typename VT>
__device__ __forceinline__ void offdiag_tiny_scale_body(const FactorArgs<VT>& a, const LevelRange& lr,
    int bid_, int nblk_)
{
    __shared__ TinyLds<VT> s;
    const int tid = threadIdx.x;
    const int q = lr.npiv + bid_ * TINY_TPB + tid;
    if(q >= lr.size) return;
    // the parked count and the task record together (neither waits for
    // the other)
    const int64_t p = (int64_t) lr.lo + q;
    const int4 t0 = ((const int4*) a.off_task)[3 * p];
    int n = a.if_cnt[q];
    if(n < 0) return;
    const int fr = t0.y;
    const int hgt = t0.z & 255;
    const int pq = t0.w;
    // a pivot block without entries right of its diagonal (nearly all at
    // level-0 fill) leaves the forward substitution with nothing to do
    const uint32_t um_any = a.piv_any[pq];
    VT* tv = &s.v[0][tid];
    int16_t* ti = &s.ix[0][tid];
    auto V = [&](int k) -> VT& { return tv[k * TINY_TPB]; };
    auto I = [&](int k) -> int16_t& { return ti[k * TINY_TPB]; };
    auto find = [&](int idx) {
        for(int k = 0; k < n; ++k)
            if(I(k) == idx) return k;
        return -1;
    };
    {
        const VT* lv = a.if_val + (int64_t) q * 1024;
        const uint16_t* li = a.if_idx + (int64_t) q * 1024;
        for(int k = 0; k < n; ++k) {
            I(k) = (int16_t) li[k];
            V(k) = lv[k];
        }
    }
    // U_ii' forward substitution over the rows in ascending order
    const VT* piv = a.if_val + (int64_t) pq * 1024;
    for(int ii = 0; ii < hgt && um_any != 0; ++ii) {
        bool any = false;
        for(int k = 0; k < n; ++k) any = any || ((I(k) >> 5) == ii && V(k) != (VT) 0);
        if(!any) continue;
        const uint32_t um = a.piv_umask[(int64_t) pq * 32 + ii];
        if(!um) continue;
        const VT d = piv[ii * hgt + ii];
        const int n0 = n;
        for(int k = 0; k < n0; ++k) {
            if((I(k) >> 5) != ii || V(k) == (VT) 0) continue;
            const int jc = I(k) & 31;
            const VT pv = V(k) / d;
            for(uint32_t bits = um; bits; bits &= bits - 1) {
                const int l = __ffs(bits) - 1;
                const int idx = 32 * l + jc;
                int m = find(idx);
                if(m < 0) {
                    if(n == TINY_T) { to_wave_form(a, lr, q); return; }
                    m = n++;
                    I(m) = (int16_t) idx;
                    V(m) = 0;
                }
                V(m) = fma(-piv[ii * hgt + l], pv, V(m));
            }
        }
    }
    // D solve (1x1 pivots): the Givens-form solve reduces to a division;
    // the divisors of eight entries are loaded together
    const double* D = a.tri + 3 * (int64_t) fr;
    for(int k0 = 0; k0 < n; k0 += 8) {
        double dv[8];
#pragma unroll
        for(int u = 0; u < 8; ++u) dv[u] = D[3 * (I(min(k0 + u, n - 1)) >> 5) + 1];
#pragma unroll
        for(int u = 0; u < 8; ++u)
            if(k0 + u < n) V(k0 + u) = (VT) ((double) V(k0 + u) / dv[u]);
    }
    // tile order (row-major)
    for(int k = 1; k < n; ++k) {
        const int16_t ik = I(k);
        const VT vk = V(k);
        int m = k - 1;
        while(m >= 0 && I(m) > ik) {
            I(m + 1) = I(m);
            V(m + 1) = V(m);
            --m;
        }
        I(m + 1) = ik;
        V(m + 1) = vk;
    }
    // row partial sums of squares (columns ascending) of the rows with
    // entries; the others stay at the zero the first phase stored (a row
    // without entries sums to +0)
    VT* rp = a.rn_part + (int64_t) q * 32;
    int k = 0;
    while(k < n) {
        const int r = I(k) >> 5;
        VT ssq = 0;
        while(k < n && (I(k) >> 5) == r) {
            ssq = fma(V(k), V(k), ssq);
            ++k;
        }
        rp[r] = ssq;
    }
    // and the compact list
    VT* oval = a.if_val + (int64_t) q * 1024;
    uint16_t* oix = a.if_idx + (int64_t) q * 1024;
    int cnt = 0;
    for(int m = 0; m < n; ++m) {
        const VT v = V(m);
        if(fabs((double) v) > NZ_EPS) {
            oix[cnt] = (uint16_t) I(m);
            oval[cnt] = v;
            ++cnt;
        }
    }
    a.if_cnt[q] = cnt;
}
// The level's pivot blocks (workgroups 0 .. npiv - 1) and the first phase
// of the one-lane off-diagonal form (the rest, 64 positions each) in one
// launch: the two read disjoint data and never wait for each other.
template<typename VT>
union PivPairsLds {
    PivLds<VT> piv;
    TinyLds<VT> tiny;
};

template<typename VT>
__device__ __forceinline__ void pivots_pairs_body(const FactorArgs<VT>& a, const LevelRange& lr,
    int bid_, int nblk_)
{
    __shared__ PivPairsLds<VT> u;
    const int w = bid_;
    if(w < lr.npiv) {
        pivots_one<VT>(a, lr, w, u.piv);
        return;
    }
    // the workgroup's positions' row partial sums start at zero, stored
    // coalesced here (the second phase writes only the rows with entries;
    // the other off-diagonal kernels write every row of theirs)
    const int q0 = lr.npiv + (w - lr.npiv) * TINY_TPB;
    const int q1 = min(q0 + TINY_TPB, lr.size);
    VT* rp = a.rn_part + (int64_t) q0 * 32;
    for(int e = (int) threadIdx.x; e < (q1 - q0) * 32; e += TINY_TPB) rp[e] = (VT) 0;
    tiny_pairs_one<VT>(a, lr, q0 + (int) threadIdx.x, (int) threadIdx.x, u.tiny);
}


template<typename VT>
__device__ __forceinline__ VT drop_norm(VT rn) { return (VT) sqrtf((float) rn); }

// row sums of squares: pivot block first, then the off-diagonal blocks of
// the row by block id (contiguous in the level list at row_first_block_offset)
template<typename VT>
__device__ __forceinline__ void rownorm_body(const FactorArgs<VT>& a, const LevelRange& lr,
    int bid_, int nblk_)
{
    const int lane = lane_id();
    const int b = a.lvlb_idx[lr.lo + bid_];
    const int i = a.bin[b].block_row;
    const int fr = a.bs[i];
    const int h = a.bs[i + 1] - fr;
    const int nblk = a.csr_off[i + 1] - a.csr_off[i] - 1;
    if(lane >= h) return;
    VT acc = a.rn[fr + lane];
    if(nblk > 0) {
        const VT* part = a.rn_part + (int64_t) (lr.npiv + a.rfo[i]) * 32 + lane;
        // sixteen loads in flight per pass; the sum stays in block order
        int t = 0;
        for(; t + 16 <= nblk; t += 16) {
            VT v[16];
#pragma unroll
            for(int u = 0; u < 16; ++u) v[u] = part[32 * (t + u)];
#pragma unroll
            for(int u = 0; u < 16; ++u) acc += v[u];
        }
        for(; t < nblk; ++t) acc += part[32 * t];
    }
    a.rn[fr + lane] = acc;
}

// Leftover pool, granted extra space and output offsets of a level's blocks
// in level-list order (kernels :1948-2008 with the atomics serialised in
// that order). The pool after a sparse block is max(pool + max_nnz - kept,
// 0); such maps p -> max(p + a, b) compose associatively, so the pool
// sequence is a scan (k_level_scan): each chunk's composed map, then the
// pools and grants, then the output offsets from the chunks' size sums.
struct PoolMap {
    int64_t a, b;
    __device__ static PoolMap shfl_up(const PoolMap& m, int o)
    {
        return {__shfl_up(m.a, o, WAVE), __shfl_up(m.b, o, WAVE)};
    }
};
struct MapOp {
    // first x, then y
    __device__ PoolMap operator()(const PoolMap& x, const PoolMap& y) const
    {
        const int64_t nb = x.b + y.a;
        return {x.a + y.a, nb > y.b ? nb : y.b};
    }
};
__device__ __forceinline__ int64_t map_apply(const PoolMap& m, int64_t p)
{
    const int64_t v = p + m.a;
    return v > m.b ? v : m.b;
}
struct Sizes {
    int64_t x, v;   // index and value entries
    __device__ static Sizes shfl_up(const Sizes& s, int o)
    {
        return {__shfl_up(s.x, o, WAVE), __shfl_up(s.v, o, WAVE)};
    }
};
struct SumOp {
    __device__ Sizes operator()(const Sizes& p, const Sizes& q) const { return {p.x + q.x, p.v + q.v}; }
};

constexpr int SC_TPB = 256;
constexpr int SC_IPT = SCAN_CHUNK / SC_TPB;
constexpr int SC_NW = SC_TPB / WAVE;
constexpr int32_t CAP_NONE = -0x7fffffff;

// in-order combination of load(0 .. count-1) over the workgroup
template<typename V, typename Op, typename Load>
__device__ V block_prefix(int count, Load load, Op op, const V& id, V* sh)
{
    const int per = (count + SC_TPB - 1) / SC_TPB;
    V v = id;
    for(int k = 0; k < per; ++k) {
        const int i = (int) threadIdx.x * per + k;
        if(i < count) v = op(v, load(i));
    }
    V tot;
    (void) block_exclusive_scan<V, Op, SC_NW>(v, op, id, sh, &tot);
    return tot;
}

// One launch per level, in chunks of SCAN_CHUNK positions. A workgroup
// takes its chunk from a ticket counter, so every chunk it later waits for
// belongs to a workgroup that is already running, whatever the dispatch
// order. Every chunk runs the threshold counts, publishes its pool map
// (release store of its flag), waits for the maps of the chunks before it
// (acquire loads), runs the pools and grants, publishes its sizes, waits
// for the sizes before it and writes the factored metadata. The waits are
// bounded only as a guard against a hang; given the tickets they end as
// soon as the earlier chunks have published.
constexpr int SCAN_SPIN_LIMIT = 1 << 22;

__device__ __forceinline__ int64_t load_published(const int64_t* p)
{
    return __hip_atomic_load(const_cast<int64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the first wave polls the flags of chunks [0, c); then the workgroup goes
// on together. A wait that gives up records 2 in *err and goes on.
__device__ void wait_chunk_flags(const int32_t* flags, int c, int32_t* err)
{
    if(threadIdx.x < WAVE) {
        int spins = 0;
        for(;;) {
            bool set = true;
            for(int i = (int) threadIdx.x; i < c; i += WAVE)
                set = set && __hip_atomic_load(const_cast<int32_t*>(flags + i), __ATOMIC_ACQUIRE,
                    __HIP_MEMORY_SCOPE_AGENT) != 0;
            if(wave_ballot(!set) == 0ull) break;
            __builtin_amdgcn_s_sleep(1);
            if(++spins > SCAN_SPIN_LIMIT) {
                if(threadIdx.x == 0 && err)
                    __hip_atomic_store(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
        }
    }
    __syncthreads();
}

template<typename VT>
__device__ __forceinline__ void level_scan_body(const FactorArgs<VT>& a, const LevelRange& lr,
    int bid_, int nblk_)
{
    __shared__ PoolMap shm[SC_NW];
    __shared__ Sizes shs[SC_NW];
    __shared__ int ticket;
    int32_t* fmap = a.scan_flags + (int64_t) lr.level * (2 * a.scan_maxch + 1);
    int32_t* fsize = fmap + a.scan_maxch;
    if(threadIdx.x == 0) ticket = atomicAdd(fsize + a.scan_maxch, 1);
    __syncthreads();
    const int c = ticket;
    const int32_t* cin = a.carry + 4 * (lr.level & 1);
    int32_t* cout = a.carry + 4 * ((lr.level + 1) & 1);
    // -- entries of a sparse position that survive |v| >= tau * sqrtf(norm)
    // (kernels :1919-1946); the thread's positions are read in three rounds
    // of independent loads (capacity / first row / entry count, then each
    // block's first entry, then that entry's row norm) instead of one
    // dependent chain per position; further entries (rare at level-0 fill)
    // are read in order
    int capk[SC_IPT], frk[SC_IPT], cntk[SC_IPT];
#pragma unroll
    for(int k = 0; k < SC_IPT; ++k) {
        const int q = c * SCAN_CHUNK + (int) threadIdx.x * SC_IPT + k;
        capk[k] = CAP_NONE;
        frk[k] = 0;
        cntk[k] = 0;
        if(q < lr.size) {
            capk[k] = a.lq_cap[lr.lo + q];
            frk[k] = a.lq_fr[lr.lo + q];
            cntk[k] = a.if_cnt[q];
        }
    }
    int ix0[SC_IPT];
    VT v0[SC_IPT];
#pragma unroll
    for(int k = 0; k < SC_IPT; ++k) {
        const int q = c * SCAN_CHUNK + (int) threadIdx.x * SC_IPT + k;
        ix0[k] = 0;
        v0[k] = (VT) 0;
        if(capk[k] >= 0 && cntk[k] > 0) {
            ix0[k] = a.if_idx[(int64_t) q * 1024];
            v0[k] = a.if_val[(int64_t) q * 1024];
        }
    }
    VT n0[SC_IPT];
#pragma unroll
    for(int k = 0; k < SC_IPT; ++k) {
        n0[k] = (VT) 0;
        if(capk[k] >= 0 && cntk[k] > 0) n0[k] = a.rn[frk[k] + (ix0[k] >> 5)];
    }
    int dcnt[SC_IPT];
    PoolMap m{0, 0};
#pragma unroll
    for(int k = 0; k < SC_IPT; ++k) {
        const int q = c * SCAN_CHUNK + (int) threadIdx.x * SC_IPT + k;
        dcnt[k] = 0;
        if(capk[k] >= 0) {
            int keep = 0;
            if(cntk[k] > 0) {
                keep = !(fabs(v0[k]) < a.threshold * drop_norm(n0[k]));
                const VT* vals = a.if_val + (int64_t) q * 1024;
                const uint16_t* ixs = a.if_idx + (int64_t) q * 1024;
                for(int e = 1; e < cntk[k]; ++e) {
                    const VT nrm = drop_norm(a.rn[frk[k] + (ixs[e] >> 5)]);
                    keep += !(fabs(vals[e]) < a.threshold * nrm);
                }
            }
            a.drop_cnt[q] = keep;
            dcnt[k] = keep;
            m = MapOp()(m, PoolMap{(int64_t) capk[k] - keep, 0});
        }
    }
    PoolMap mtot;
    const PoolMap mex = block_exclusive_scan<PoolMap, MapOp, SC_NW>(m, MapOp(), PoolMap{0, 0}, shm,
        &mtot);
    if(threadIdx.x == 0) {
        a.chunk[4 * c + 0] = mtot.a;
        a.chunk[4 * c + 1] = mtot.b;
        __hip_atomic_store(fmap + c, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    wait_chunk_flags(fmap, c, a.err);
    // -- pools, grants and kept counts
    const int64_t* ch = a.chunk;
    const PoolMap before = block_prefix<PoolMap>(c,
        [ch](int i) { return PoolMap{load_published(ch + 4 * i), load_published(ch + 4 * i + 1)}; },
        MapOp(), PoolMap{0, 0}, shm);
    const int64_t pool_c = map_apply(before, cin[0]);
    int64_t pl = map_apply(mex, pool_c);
    int kp[SC_IPT], gr[SC_IPT];
    Sizes sz{0, 0};
#pragma unroll
    for(int k = 0; k < SC_IPT; ++k) {
        const int q = c * SCAN_CHUNK + (int) threadIdx.x * SC_IPT + k;
        kp[k] = 0;
        gr[k] = 0;
        if(capk[k] >= 0) {
            int64_t g = 0;
            if(dcnt[k] > capk[k]) {
                const int64_t e = (int64_t) dcnt[k] - capk[k];
                g = e < pl ? e : pl;
            }
            const int64_t room = capk[k] + g;
            kp[k] = (int32_t) (dcnt[k] < room ? dcnt[k] : room);
            gr[k] = (int32_t) g;
            const int64_t np = pl + capk[k] - dcnt[k];
            pl = np > 0 ? np : 0;
            a.keep[q] = kp[k];
            a.granted[q] = gr[k];
            sz.x += kp[k];
            sz.v += kp[k];
        } else if(capk[k] != CAP_NONE) {
            sz.v += -1 - capk[k];
        }
    }
    Sizes stot;
    const Sizes sex = block_exclusive_scan<Sizes, SumOp, SC_NW>(sz, SumOp(), Sizes{0, 0}, shs, &stot);
    if(threadIdx.x == 0) {
        a.chunk[4 * c + 2] = stot.x;
        a.chunk[4 * c + 3] = stot.v;
        __hip_atomic_store(fsize + c, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    wait_chunk_flags(fsize, c, a.err);
    // -- output offsets and factored metadata
    const Sizes sbefore = block_prefix<Sizes>(c,
        [ch](int i) { return Sizes{load_published(ch + 4 * i + 2), load_published(ch + 4 * i + 3)}; },
        SumOp(), Sizes{0, 0}, shs);
    int64_t px = cin[1] + sbefore.x + sex.x, pv = cin[2] + sbefore.v + sex.v;
#pragma unroll
    for(int k = 0; k < SC_IPT; ++k) {
        if(capk[k] == CAP_NONE) continue;
        const int q = c * SCAN_CHUNK + (int) threadIdx.x * SC_IPT + k;
        const int b = a.lvlb_idx[lr.lo + q];
        BlockMeta o = a.bin[b];
        int64_t nv;
        if(capk[k] >= 0) {
            o.nnz = kp[k];
            o.max_nnz = capk[k] + gr[k];
            if(dcnt[k] > capk[k]) o.got_extra = gr[k];
            a.overflow[q] = (uint8_t) (dcnt[k] > capk[k] + gr[k]);
            nv = kp[k];
        } else {
            o.nnz = 0;
            a.overflow[q] = 0;
            nv = -1 - capk[k];
        }
        o.col_ptr = (int32_t) px;
        o.val_ptr = (int32_t) pv;
        px += (capk[k] >= 0) ? kp[k] : 0;
        pv += nv;
        a.bout[o.id] = o;
    }
    if(threadIdx.x == 0 && c == nblk_ - 1) {
        cout[0] = (int32_t) map_apply(mtot, pool_c);
        cout[1] = (int32_t) (cin[1] + sbefore.x + stot.x);
        cout[2] = (int32_t) (cin[2] + sbefore.v + stot.v);
    }
}

// final compacted storage (kernels :2011-2073, d_save_block :85-173), one
// wave per block
template<typename VT>
__device__ void store_one(const FactorArgs<VT>& a, const LevelRange& lr, int q, VT* nrm, VT* key,
    uint16_t* kix)
{
    const int lane = lane_id();
    const int b = a.lvlb_idx[lr.lo + q];
    const BlockMeta mo = a.bout[b];
    const int fr = a.bs[mo.block_row];
    const int h = a.bs[mo.block_row + 1] - fr;
    const int w = a.bs[mo.block_col + 1] - a.bs[mo.block_col];
    const VT* vals = a.if_val + (int64_t) q * 1024;
    VT* out_v = a.val_out + mo.val_ptr;
    if(mo.format == FMT_DENSE) {
        for(int e = lane; e < h * w; e += 64) out_v[e] = vals[e];
        return;
    }
    const uint16_t* ixs = a.if_idx + (int64_t) q * 1024;
    int32_t* out_i = a.ix_out + mo.col_ptr;
    if(lane < h) nrm[lane] = drop_norm(a.rn[fr + lane]);
    wave_sync();
    const int cnt = a.if_cnt[q];
    if(!a.overflow[q]) {
        int base = 0;
        for(int e0 = 0; e0 < cnt; e0 += 64) {
            const int e = e0 + lane;
            bool k = false;
            VT v = 0;
            int ix = 0;
            if(e < cnt) {
                v = vals[e];
                ix = ixs[e];
                k = !(fabs(v) < a.threshold * nrm[ix >> 5]);
            }
            const uint64_t mask = wave_ballot(k);
            if(k) {
                const int pos = base + lanes_below(mask);
                out_i[pos] = (ix >> 5) * w + (ix & 31);
                out_v[pos] = v;
            }
            base += __popcll(mask);
        }
        return;
    }
    // over capacity: keep the max_nnz largest |v|, ties to the lower tile
    // index, written in that order (CUB BlockRadixSort descending, stable)
    for(int e = lane; e < cnt; e += 64) {
        const VT v = vals[e];
        const int ix = ixs[e];
        const bool k = !(fabs(v) < a.threshold * nrm[ix >> 5]);
        key[e] = k ? (VT) fabs(v) : (VT) -1;
        kix[e] = (uint16_t) ix;
    }
    wave_sync();
    for(int e = lane; e < cnt; e += 64) {
        const VT ke = key[e];
        if(ke < (VT) 0) continue;
        const int ie = kix[e];
        int rank = 0;
        for(int f = 0; f < cnt; ++f) {
            const VT kf = key[f];
            rank += (kf > ke) || (kf == ke && (int) kix[f] < ie);
        }
        if(rank < mo.max_nnz) {
            out_i[rank] = (ie >> 5) * w + (ie & 31);
            out_v[rank] = vals[e];
        }
    }
}

// the positions k_store_tiny leaves to the wave form, grid-stride
template<typename VT>
__device__ __forceinline__ void store_body(const FactorArgs<VT>& a, const LevelRange& lr,
    int bid_, int nblk_)
{
    __shared__ VT nrm[32];
    __shared__ VT key[1024];
    __shared__ uint16_t kix[1024];
    const int ns = a.store_cnt[lr.level];
    for(int t = bid_; t < ns; t += nblk_) {
        store_one(a, lr, a.store_list[t], nrm, key, kix);
        wave_sync();
    }
}

// one lane per sparse block with at most STORE_TINY pre-drop entries, the
// same selection and order as store_one; the rest go to the wave form
constexpr int STORE_TINY = 16;

template<typename VT>
__device__ __forceinline__ void store_tiny_body(const FactorArgs<VT>& a, const LevelRange& lr,
    int bid_, int nblk_)
{
    const int q = bid_ * 256 + (int) threadIdx.x;
    if(q >= lr.size) return;
    const int cap = a.lq_cap[lr.lo + q];
    const int cnt = a.if_cnt[q];
    if(cap < 0 || cnt > STORE_TINY) {
        a.store_list[atomicAdd(a.store_cnt + lr.level, 1)] = q;
        return;
    }
    const int b = a.lvlb_idx[lr.lo + q];
    const BlockMeta mo = a.bout[b];
    const int fr = a.lq_fr[lr.lo + q];
    const int w = a.bs[mo.block_col + 1] - a.bs[mo.block_col];
    const VT* vals = a.if_val + (int64_t) q * 1024;
    const uint16_t* ixs = a.if_idx + (int64_t) q * 1024;
    VT* out_v = a.val_out + mo.val_ptr;
    int32_t* out_i = a.ix_out + mo.col_ptr;
    VT key[STORE_TINY];
    int ixk[STORE_TINY];
#pragma unroll
    for(int e = 0; e < STORE_TINY; ++e) {
        key[e] = (VT) -1;
        ixk[e] = 0;
        if(e < cnt) {
            const VT v = vals[e];
            const int ix = ixs[e];
            const bool k = !(fabs(v) < a.threshold * drop_norm(a.rn[fr + (ix >> 5)]));
            key[e] = k ? (VT) fabs(v) : (VT) -1;
            ixk[e] = ix;
        }
    }
    if(!a.overflow[q]) {
        int pos = 0;
#pragma unroll
        for(int e = 0; e < STORE_TINY; ++e) {
            if(key[e] >= (VT) 0) {
                out_i[pos] = (ixk[e] >> 5) * w + (ixk[e] & 31);
                out_v[pos] = vals[e];
                ++pos;
            }
        }
        return;
    }
#pragma unroll
    for(int e = 0; e < STORE_TINY; ++e) {
        if(key[e] < (VT) 0) continue;
        int rank = 0;
#pragma unroll
        for(int f = 0; f < STORE_TINY; ++f)
            rank += (f < cnt) && ((key[f] > key[e]) || (key[f] == key[e] && ixk[f] < ixk[e]));
        if(rank < mo.max_nnz) {
            out_i[rank] = (ixk[e] >> 5) * w + (ixk[e] & 31);
            out_v[rank] = vals[e];
        }
    }
}

// k_pivot_col_blocks (kernels :1352-1418) with d_col_pivot_block
// (:1290-1348): one wave per block row of the level; the stored blocks above
// its pivot block (earlier levels, final storage) take its pivoting order in
// their columns: sparse entries get new column indices in place, dense rows
// are gathered column by column.
// the single-system kernels: the bodies at their own grid position
template<typename VT>
__global__ __launch_bounds__(64) void k_offdiag_general(FactorArgs<VT> a, LevelRange lr)
{
    offdiag_general_body<VT>(a, lr, blockIdx.x, gridDim.x);
}

template<typename VT>
__global__ __launch_bounds__(64, 3) void k_offdiag(FactorArgs<VT> a, LevelRange lr)
{
    offdiag_body<VT>(a, lr, blockIdx.x, gridDim.x);
}

template<typename VT>
__global__ __launch_bounds__(TINY_TPB) void k_offdiag_tiny_scale(FactorArgs<VT> a, LevelRange lr)
{
    offdiag_tiny_scale_body<VT>(a, lr, blockIdx.x, gridDim.x);
}

template<typename VT>
__global__ __launch_bounds__(64) void k_pivots_pairs(FactorArgs<VT> a, LevelRange lr)
{
    pivots_pairs_body<VT>(a, lr, blockIdx.x, gridDim.x);
}

template<typename VT>
__global__ __launch_bounds__(64) void k_rownorm(FactorArgs<VT> a, LevelRange lr)
{
    rownorm_body<VT>(a, lr, blockIdx.x, gridDim.x);
}

template<typename VT>
__global__ __launch_bounds__(SC_TPB) void k_level_scan(FactorArgs<VT> a, LevelRange lr)
{
    level_scan_body<VT>(a, lr, blockIdx.x, gridDim.x);
}

template<typename VT>
__global__ __launch_bounds__(64) void k_store(FactorArgs<VT> a, LevelRange lr)
{
    store_body<VT>(a, lr, blockIdx.x, gridDim.x);
}

template<typename VT>
__global__ __launch_bounds__(256) void k_store_tiny(FactorArgs<VT> a, LevelRange lr)
{
    store_tiny_body<VT>(a, lr, blockIdx.x, gridDim.x);
}

template<typename VT>
__global__ __launch_bounds__(64) void k_pivot_cols(FactorArgs<VT> a, LevelRange lr)
{
    __shared__ int inv[32];
    const int lane = lane_id();
    const int i = a.lvl_rows[lr.row_lo + blockIdx.x];
    const int fr = a.bs[i];
    const int w = a.bs[i + 1] - fr;
    const int lp = (lane < w) ? a.perm[fr + lane] - fr : lane;
    if(lane < w) inv[lp] = lane;
    if(wave_ballot(lane < w && lp != lane) == 0) return;
    wave_sync();
    for(int ci = a.csc_off[i]; ci < a.csc_off[i + 1] - 1; ++ci) {
        const BlockMeta m = a.bout[a.csc_idx[ci]];
        if(m.format == FMT_SPARSE) {
            int32_t* ix = a.ix_out + m.col_ptr;
            for(int e = lane; e < m.nnz; e += 64) {
                const int v = ix[e];
                ix[e] = (v / w) * w + inv[v % w];
            }
        } else {
            const int h = a.bs[m.block_row + 1] - a.bs[m.block_row];
            VT* v = a.val_out + m.val_ptr;
            for(int r = 0; r < h; ++r) {
                const VT x = (lane < w) ? v[(int64_t) r * w + lp] : (VT) 0;
                wave_sync();
                if(lane < w) v[(int64_t) r * w + lane] = x;
                wave_sync();
            }
        }
    }
}

template<typename VT>
__global__ void k_zero_carry(int32_t* carry)
{
    if(threadIdx.x < 8) carry[threadIdx.x] = 0;
}

template<typename VT>
__global__ void k_gather(const double* vals, const int64_t* slot, int64_t nnz, VT* val_in)
{
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if(q < nnz && slot[q] >= 0) val_in[slot[q]] = (VT) vals[q];
}

template<typename VT>
__global__ void k_zero(VT* a, int64_t n)
{
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if(q < n) a[q] = (VT) 0;
}

// ---------------------------------------------------------------------------
// Batched refactor: one launch per step covers every listed system; item
// {s, q} is workgroup q of the single-system launch on system s (the same
// grid position and grid size, so the same results).

// factor_reset_pivots, the zeroing of gather_values and factor_begin
template<typename VT>
__global__ void bk_factor_begin(const FactorSys<VT>* fs, const int2* items)
{
    const int2 it = items[blockIdx.x];
    const FactorSys<VT>& F = fs[it.x];
    const FactorArgs<VT>& a = F.a;
    const int64_t g = (int64_t) it.y * 256 + threadIdx.x;
    if(g < a.n) {
        a.is_piv[g] = F.static_is_piv[g];
        a.perm[g] = (int32_t) g;
    }
    if(g < F.val_space) F.val_in[g] = (VT) 0;
    const int64_t nf = (a.num_levels > 0 && a.scan_maxch > 0)
        ? (int64_t) (2 * a.scan_maxch + 1) * a.num_levels : 0;
    if(g < nf) a.scan_flags[g] = 0;
    if(g < a.num_levels) {
        a.defer_cnt[g] = 0;
        a.wave_cnt[g] = 0;
        a.store_cnt[g] = 0;
    }
    if(it.y == 0 && threadIdx.x < 8) a.carry[threadIdx.x] = 0;
}

// the gather of gather_values
template<typename VT>
__global__ void bk_gather(const FactorSys<VT>* fs, const int2* items)
{
    const int2 it = items[blockIdx.x];
    const FactorSys<VT>& F = fs[it.x];
    const int64_t q = (int64_t) it.y * 256 + threadIdx.x;
    if(q < F.nnz && F.value_slot[q] >= 0) F.val_in[F.value_slot[q]] = (VT) F.vals[q];
}

template<typename VT>
__global__ __launch_bounds__(64) void bk_pivots_pairs(const FactorSys<VT>* fs, const int2* items,
    int32_t level)
{
    const int2 it = items[blockIdx.x];
    const FactorSys<VT>& F = fs[it.x];
    const LevelRange lr = F.lr[level];
    pivots_pairs_body<VT>(F.a, lr, it.y, fl_grid_pivots_pairs(lr));
}

template<typename VT>
__global__ __launch_bounds__(TINY_TPB) void bk_offdiag_tiny_scale(const FactorSys<VT>* fs,
    const int2* items, int32_t level)
{
    const int2 it = items[blockIdx.x];
    const FactorSys<VT>& F = fs[it.x];
    const LevelRange lr = F.lr[level];
    offdiag_tiny_scale_body<VT>(F.a, lr, it.y, fl_tiny_blocks(lr));
}

template<typename VT>
__global__ __launch_bounds__(64, 3) void bk_offdiag(const FactorSys<VT>* fs, const int2* items,
    int32_t level)
{
    const int2 it = items[blockIdx.x];
    const FactorSys<VT>& F = fs[it.x];
    const LevelRange lr = F.lr[level];
    offdiag_body<VT>(F.a, lr, it.y, fl_grid_tasks(lr.size - lr.npiv, BATCH_TASK_GRID));
}

template<typename VT>
__global__ __launch_bounds__(64) void bk_offdiag_general(const FactorSys<VT>* fs,
    const int2* items, int32_t level)
{
    const int2 it = items[blockIdx.x];
    const FactorSys<VT>& F = fs[it.x];
    const LevelRange lr = F.lr[level];
    offdiag_general_body<VT>(F.a, lr, it.y, fl_grid_tasks(lr.size - lr.npiv, BATCH_TASK_GRID));
}

template<typename VT>
__global__ __launch_bounds__(64) void bk_rownorm(const FactorSys<VT>* fs, const int2* items,
    int32_t level)
{
    const int2 it = items[blockIdx.x];
    const FactorSys<VT>& F = fs[it.x];
    const LevelRange lr = F.lr[level];
    rownorm_body<VT>(F.a, lr, it.y, lr.npiv);
}

template<typename VT>
__global__ __launch_bounds__(SC_TPB) void bk_level_scan(const FactorSys<VT>* fs, const int2* items,
    int32_t level)
{
    const int2 it = items[blockIdx.x];
    const FactorSys<VT>& F = fs[it.x];
    const LevelRange lr = F.lr[level];
    level_scan_body<VT>(F.a, lr, it.y, fl_grid_level_scan(lr));
}

template<typename VT>
__global__ __launch_bounds__(256) void bk_store_tiny(const FactorSys<VT>* fs, const int2* items,
    int32_t level)
{
    const int2 it = items[blockIdx.x];
    const FactorSys<VT>& F = fs[it.x];
    const LevelRange lr = F.lr[level];
    store_tiny_body<VT>(F.a, lr, it.y, fl_grid_store_tiny(lr));
}

template<typename VT>
__global__ __launch_bounds__(64) void bk_store(const FactorSys<VT>* fs, const int2* items,
    int32_t level)
{
    const int2 it = items[blockIdx.x];
    const FactorSys<VT>& F = fs[it.x];
    const LevelRange lr = F.lr[level];
    store_body<VT>(F.a, lr, it.y, fl_grid_tasks(lr.size, BATCH_TASK_GRID));
}

}  // namespace

__global__ void k_reset_pivots(int32_t n, const int32_t* sp, int32_t* is_piv, int32_t* perm)
{
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if(g < n) {
        is_piv[g] = sp[g];
        perm[g] = g;
    }
}

void factor_reset_pivots(int32_t n, const int32_t* static_is_piv, int32_t* is_piv, int32_t* perm,
    void* stream)
{
    if(n > 0)
        k_reset_pivots<<<(n + 255) / 256, 256, 0, (hipStream_t) stream>>>(n, static_is_piv, is_piv,
            perm);
}

template<typename VT>
void gather_values(const double* vals, const int64_t* slot, int64_t nnz, VT* val_in,
    int64_t val_space, void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    if(val_space > 0) k_zero<VT><<<(unsigned) ((val_space + 255) / 256), 256, 0, st>>>(val_in, val_space);
    if(nnz > 0) k_gather<VT><<<(unsigned) ((nnz + 255) / 256), 256, 0, st>>>(vals, slot, nnz, val_in);
}

template<typename VT>
void factor_begin(const FactorArgs<VT>& a, void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    k_zero_carry<VT><<<1, 64, 0, st>>>(a.carry);
    if(a.num_levels > 0 && a.scan_maxch > 0) {
        const int64_t nf = (int64_t) (2 * a.scan_maxch + 1) * a.num_levels;
        k_zero<int32_t><<<(unsigned) ((nf + 255) / 256), 256, 0, st>>>(a.scan_flags, nf);
    }
    if(a.num_levels > 0) {
        const unsigned g = (unsigned) ((a.num_levels + 255) / 256);
        k_zero<int32_t><<<g, 256, 0, st>>>(a.defer_cnt, a.num_levels);
        k_zero<int32_t><<<g, 256, 0, st>>>(a.wave_cnt, a.num_levels);
        k_zero<int32_t><<<g, 256, 0, st>>>(a.store_cnt, a.num_levels);
    }
}

template<typename VT>
void factor_level(const FactorArgs<VT>& a, const LevelRange& r, void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    const int noff = r.size - r.npiv;
    const int tiny_blocks = fl_tiny_blocks(r);
    if(fl_grid_pivots_pairs(r) > 0)
        k_pivots_pairs<VT><<<fl_grid_pivots_pairs(r), 64, 0, st>>>(a, r);
    if(noff > 0) {
        const int grid = fl_grid_tasks(noff, 2048);
        k_offdiag_tiny_scale<VT><<<tiny_blocks, TINY_TPB, 0, st>>>(a, r);
        k_offdiag<VT><<<grid, 64, 0, st>>>(a, r);
        k_offdiag_general<VT><<<grid, 64, 0, st>>>(a, r);
    }
    if(a.pivoting && r.npiv > 0) k_pivot_cols<VT><<<r.npiv, 64, 0, st>>>(a, r);
    if(r.npiv > 0) k_rownorm<VT><<<r.npiv, 64, 0, st>>>(a, r);
    k_level_scan<VT><<<fl_grid_level_scan(r), SC_TPB, 0, st>>>(a, r);
    k_store_tiny<VT><<<fl_grid_store_tiny(r), 256, 0, st>>>(a, r);
    k_store<VT><<<fl_grid_tasks(r.size, 2048), 64, 0, st>>>(a, r);
}

template void gather_values<double>(const double*, const int64_t*, int64_t, double*, int64_t, void*);
template void gather_values<float>(const double*, const int64_t*, int64_t, float*, int64_t, void*);
template void factor_begin<double>(const FactorArgs<double>&, void*);
template void factor_begin<float>(const FactorArgs<float>&, void*);
template void factor_level<double>(const FactorArgs<double>&, const LevelRange&, void*);
template void factor_level<float>(const FactorArgs<float>&, const LevelRange&, void*);


int64_t factor_batch_chunks(int32_t n, int64_t nnz, int64_t val_space, int32_t num_levels,
    int32_t scan_maxch)
{
    int64_t m = std::max<int64_t>(std::max<int64_t>(n, nnz), std::max<int64_t>(val_space, 1));
    if(num_levels > 0 && scan_maxch > 0)
        m = std::max<int64_t>(m, (int64_t) (2 * scan_maxch + 1) * num_levels);
    m = std::max<int64_t>(m, num_levels);
    return (m + 255) / 256;
}

template<typename VT>
void factor_batch_begin(const FactorSys<VT>* fs, FactorItems chunks, hipStream_t st)
{
    if(chunks.count <= 0) return;
    bk_factor_begin<VT><<<(unsigned) chunks.count, 256, 0, st>>>(fs, chunks.items);
    bk_gather<VT><<<(unsigned) chunks.count, 256, 0, st>>>(fs, chunks.items);
}

// the launches of factor_level, in its order, each over the level's items
// of every system that has the level
template<typename VT>
void factor_batch_level(const FactorSys<VT>* fs, const FactorBatchLevel& L, int32_t level,
    hipStream_t st)
{
    auto grid = [](const FactorItems& f) { return (unsigned) f.count; };
    if(L.pp.count) bk_pivots_pairs<VT><<<grid(L.pp), 64, 0, st>>>(fs, L.pp.items, level);
    if(L.ts.count) bk_offdiag_tiny_scale<VT><<<grid(L.ts), TINY_TPB, 0, st>>>(fs, L.ts.items, level);
    if(L.off.count) {
        bk_offdiag<VT><<<grid(L.off), 64, 0, st>>>(fs, L.off.items, level);
        bk_offdiag_general<VT><<<grid(L.off), 64, 0, st>>>(fs, L.off.items, level);
    }
    if(L.rn.count) bk_rownorm<VT><<<grid(L.rn), 64, 0, st>>>(fs, L.rn.items, level);
    if(L.sc.count) bk_level_scan<VT><<<grid(L.sc), SC_TPB, 0, st>>>(fs, L.sc.items, level);
    if(L.stt.count) bk_store_tiny<VT><<<grid(L.stt), 256, 0, st>>>(fs, L.stt.items, level);
    if(L.st.count) bk_store<VT><<<grid(L.st), 64, 0, st>>>(fs, L.st.items, level);
}

template void factor_batch_begin<double>(const FactorSys<double>*, FactorItems, hipStream_t);
template void factor_batch_begin<float>(const FactorSys<float>*, FactorItems, hipStream_t);
template void factor_batch_level<double>(const FactorSys<double>*, const FactorBatchLevel&, int32_t, hipStream_t);
template void factor_batch_level<float>(const FactorSys<float>*, const FactorBatchLevel&, int32_t, hipStream_t);

}  // namespace culip
