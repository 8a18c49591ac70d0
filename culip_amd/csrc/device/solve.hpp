// Block triangular solves with the BILDLt factor on the device.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "factor.hpp"
#include "spmv.hpp"

namespace culip {

// Solve layout built once per factorization from the block storage.
// For the U solve every fine row r holds the entries of the off-diagonal
// blocks of its block row (uo) and row r of U_ii^-1 (ui); for the U' solve
// every fine row c holds column c of the off-diagonal blocks above its
// block (to) and column c of U_jj^-1 (ti). Each row's entries come in block
// order, so all sums run in a fixed order. lr* = local row of each entry.
template<typename VT>
struct SolvePlan {
    int32_t* uo_ptr; int32_t* uo_col; VT* uo_val; uint8_t* uo_lr;
    int32_t* ui_ptr; int32_t* ui_col; VT* ui_val; uint8_t* ui_lr;
    int32_t* to_ptr; int32_t* to_col; VT* to_val; uint8_t* to_lr;
    int32_t* ti_ptr; int32_t* ti_col; VT* ti_val; uint8_t* ti_lr;
    int4* task_u;   // per block row in level order: {fr, h, uo range}, {ui range}
    int4* task_t;   // the same for the U' solve (to, ti)
};

struct SolveTopo {
    int32_t n, nb, num_levels;
    const int32_t* bs;          // block starts
    const int32_t* lvl_rows;    // block rows by level (device)
    const int32_t* lvl_off_h;   // level offsets (host)
    const int32_t* csr_off;     // coarse CSR offsets (block ids of a row)
    const int32_t* csc_off;     // coarse CSC offsets
    const int32_t* csc_idx;     // block ids of a column by row
    const BlockMeta* bout;      // factored metadata
    const int32_t* ix_out;
    const double* tri;
    const int32_t* is_piv;
    // single-launch solves
    double* chain_pub;          // 2n published values, two halves used in turn
    int32_t* err;               // set when a dependency wait gave up
};

// Counting pass of the layout: fills the four count arrays (n entries each,
// stored in the *_ptr arrays shifted by one) and the pivot-block inverses.
template<typename VT>
void solve_plan_count(const SolveTopo& t, const VT* val_out, VT* dinv, SolvePlan<VT>& p,
    void* stream);
// Fill pass after the *_ptr arrays hold exclusive prefix sums.
template<typename VT>
void solve_plan_fill(const SolveTopo& t, const VT* val_out, const VT* dinv, SolvePlan<VT>& p,
    void* stream);

// exclusive prefix sum of cnt[1..n] into ptr[0..n] (ptr[0] = 0), in place;
// tmp holds scan_tmp_ints(n) ints of scratch
size_t scan_tmp_ints(int32_t n);
void scan_counts(int32_t* ptr, int32_t n, int32_t* total_dev, int32_t* tmp, void* stream);

// chunks of the prefix sums of scan_counts over n entries
int32_t scan_chunks(int32_t n);

// One system of a batched layout build (device table entry).
template<typename VT>
struct PlanSys {
    SolveTopo t;
    SolvePlan<VT> p;
    const VT* val_out = nullptr;
    int32_t* totals = nullptr;    // 4 entry totals (uo, ui, to, ti)
    int32_t* scan_tmp = nullptr;  // scan_tmp_ints(n) ints
};
// solve_plan_count, the four scan_counts and solve_plan_fill of every
// listed system: blk = {system, block row} items, scan = {system, chunk <
// scan_chunks(n)}, task = {system, chunk of 256 block rows}; hm: the largest
// block height of the listed systems (1..32; the LDS of the pivot-block
// inverses is hm x hm per block row)
template<typename VT>
void plan_batch(const PlanSys<VT>* ps, FactorItems blk, FactorItems scan, FactorItems task,
    int32_t hm, hipStream_t st);

// x <- U \ x (in place); x <- U' \ x; x <- D \ x.
template<typename VT>
void trsv_upper(const SolveTopo& t, const SolvePlan<VT>& p, double* x, void* stream);
template<typename VT>
void trsv_upper_t(const SolveTopo& t, const SolvePlan<VT>& p, double* x, void* stream);
// single-launch solves that may run on one GPU at the same time
constexpr int TRSV_SINGLE_MAX_CONCURRENT = 4;

// The same solves as one launch each (at most TRSV_SINGLE_MAX_CONCURRENT at
// a time on one GPU),
// block rows ordered by level; readers wait for each value they need in
// chain_pub. A call
// publishes into half `parity` of chain_pub and presets the other half for
// the next call, which must pass the other parity; chain_reset presets both
// halves once. Returns the launch's hipError_t.
template<typename VT>
int trsv_upper_single(const SolveTopo& t, const SolvePlan<VT>& p, double* x, int32_t parity,
    void* stream);
template<typename VT>
int trsv_upper_t_single(const SolveTopo& t, const SolvePlan<VT>& p, double* x, int32_t parity,
    void* stream);
void chain_reset(double* pub, int64_t count, void* stream);
// x = U' \ (K src) as one single-launch solve: every block row forms its
// rows of K src itself (in k_csr_spmv's order) before it waits for its
// inputs; K: the CSR matrix on the device, n x n
template<typename VT>
int trsv_upper_t_single_product(const SolveTopo& t, const SolvePlan<VT>& p, const DevCsr* K,
    const double* src, double* x, int32_t parity, void* stream);
// x <- approximately U \ b (U' \ b when transpose) by `sweeps` block-Jacobi
// sweeps from 0; b_tmp holds b (n doubles), x_a / x_b are scratch vectors
template<typename VT>
void jacobi_sweeps(const SolveTopo& t, const SolvePlan<VT>& p, bool transpose, double* x,
    int32_t sweeps, double* b_tmp, double* x_a, double* x_b, void* stream);
// the same sweeps on b = K src, formed by the first sweep (into b_tmp)
template<typename VT>
void jacobi_sweeps_product(const SolveTopo& t, const SolvePlan<VT>& p, bool transpose,
    const DevCsr* K, const double* src, double* x, int32_t sweeps, double* b_tmp, double* x_a,
    double* x_b, void* stream);
void solve_D(int32_t n, const double* tri, const int32_t* is_piv, const double* b, double* x,
    void* stream);
// k_permute_Pt: out[i] = in[perm[i]]; k_permute_P: out[perm[i]] = in[i]
// (helper_kernels.impl.cuh:776-814); out must not be in
void permute_pull(int32_t n, const int32_t* perm, const double* in, double* out, void* stream);
void permute_push(int32_t n, const int32_t* perm, const double* in, double* out, void* stream);

}  // namespace culip
