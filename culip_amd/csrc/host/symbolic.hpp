// Symbolic phase of BILDLt on the host: everything derived from the fine
// pattern, block starts, pivot starts, fill level, fill factor and the value
// type T (block_ildlt.impl.cuh:44-1369, 2050-2150). Deterministic: where the
// reference's OpenMP frontier order varies, the single-thread order is used.
#pragma once

#include "common.hpp"

namespace culip {

enum : int32_t { FMT_SPARSE = 0, FMT_DENSE = 1 };

// matrix_block (block_ildlt.cuh:32-86), one row of this table per block id.
struct BlockMeta {
    int32_t id, format, nnz, max_nnz, got_extra, block_row, block_col, col_ptr, val_ptr;
};
static_assert(sizeof(BlockMeta) == 36, "matrix_block is nine int32");

struct Symbolic {
    bool f32 = false;            // BlockiLDLt<float> semantics
    int32_t n = 0, nb = 0;
    int32_t fill_level = 0;
    double fill_factor = 1.0;

    std::vector<int32_t> block_starts;   // nb + 1
    std::vector<int32_t> row_block;      // n: fine row -> block
    std::vector<int32_t> is_piv;         // n

    // statistics printed by the reference
    int32_t fine_u_nnz = 0, init_coarse_nnz = 0, fill_in_nnz = 0;
    int32_t blks_sparse = 0, blks_dense = 0;

    int32_t coarse_nnz = 0;
    std::vector<int32_t> csr_off, csr_col;   // coarse CSR (block id == position)
    std::vector<int32_t> csc_off, csc_idx;   // block ids sorted by (col, row)
    std::vector<int32_t> a_block_nnz;        // input entries per block

    int32_t num_levels = 0, max_level_size = 0, max_level_blocks = 0;
    std::vector<int32_t> lvl_off, lvl_rows, row_level;
    std::vector<int32_t> lvlb_off, lvlb_idx, block_in_lvl;
    std::vector<int32_t> row_first_block_offset;
    std::vector<int32_t> solve_off, solve_idx;

    std::vector<BlockMeta> blocks;           // pre-factor metadata
    int64_t ix_space = 0, val_space = 0;
    std::vector<int32_t> ix_store;           // sparse local indices (input)
    std::vector<int64_t> value_slot;         // fine entry -> val_store slot, -1 if row > col

    // Schur-update pairs of each off-diagonal block (i, j): blocks (k, i) and
    // (k, j) for every common k < i, in ascending k (the merge walk of
    // k_update_scale_offdiagonal_blocks, kernels :1692-1727).
    std::vector<int64_t> upd_off;            // coarse_nnz + 1
    std::vector<int32_t> upd_ki, upd_kj;
};

// Runs the full symbolic phase. Throws Error on invalid input.
void symbolic_analyze(Symbolic& s, const HostCsr& A, const std::vector<int32_t>& piv_starts,
    const std::vector<int32_t>& block_starts_in, int32_t fill_level, double fill_factor, bool f32);

// Fills the input value store (val_space entries) from fine values.
void symbolic_gather_values(const Symbolic& s, const double* fine_vals, double* val_store);

}  // namespace culip
