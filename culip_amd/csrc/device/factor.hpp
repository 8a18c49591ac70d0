// Numeric BILDLt factor on the device: shared declarations.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

#include "../host/symbolic.hpp"

namespace culip {

// Device pointers used by the factor kernels of one system.
template<typename VT>
struct FactorArgs {
    int32_t n = 0, nb = 0;
    const int32_t* bs = nullptr;          // block starts (nb + 1)
    // n: pivot starts of this factor (a copy of the static ones; numeric
    // pivoting rewrites it as it picks 1x1 and 2x2 pivots)
    int32_t* is_piv = nullptr;
    // n: new -> old rows (identity, interchanged inside pivot blocks by
    // numeric pivoting); pivoting: 0 static, 1 Bunch-Kaufman, 2 rook
    int32_t* perm = nullptr;
    int32_t pivoting = 0;
    // 1: 2x2 pivot steps and the Givens-form D solves in FP64 (an FP32
    // factor's default); 0: in the factor's type, as BlockiLDLt<T>
    int32_t piv64 = 0;
    const BlockMeta* bin = nullptr;       // symbolic metadata (input stores)
    BlockMeta* bout = nullptr;            // factored metadata (output stores)
    const int32_t* csr_off = nullptr;     // coarse CSR offsets (block ids)
    const int32_t* csc_off = nullptr;
    const int32_t* csc_idx = nullptr;
    const int64_t* upd_off = nullptr;     // Schur pairs per block
    const int32_t* upd_ki = nullptr;
    const int32_t* upd_kj = nullptr;
    const int32_t* lvlb_idx = nullptr;    // level block list
    const int32_t* block_in_lvl = nullptr;
    const int32_t* lvl_rows = nullptr;
    const int32_t* rfo = nullptr;         // row_first_block_offset
    const int32_t* ix_in = nullptr;       // input stores
    const VT* val_in = nullptr;
    int32_t* ix_out = nullptr;            // output stores
    VT* val_out = nullptr;
    double* tri = nullptr;                // 3n, tridiagonal D
    VT* rn = nullptr;                     // n, row sums of squares
    // per level position scratch
    VT* if_val = nullptr;                 // 1024 per position
    uint16_t* if_idx = nullptr;           // 1024 per position
    int32_t* if_cnt = nullptr;
    VT* rn_part = nullptr;                // 32 per position
    int32_t* drop_cnt = nullptr;
    int32_t* keep = nullptr;              // entries kept per position
    int32_t* granted = nullptr;           // leftover space granted per position
    uint8_t* overflow = nullptr;
    int32_t* piv_loc = nullptr;           // nb
    // static, in level-list order: capacity of a sparse block (max_nnz >= 0)
    // or -1 - rows * cols for a dense block; first fine row of its block row
    const int32_t* lq_cap = nullptr;
    const int32_t* lq_fr = nullptr;
    int64_t* chunk = nullptr;             // per scan chunk: pool map (2), sizes (2)
    // per level: two flags per scan chunk (pool map, sizes published) and a
    // chunk ticket counter (2 scan_maxch + 1 words); zeroed at the start of
    // every factor
    int32_t* scan_flags = nullptr;
    int32_t scan_maxch = 0;
    // pinned host flag: 2 when a scan chunk's bounded wait gave up
    int32_t* err = nullptr;
    // static, per level-list position of an off-diagonal block, 12 int32:
    // {block, first row, rows | cols << 8 | all-1x1 << 16, pivot position},
    // {input format, nnz, col_ptr, val_ptr}, {first pair (low 32 bits),
    // pairs, flags, first pair (high 32 bits)} with flags bit 0: small
    // enough for the one-lane kernel
    const int32_t* off_task = nullptr;
    // static, per Schur pair: {block (k, i), block (k, j), first row of k, rows of k}
    const int32_t* upd4 = nullptr;
    uint32_t* piv_umask = nullptr;        // 32 per pivot position: nonzero columns > row
    uint32_t* piv_any = nullptr;          // per pivot position: any such column
    int32_t* defer_list = nullptr;        // positions left to the general kernel
    int32_t* defer_cnt = nullptr;         // one counter per level
    int32_t* wave_list = nullptr;         // positions left to the one-wave kernel
    int32_t* wave_cnt = nullptr;          // one counter per level
    int32_t* store_list = nullptr;        // positions stored by one wave each
    int32_t* store_cnt = nullptr;         // one counter per level
    int32_t num_levels = 0;
    // two sets of {pool, out_ix, out_val, 0}; level l reads set l & 1 and
    // writes set (l + 1) & 1
    int32_t* carry = nullptr;
    VT threshold = 0;
};

// positions per chunk of the level scans (256 threads x 8)
constexpr int SCAN_CHUNK = 2048;

struct LevelRange {
    int32_t lo;        // offset of the level in the level block list
    int32_t npiv;      // pivot blocks (first in the level)
    int32_t size;      // blocks in the level
    int32_t row_lo;    // offset of the level in lvl_rows
    int32_t level;     // level index
};

// threads of a k_offdiag_tiny_scale workgroup (one per one-lane block)
constexpr int TINY_TPB = 64;
// per-system grid of the loops over device-counted tasks (k_offdiag,
// k_offdiag_general, k_store) in a batched refactor: each task is one
// workgroup's work, so the grid changes no result
constexpr int BATCH_TASK_GRID = 64;

// Workgroups of the launches of one level (factor_level); the batched
// refactor launches the same workgroups for every system of a batch.
__host__ __device__ inline int32_t fl_tiny_blocks(const LevelRange& r)
{
    const int32_t noff = r.size - r.npiv;
    return noff > 0 ? (noff + TINY_TPB - 1) / TINY_TPB : 0;
}
__host__ __device__ inline int32_t fl_grid_pivots_pairs(const LevelRange& r)
{
    return r.npiv + fl_tiny_blocks(r);
}
__host__ __device__ inline int32_t fl_grid_tasks(int32_t count, int32_t cap)
{
    return count > 0 ? (count < cap ? count : cap) : 0;
}
__host__ __device__ inline int32_t fl_grid_level_scan(const LevelRange& r)
{
    return (r.size + SCAN_CHUNK - 1) / SCAN_CHUNK;
}
__host__ __device__ inline int32_t fl_grid_store_tiny(const LevelRange& r)
{
    return (r.size + 255) / 256;
}

// One system of a batched refactor (device table entry).
template<typename VT>
struct FactorSys {
    FactorArgs<VT> a;
    const LevelRange* lr = nullptr;       // per level (device)
    const double* vals = nullptr;         // K's values
    const int64_t* value_slot = nullptr;  // their slots in val_in (or -1)
    int64_t nnz = 0, val_space = 0;
    VT* val_in = nullptr;
    const int32_t* static_is_piv = nullptr;
};

// work items {system, workgroup of that system} of one batched launch
struct FactorItems {
    const int2* items = nullptr;
    int32_t count = 0;
};
// the launches of one level of a batched refactor
struct FactorBatchLevel {
    FactorItems pp;    // k_pivots_pairs
    FactorItems ts;    // k_offdiag_tiny_scale
    FactorItems off;   // k_offdiag and k_offdiag_general (BATCH_TASK_GRID per system)
    FactorItems rn;    // k_rownorm
    FactorItems sc;    // k_level_scan
    FactorItems stt;   // k_store_tiny
    FactorItems st;    // k_store (BATCH_TASK_GRID per system)
};

// Host entry points (factor.hip). All launches go to `stream`.
// Batched refactor: the start of every listed system's factor (pivot
// starts, zeroed and gathered input values, zeroed level counters), over
// chunk items {system, 256-entry chunk}; then one level of all systems.
template<typename VT>
void factor_batch_begin(const FactorSys<VT>* fs, FactorItems chunks, hipStream_t st);
template<typename VT>
void factor_batch_level(const FactorSys<VT>* fs, const FactorBatchLevel& L, int32_t level,
    hipStream_t st);
// chunks of 256 entries factor_batch_begin needs for a system
int64_t factor_batch_chunks(int32_t n, int64_t nnz, int64_t val_space, int32_t num_levels,
    int32_t scan_maxch);
template<typename VT>
void factor_begin(const FactorArgs<VT>& a, void* stream);
template<typename VT>
void factor_level(const FactorArgs<VT>& a, const LevelRange& r, void* stream);
// start of a factor: is_piv <- the static pivot starts, perm <- identity
void factor_reset_pivots(int32_t n, const int32_t* static_is_piv, int32_t* is_piv, int32_t* perm,
    void* stream);
// val_in[slot[q]] = vals[q] for slot[q] >= 0 (val_in zeroed first)
template<typename VT>
void gather_values(const double* vals, const int64_t* slot, int64_t nnz, VT* val_in,
    int64_t val_space, void* stream);

}  // namespace culip
