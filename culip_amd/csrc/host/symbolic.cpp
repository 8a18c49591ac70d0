// BILDLt symbolic phase (host). Follows libs/la/preconditioner/
// block_ildlt.impl.cuh; each step cites the lines it reproduces. The output
// is bit-identical to the reference's structures for the same inputs and T,
// with the single-thread order where the reference's OpenMP order varies.
#include "symbolic.hpp"
#include "parallel.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace culip {

namespace {

// counting sort of items by key in [0, nkeys): stable
template<typename KeyFn>
std::vector<int32_t> stable_bucket(int32_t nitems, int32_t nkeys, KeyFn key)
{
    std::vector<int32_t> start(nkeys + 1, 0), out(nitems);
    for(int32_t i = 0; i < nitems; ++i) start[key(i) + 1]++;
    for(int32_t k = 0; k < nkeys; ++k) start[k + 1] += start[k];
    for(int32_t i = 0; i < nitems; ++i) out[start[key(i)]++] = i;
    return out;
}

// should_be_sparse (:1027-1036): (sizeof(int) + sizeof(T)) * nnz as size_t
// against 0.5 * sizeof(T) * rows * cols evaluated in T.
template<typename T>
bool should_be_sparse(int32_t nnz, int32_t rows, int32_t cols)
{
    const T slf = (T) 0.5;
    const unsigned long lhs = (unsigned long) (sizeof(int32_t) + sizeof(T)) * (unsigned long) nnz;
    const T rhs = ((slf * (T) sizeof(T)) * (T) rows) * (T) cols;
    return (T) lhs < rhs;
}

template<typename T>
int32_t ceil_in_t(T v) { return (int32_t) std::ceil(v); }

}  // namespace

// build_initial_coarse (:352-484): coarse CSR of triu(A)'s block pattern and
// the number of fine entries per block. Block rows are independent (their
// fine rows are contiguous): counted, sorted and deduplicated in parallel,
// placed by prefix sums.
static void build_initial_coarse(Symbolic& s, const HostCsr& A)
{
    const int32_t nb = s.nb;
    const int64_t chunk = 256;
    // entries with row <= col per block row
    std::vector<int32_t> cnt(nb + 1, 0);
    parallel_chunks(nb, chunk, [&](int, int64_t lo, int64_t hi) {
        for(int64_t i = lo; i < hi; ++i) {
            int32_t c = 0;
            for(int32_t r = s.block_starts[i]; r < s.block_starts[i + 1]; ++r)
                for(int32_t q = A.row_ptr[r]; q < A.row_ptr[r + 1]; ++q) c += (r <= A.col[q]);
            cnt[i + 1] = c;
        }
    });
    for(int32_t i = 0; i < nb; ++i) cnt[i + 1] += cnt[i];
    s.fine_u_nnz = cnt[nb];
    // their block columns, sorted per block row; distinct columns counted
    std::vector<int32_t> bcol(cnt[nb]), ucnt(nb + 1, 0);
    parallel_chunks(nb, chunk, [&](int, int64_t lo, int64_t hi) {
        for(int64_t i = lo; i < hi; ++i) {
            int32_t at = cnt[i];
            for(int32_t r = s.block_starts[i]; r < s.block_starts[i + 1]; ++r)
                for(int32_t q = A.row_ptr[r]; q < A.row_ptr[r + 1]; ++q)
                    if(r <= A.col[q]) bcol[at++] = s.row_block[A.col[q]];
            std::sort(bcol.begin() + cnt[i], bcol.begin() + cnt[i + 1]);
            int32_t u = 0;
            for(int32_t q = cnt[i]; q < cnt[i + 1]; ++q) u += (q == cnt[i] || bcol[q] != bcol[q - 1]);
            ucnt[i + 1] = u;
        }
    });
    for(int32_t i = 0; i < nb; ++i) ucnt[i + 1] += ucnt[i];
    s.csr_off = ucnt;
    s.csr_col.assign(ucnt[nb], 0);
    s.a_block_nnz.assign(ucnt[nb], 0);
    parallel_chunks(nb, chunk, [&](int, int64_t lo, int64_t hi) {
        for(int64_t i = lo; i < hi; ++i) {
            int32_t at = ucnt[i] - 1;
            for(int32_t q = cnt[i]; q < cnt[i + 1]; ++q) {
                if(q == cnt[i] || bcol[q] != bcol[q - 1]) s.csr_col[++at] = bcol[q];
                s.a_block_nnz[at]++;
            }
        }
    });
    s.coarse_nnz = s.init_coarse_nnz = (int32_t) s.csr_col.size();
}

// determine_coarse_level_fill_and_diag (:492-728). Level-of-fill blocks are
// found by the BFS over fine rows of Hysom-Pothen (:510-558); a missing
// diagonal block is added when the row's first coarse entry is not the
// diagonal (:561-567). Entries are merged and sorted by (row, col).
static void level_fill(Symbolic& s, const HostCsr& A, int32_t max_level)
{
    const int32_t nb = s.nb;
    std::vector<std::vector<int32_t>> add(nb);
    const int32_t old_nnz = s.coarse_nnz;
    const int64_t chunk = 64;
    const int workers = parallel_workers(nb, chunk);
    // per-thread BFS state (visited stamps are block row ids, so they never
    // need clearing)
    std::vector<std::vector<int32_t>> visited(workers), length(workers), queue(workers);
    if(max_level > 0)
        for(int w = 0; w < workers; ++w) {
            visited[w].assign(nb, -1);
            length[w].assign(nb, 0);
            queue[w].assign(nb, 0);
        }
    parallel_chunks(nb, chunk, [&](int w, int64_t lo, int64_t hi) {
        for(int32_t i = (int32_t) lo; i < (int32_t) hi; ++i) {
            if(max_level > 0) {
                std::vector<int32_t>& vis = visited[w];
                std::vector<int32_t>& len = length[w];
                std::vector<int32_t>& que = queue[w];
                int32_t head = 0, tail = 0;
                que[tail++] = i;
                vis[i] = i;
                len[i] = 0;
                while(head < tail) {
                    const int32_t cur = que[head++];
                    for(int32_t r = s.block_starts[cur]; r < s.block_starts[cur + 1]; ++r)
                        for(int32_t q = A.row_ptr[r]; q < A.row_ptr[r + 1]; ++q) {
                            const int32_t kc = s.row_block[A.col[q]];
                            if(vis[kc] == i) continue;
                            vis[kc] = i;
                            len[kc] = 0;
                            if(kc < i && len[cur] < max_level) {
                                que[tail++] = kc;
                                len[kc] = len[cur] + 1;
                            }
                            if(kc > i && len[cur] > 0) add[i].push_back(kc);
                        }
                }
            }
            // diagonal check on the row's first entry only; an empty
            // trailing row (first entry index == nnz) counts as missing
            const int32_t f = s.csr_off[i];
            const bool has_diag = (f < old_nnz) && s.csr_col[f] == i;
            if(!has_diag) add[i].push_back(i);
        }
    });
    int32_t new_nnz = 0;
    for(auto& a : add) new_nnz += (int32_t) a.size();
    s.fill_in_nnz = new_nnz;
    if(new_nnz == 0) return;

    // merged rows, sorted by column, placed by a prefix sum
    std::vector<int32_t> off(nb + 1, 0);
    for(int32_t i = 0; i < nb; ++i)
        off[i + 1] = off[i] + (s.csr_off[i + 1] - s.csr_off[i]) + (int32_t) add[i].size();
    std::vector<int32_t> col(off[nb]), bn(off[nb]);
    parallel_chunks(nb, chunk, [&](int, int64_t lo, int64_t hi) {
        std::vector<std::pair<int32_t, int32_t>> row;
        for(int64_t i = lo; i < hi; ++i) {
            row.clear();
            for(int32_t q = s.csr_off[i]; q < s.csr_off[i + 1]; ++q)
                row.emplace_back(s.csr_col[q], s.a_block_nnz[q]);
            for(int32_t c : add[i]) row.emplace_back(c, 0);
            std::sort(row.begin(), row.end());
            int32_t at = off[i];
            for(auto& e : row) { col[at] = e.first; bn[at] = e.second; ++at; }
        }
    });
    s.csr_off.swap(off);
    s.csr_col.swap(col);
    s.a_block_nnz.swap(bn);
    s.coarse_nnz = (int32_t) s.csr_col.size();
}

// build_block_csc (:732-762) + the block-id rewrite of build_blocks
// (:1268-1279): block ids of each column in ascending row order.
static void build_csc(Symbolic& s)
{
    const int32_t nb = s.nb, K = s.coarse_nnz;
    s.csc_off.assign(nb + 1, 0);
    for(int32_t q = 0; q < K; ++q) s.csc_off[s.csr_col[q] + 1]++;
    for(int32_t j = 0; j < nb; ++j) s.csc_off[j + 1] += s.csc_off[j];
    s.csc_idx.assign(K, 0);
    std::vector<int32_t> fillp(s.csc_off.begin(), s.csc_off.end() - 1);
    for(int32_t i = 0; i < nb; ++i)
        for(int32_t q = s.csr_off[i]; q < s.csr_off[i + 1]; ++q)
            s.csc_idx[fillp[s.csr_col[q]]++] = q;
}

// find_level_sets (:766-1011) in the single-thread frontier order.
static void find_level_sets(Symbolic& s)
{
    const int32_t nb = s.nb, K = s.coarse_nnz;
    std::vector<int32_t> dep(nb);
    for(int32_t i = 0; i < nb; ++i) dep[i] = s.csc_off[i + 1] - s.csc_off[i];
    s.lvl_rows.assign(nb, 0);
    s.row_level.assign(nb, 0);
    s.lvl_off.assign(1, 0);
    std::vector<int32_t> lvl_blocks;
    auto rowlen = [&](int32_t i) { return s.csr_off[i + 1] - s.csr_off[i]; };

    int32_t offset = 0;
    for(int32_t i = 0; i < nb; ++i) if(dep[i] == 1) s.lvl_rows[offset++] = i;
    s.lvl_off.push_back(offset);
    int32_t blocks0 = 0;
    for(int32_t q = 0; q < offset; ++q) blocks0 += rowlen(s.lvl_rows[q]);
    lvl_blocks.push_back(blocks0);
    s.max_level_size = offset;
    s.max_level_blocks = blocks0;

    int32_t level = 0, prev = 0;
    while(prev < offset) {
        ++level;
        int32_t level_blocks = 0;
        const int32_t cur_off = offset;
        for(int32_t q = prev; q < cur_off; ++q) {
            const int32_t cur = s.lvl_rows[q];
            for(int32_t e = s.csr_off[cur]; e < s.csr_off[cur + 1]; ++e) {
                const int32_t jc = s.csr_col[e];
                if(--dep[jc] == 1) {
                    CULIP_REQUIRE(offset < nb, "level sets: inconsistent coarse pattern");
                    s.lvl_rows[offset++] = jc;
                    s.row_level[jc] = level;
                    level_blocks += rowlen(jc);
                }
            }
        }
        lvl_blocks.push_back(level_blocks);
        s.max_level_size = std::max(s.max_level_size, offset - cur_off);
        s.max_level_blocks = std::max(s.max_level_blocks, level_blocks);
        s.lvl_off.push_back(offset);
        prev = cur_off;
    }
    s.num_levels = level;
    s.lvl_off.resize(level + 1);
    CULIP_REQUIRE(s.lvl_off[level] == nb,
        "level sets: not every block row was scheduled (missing diagonal block?)");

    s.blocks.assign(K, BlockMeta{});
    parallel_chunks(nb, 256, [&](int, int64_t lo, int64_t hi) {
        for(int64_t i = lo; i < hi; ++i)
            for(int32_t q = s.csr_off[i]; q < s.csr_off[i + 1]; ++q) {
                s.blocks[q].id = q;
                s.blocks[q].block_row = (int32_t) i;
                s.blocks[q].block_col = s.csr_col[q];
            }
    });

    // level block list (:916-967): by (row level, pivot first), stable
    s.lvlb_idx = stable_bucket(K, 2 * level, [&](int32_t b) {
        const BlockMeta& m = s.blocks[b];
        return 2 * s.row_level[m.block_row] + (m.block_row == m.block_col ? 0 : 1);
    });
    s.lvlb_off.assign(level + 1, 0);
    for(int32_t l = 0; l < level; ++l) s.lvlb_off[l + 1] = s.lvlb_off[l] + lvl_blocks[l];
    s.block_in_lvl.assign(K, 0);
    for(int32_t p = 0; p < K; ++p) s.block_in_lvl[s.lvlb_idx[p]] = p;

    // row_first_block_offset (:984-1010); rows without off-diagonal blocks
    // keep -1 (the reference leaves them unwritten)
    s.row_first_block_offset.assign(nb, -1);
    for(int32_t l = 0; l < level; ++l) {
        const int32_t np = s.lvl_off[l + 1] - s.lvl_off[l];
        int32_t prev_row = -1;
        for(int32_t j = s.lvlb_off[l] + np; j < s.lvlb_off[l + 1]; ++j) {
            const int32_t r = s.blocks[s.lvlb_idx[j]].block_row;
            if(r != prev_row) {
                s.row_first_block_offset[r] = j - s.lvlb_off[l] - np;
                prev_row = r;
            }
        }
    }
}

// build_blocks (:1015-1280): sizes, formats, storage offsets in level-list
// order, gather of the input entries.
template<typename T>
static void build_blocks(Symbolic& s, const HostCsr& A, T fill_factor)
{
    const int32_t nb = s.nb, K = s.coarse_nnz;
    std::vector<int64_t> ixs(K + 1, 0), vls(K + 1, 0);
    const int64_t chunk = 256;
    std::vector<int32_t> nsparse(nb, 0);
    // block rows whose average block size is undefined (the lowest is
    // reported, whatever the schedule)
    std::vector<uint8_t> no_entries(nb, 0);
    parallel_chunks(nb, chunk, [&](int, int64_t lo, int64_t hi) {
    for(int32_t i = (int32_t) lo; i < (int32_t) hi; ++i) {
        int32_t sum_nnz = 0, sum_blocks = 0;
        bool need_avg = false;
        for(int32_t q = s.csr_off[i]; q < s.csr_off[i + 1]; ++q) {
            sum_nnz += s.a_block_nnz[q];
            sum_blocks += (s.a_block_nnz[q] != 0);
            need_avg |= (s.a_block_nnz[q] == 0);
        }
        int32_t avg_nnz = 0;
        if(need_avg) {
            if(sum_blocks == 0) {
                no_entries[i] = 1;
                continue;
            }
            avg_nnz = ceil_in_t<T>((T) sum_nnz / (T) sum_blocks);
        }
        const int32_t rows = s.block_starts[i + 1] - s.block_starts[i];
        for(int32_t q = s.csr_off[i]; q < s.csr_off[i + 1]; ++q) {
            BlockMeta& m = s.blocks[q];
            m.id = q;
            m.block_row = i;
            m.block_col = s.csr_col[q];
            m.nnz = s.a_block_nnz[q];
            m.got_extra = 0;
            const int32_t cols = s.block_starts[m.block_col + 1] - s.block_starts[m.block_col];
            const int32_t src = (m.nnz == 0) ? avg_nnz : m.nnz;
            m.max_nnz = std::min(rows * cols, ceil_in_t<T>(fill_factor * (T) src));
            const int32_t li = s.block_in_lvl[q];
            if(m.block_row != m.block_col && (double) m.block_row < 0.9 * (double) nb &&
                should_be_sparse<T>(m.max_nnz, rows, cols)) {
                m.format = FMT_SPARSE;
                ixs[li] = m.max_nnz;
                vls[li] = m.max_nnz;
                nsparse[i]++;
            } else {
                m.format = FMT_DENSE;
                ixs[li] = 0;
                vls[li] = (int64_t) rows * cols;
            }
        }
    }
    });
    for(int32_t i = 0; i < nb; ++i)
        if(no_entries[i])
            fail(CULIP_ERR_INVALID, "build_blocks: block row " + std::to_string(i) +
                " has no input entries (average block size undefined)");
    s.blks_sparse = 0;
    for(int32_t i = 0; i < nb; ++i) s.blks_sparse += nsparse[i];
    s.blks_dense = K - s.blks_sparse;
    int64_t a = 0, v = 0;
    for(int32_t p = 0; p <= K; ++p) {
        const int64_t ca = ixs[p], cv = vls[p];
        ixs[p] = a; vls[p] = v; a += ca; v += cv;
    }
    s.ix_space = ixs[K];
    s.val_space = vls[K];
    CULIP_REQUIRE(s.val_space < (int64_t) INT32_MAX,
        "build_blocks: value storage exceeds int32 offsets (mat_int_t)");
    parallel_chunks(K, 1 << 16, [&](int, int64_t lo, int64_t hi) {
        for(int64_t b = lo; b < hi; ++b) {
            s.blocks[b].col_ptr = (int32_t) ixs[s.block_in_lvl[b]];
            s.blocks[b].val_ptr = (int32_t) vls[s.block_in_lvl[b]];
        }
    });

    // gather map (:1214-1266): sparse entries in CSR scan order of the
    // block's fine rows, dense entries row-major; a block row's fine rows
    // fill only its own blocks, so block rows run in parallel
    s.ix_store.assign(s.ix_space, -1);
    s.value_slot.assign(A.nnz(), -1);
    std::vector<int32_t> rec(K, 0);
    parallel_chunks(nb, chunk, [&](int, int64_t lo, int64_t hi) {
    for(int32_t r = s.block_starts[lo]; r < s.block_starts[hi]; ++r) {
        const int32_t bi = s.row_block[r];
        const int32_t lr = r - s.block_starts[bi];
        int32_t e = s.csr_off[bi];
        for(int32_t q = A.row_ptr[r]; q < A.row_ptr[r + 1]; ++q) {
            const int32_t c = A.col[q];
            if(c < r) continue;
            const int32_t bj = s.row_block[c];
            while(s.csr_col[e] != bj) ++e;   // columns sorted: monotone walk
            const BlockMeta& m = s.blocks[e];
            const int32_t cs = s.block_starts[bj + 1] - s.block_starts[bj];
            const int32_t lc = c - s.block_starts[bj];
            if(m.format == FMT_SPARSE) {
                s.ix_store[(int64_t) m.col_ptr + rec[e]] = lr * cs + lc;
                s.value_slot[q] = (int64_t) m.val_ptr + rec[e];
            } else {
                s.value_slot[q] = (int64_t) m.val_ptr + (int64_t) lr * cs + lc;
            }
            rec[e]++;
        }
    }
    });
}

// solve_analysis_block (:2050-2150)
static void solve_analysis(Symbolic& s)
{
    const int32_t K = s.coarse_nnz, L = s.num_levels;
    s.solve_idx = stable_bucket(K, 2 * L, [&](int32_t b) {
        const BlockMeta& m = s.blocks[b];
        return 2 * s.row_level[m.block_col] + (m.block_row == m.block_col ? 0 : 1);
    });
    s.solve_off.assign(L + 1, 0);
    for(int32_t j = 0; j < s.nb; ++j)
        s.solve_off[s.row_level[j] + 1] += s.csc_off[j + 1] - s.csc_off[j];
    for(int32_t l = 0; l < L; ++l) s.solve_off[l + 1] += s.solve_off[l];
}

// Schur-update pairs for each off-diagonal block (merge walk, kernels
// :1692-1727), computed once per pattern: for block (i, j), j != i, the
// blocks (k, i) and (k, j) for every k < i present in both columns, in
// ascending k. Block rows are independent: a first parallel pass counts the
// pairs of every block, a prefix sum places them, a second pass writes them.
static void update_pairs(Symbolic& s)
{
    const int32_t K = s.coarse_nnz, nb = s.nb;
    // block row of every column entry, in column order
    std::vector<int32_t> crow(K);
    parallel_chunks(K, 1 << 16, [&](int, int64_t lo, int64_t hi) {
        for(int64_t p = lo; p < hi; ++p) crow[p] = s.blocks[s.csc_idx[p]].block_row;
    });
    const int64_t chunk = 64;
    const int workers = parallel_workers(nb, chunk);
    std::vector<std::vector<int32_t>> marks(workers, std::vector<int32_t>(nb, -1));
    // emit(q, block (k, i), block (k, j)) for the pairs of block row i
    auto row_pairs = [&](int32_t i, std::vector<int32_t>& mark, auto&& emit) {
        for(int32_t p = s.csc_off[i]; p < s.csc_off[i + 1] - 1; ++p) mark[crow[p]] = s.csc_idx[p];
        for(int32_t q = s.csr_off[i]; q < s.csr_off[i + 1]; ++q) {
            const int32_t j = s.csr_col[q];
            if(j == i) continue;
            for(int32_t p = s.csc_off[j]; p < s.csc_off[j + 1] - 1; ++p) {
                const int32_t k = crow[p];
                if(k >= i) break;
                if(mark[k] >= 0) emit(q, mark[k], s.csc_idx[p]);
            }
        }
        for(int32_t p = s.csc_off[i]; p < s.csc_off[i + 1] - 1; ++p) mark[crow[p]] = -1;
    };
    s.upd_off.assign(K + 1, 0);
    parallel_chunks(nb, chunk, [&](int w, int64_t lo, int64_t hi) {
        for(int64_t i = lo; i < hi; ++i)
            row_pairs((int32_t) i, marks[w], [&](int32_t q, int32_t, int32_t) { s.upd_off[q + 1]++; });
    });
    for(int32_t q = 0; q < K; ++q) s.upd_off[q + 1] += s.upd_off[q];
    // 24 bytes per pair on the device (pair lists and the task records):
    // 2^33 pairs is 206 GB, most of the GPU
    if(s.upd_off[K] > ((int64_t) 1 << 33))
        fail(CULIP_ERR_INVALID, "symbolic: " + std::to_string(s.upd_off[K]) +
            " Schur pairs exceed the 2^33 supported (fill level too high for this pattern)");
    s.upd_ki.assign(s.upd_off[K], 0);
    s.upd_kj.assign(s.upd_off[K], 0);
    parallel_chunks(nb, chunk, [&](int w, int64_t lo, int64_t hi) {
        for(int64_t i = lo; i < hi; ++i) {
            int64_t at = -1;
            int32_t cur = -1;
            row_pairs((int32_t) i, marks[w], [&](int32_t q, int32_t bki, int32_t bkj) {
                if(q != cur) { cur = q; at = s.upd_off[q]; }
                s.upd_ki[at] = bki;
                s.upd_kj[at] = bkj;
                ++at;
            });
        }
    });
}

void symbolic_analyze(Symbolic& s, const HostCsr& A, const std::vector<int32_t>& piv_starts,
    const std::vector<int32_t>& bs_in, int32_t fill_level, double fill_factor, bool f32)
{
    CULIP_REQUIRE(A.m == A.n && A.m > 0, "symbolic: matrix must be square and non-empty");
    CULIP_REQUIRE(fill_level >= 0, "symbolic: fill level must be >= 0");
    CULIP_REQUIRE(fill_factor >= 1.0, "symbolic: fill factor must be >= 1");
    const int32_t n = A.m;
    s.f32 = f32;
    s.n = n;
    s.fill_level = fill_level;
    s.fill_factor = fill_factor;
    s.nb = (int32_t) bs_in.size();
    CULIP_REQUIRE(s.nb > 0 && bs_in[0] == 0, "symbolic: block starts must begin at 0");
    s.block_starts = bs_in;
    s.block_starts.push_back(n);
    for(int32_t b = 0; b < s.nb; ++b) {
        const int32_t sz = s.block_starts[b + 1] - s.block_starts[b];
        CULIP_REQUIRE(sz >= 1 && sz <= 32, "symbolic: block sizes must be in [1, 32]");
    }
    s.row_block.assign(n, 0);
    for(int32_t b = 0; b < s.nb; ++b)
        for(int32_t r = s.block_starts[b]; r < s.block_starts[b + 1]; ++r) s.row_block[r] = b;
    // is_piv scatter (upload_data :1292-1302); every block must start a pivot
    s.is_piv.assign(n, 0);
    for(int32_t p : piv_starts) {
        CULIP_REQUIRE(p >= 0 && p < n, "symbolic: pivot start out of range");
        s.is_piv[p] = 1;
    }
    for(int32_t b = 0; b < s.nb; ++b)
        CULIP_REQUIRE(s.is_piv[s.block_starts[b]],
            "symbolic: every block start must be a pivot start (a 2x2 pivot may not straddle blocks)");
    for(int32_t r = 0; r + 1 < n; ++r)
        CULIP_REQUIRE(s.is_piv[r] || s.is_piv[r + 1] || false || (r > 0 && s.is_piv[r - 1]),
            "symbolic: pivots must be 1x1 or 2x2");
    for(int32_t r = 0; r < n; ++r) {
        for(int32_t q = A.row_ptr[r]; q < A.row_ptr[r + 1]; ++q) {
            CULIP_REQUIRE(A.col[q] >= 0 && A.col[q] < n, "symbolic: column index out of range");
            CULIP_REQUIRE(q == A.row_ptr[r] || A.col[q - 1] <= A.col[q],
                "symbolic: CSR columns must be sorted per row");
        }
    }

    build_initial_coarse(s, A);
    level_fill(s, A, fill_level);
    build_csc(s);
    find_level_sets(s);
    if(f32) build_blocks<float>(s, A, (float) fill_factor);
    else build_blocks<double>(s, A, fill_factor);
    solve_analysis(s);
    update_pairs(s);
}

void symbolic_gather_values(const Symbolic& s, const double* fine_vals, double* val_store)
{
    std::fill(val_store, val_store + s.val_space, 0.0);
    for(size_t q = 0; q < s.value_slot.size(); ++q)
        if(s.value_slot[q] >= 0) val_store[s.value_slot[q]] = fine_vals[q];
}

}  // namespace culip
