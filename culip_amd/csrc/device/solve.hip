// Solve layout construction and level-scheduled block triangular solves.
//
// Reference: solve_U_block / solve_Ut_block (block_ildlt.impl.cuh:2378-2426,
// 2557-2606) with k_solve_blocks_offdiagonal_atomic / _diagonal_atomic
// (block_ildlt_kernels.impl.cuh:2354-2606): per level, x[rows] -= U_blk x
// for the level's off-diagonal blocks, then a 32-step substitution with the
// pivot block. Here the factor is first re-laid out per fine row (entries of
// a block row contiguous), the pivot blocks are inverted once, and each level
// is one launch: one wave per block row gathers its rows with coalesced loads
// and a segmented wave reduction (fixed order, no atomics), then applies
// U_ii^-1. k_solve_D follows helper_kernels.impl.cuh:720-772.
// Factor entries are read in their storage precision (FP32 for an FP32
// factor) and every product and sum runs in FP64: the preconditioner stays
// one fixed linear operator to FP64 rounding, which SQMR's short
// recurrences need (FP32 sums make it vary by ~1e-7 from apply to apply and
// cost the recurrence most of its progress below that level).
#include "solve.hpp"
#include "batch.hpp"
#include "wave.hpp"

#include <hip/hip_runtime.h>

namespace culip {

namespace {

// columns of U_ii^-1 by back substitution, in double; X[r][c] at X[hm r + c]
// (hm x hm LDS, hm >= h: the largest block height of the launch)
template<typename VT>
__device__ void diag_inverse(const VT* U, int h, double* X, int hm)
{
    const int c = lane_id();
    // a tile with no entry right of its diagonal (most pivot blocks of a
    // sparse pattern) has the diagonal inverse: the back substitution below
    // would only add exact zeros to it
    bool offd = false;
    if(c < h)
        for(int k = c + 1; k < h; ++k) offd = offd || (U[c * h + k] != (VT) 0);
    if(wave_ballot(offd) == 0) {
        for(int e = c; e < hm * hm; e += 64) {
            const int r = e / hm, q = e - r * hm;
            X[e] = (r == q && r < h) ? 1.0 / (double) U[r * h + r] : 0.0;
        }
        wave_sync();
        return;
    }
    for(int r = hm - 1; r >= 0; --r) {
        if(c < hm) {
            double v = 0.0;
            if(c < h && r < h && r <= c) {
                if(r == c) {
                    v = 1.0 / (double) U[r * h + r];
                } else {
                    double s = 0.0;
                    for(int k = r + 1; k <= c; ++k) s = fma((double) U[r * h + k], X[hm * k + c], s);
                    v = -s / (double) U[r * h + r];
                }
            }
            X[hm * r + c] = v;
        }
    }
    wave_sync();
}

// up to 64 factored blocks staged one per lane: metadata, and the entries
// of sparse blocks with at most PLAN_E of them
constexpr int PLAN_E = 8;

template<typename VT>
struct PlanStage {
    int32_t fmt[64], nnz[64], cp[64], vp[64], off[64], ext[64];
    uint8_t small[64];
    uint8_t rr[64 * PLAN_E], cc[64 * PLAN_E];   // local row and column of each entry
    VT v[64 * PLAN_E];
};

// off / ext: first fine index and size of the block's other dimension; w:
// the block's width (the stride of its entry indices)
template<typename VT>
__device__ void plan_stage(PlanStage<VT>& sg, int k, const BlockMeta& m, int off, int ext, int w,
    const int32_t* ix_out, const VT* val_out)
{
    sg.fmt[k] = m.format;
    sg.nnz[k] = m.nnz;
    sg.cp[k] = m.col_ptr;
    sg.vp[k] = m.val_ptr;
    sg.off[k] = off;
    sg.ext[k] = ext;
    const bool sm = m.format == FMT_SPARSE && m.nnz <= PLAN_E;
    sg.small[k] = sm ? 1 : 0;
    if(sm)
        for(int e = 0; e < m.nnz; ++e) {
            const int q = ix_out[(int64_t) m.col_ptr + e];
            const int r = q / w;
            sg.rr[PLAN_E * k + e] = (uint8_t) r;
            sg.cc[PLAN_E * k + e] = (uint8_t) (q - r * w);
            sg.v[PLAN_E * k + e] = val_out[(int64_t) m.val_ptr + e];
        }
}

// true when every staged block of the round is small with at most one
// entry (nearly every off-diagonal block at level-0 fill): the round is then
// placed in parallel instead of walked block by block
template<typename VT>
__device__ __forceinline__ bool plan_one_entry_round(const PlanStage<VT>& sg, int nblk)
{
    const int lane = lane_id();
    const bool simple = lane >= nblk || (sg.small[lane] && sg.nnz[lane] <= 1);
    return wave_ballot(!simple) == 0;
}

// For a one-entry round: key = the local row (by_row) or column of lane k's
// entry (64: none), rank = entries of the same key on lower lanes (earlier
// blocks), and on lane r < ext the round's number of entries with key r.
template<typename VT>
__device__ __forceinline__ void plan_rank(const PlanStage<VT>& sg, int nblk, int ext, bool by_row,
    int& key, int& rank, int& add)
{
    const int lane = lane_id();
    const bool has = lane < nblk && sg.nnz[lane] == 1;
    key = has ? (by_row ? sg.rr[PLAN_E * lane] : sg.cc[PLAN_E * lane]) : 64;
    rank = 0;
    add = 0;
    for(int r = 0; r < ext; ++r) {
        const uint64_t m = wave_ballot(key == r);
        if(key == r) rank = lanes_below(m);
        if(lane == r) add = __popcll(m);
    }
}

// X: hm x hm doubles of LDS (the launch's dynamic LDS), hm >= every block
// height of the launch
template<typename VT, bool FILL>
__device__ __forceinline__ void plan_rows_body(const SolveTopo& t, const VT* val_out,
    const SolvePlan<VT>& p, int bid_, double* X, int hm)
{
    const int lane = lane_id();
    const int i = bid_;
    const int fr = t.bs[i];
    const int h = t.bs[i + 1] - fr;
    // pivot block inverse rows (ui) and columns (ti)
    const BlockMeta dm = t.bout[t.csr_off[i]];
    diag_inverse(val_out + dm.val_ptr, h, X, hm);
    if(lane < h) {
        int cr = 0, cc = 0;
        int32_t wr = 0, wc = 0;
        if(FILL) { wr = p.ui_ptr[fr + lane]; wc = p.ti_ptr[fr + lane]; }
        for(int c = 0; c < h; ++c) {
            const double v = X[hm * lane + c];  // row lane of U^-1
            if(v != 0.0) {
                if(FILL) { p.ui_col[wr + cr] = fr + c; p.ui_val[wr + cr] = (VT) v; p.ui_lr[wr + cr] = (uint8_t) lane; }
                ++cr;
            }
            const double u = X[hm * c + lane];  // column lane of U^-1
            if(u != 0.0) {
                if(FILL) { p.ti_col[wc + cc] = fr + c; p.ti_val[wc + cc] = (VT) u; p.ti_lr[wc + cc] = (uint8_t) lane; }
                ++cc;
            }
        }
        if(!FILL) { p.ui_ptr[fr + lane + 1] = cr; p.ti_ptr[fr + lane + 1] = cc; }
    }
    // off-diagonal blocks of block row i (uo), by block id; 64 blocks at a
    // time are staged (metadata and small entry lists) one per lane, then
    // walked in order by the lanes of the fine rows
    __shared__ PlanStage<VT> sg;
    int cnt = 0;
    const int32_t base = (FILL && lane < h) ? p.uo_ptr[fr + lane] : 0;
    const int b0 = t.csr_off[i] + 1, b1 = t.csr_off[i + 1];
    for(int c0 = b0; c0 < b1; c0 += 64) {
        const int nblk = min(64, b1 - c0);
        wave_sync();
        if(lane < nblk) {
            const BlockMeta m = t.bout[c0 + lane];
            const int wb = t.bs[m.block_col + 1] - t.bs[m.block_col];
            plan_stage(sg, lane, m, t.bs[m.block_col], wb, wb, t.ix_out, val_out);
        }
        wave_sync();
        if(plan_one_entry_round(sg, nblk)) {
            // every staged block holds at most one entry: lane k places block
            // k's entry after the entries of its row from earlier blocks
            int key, rank, add;
            plan_rank(sg, nblk, h, true, key, rank, add);
            const int prev = __shfl(cnt, key & 63, 64);
            const int32_t rbase = __shfl(base, key & 63, 64);
            if(FILL && key < 64) {
                const int32_t pos = rbase + prev + rank;
                p.uo_col[pos] = sg.off[lane] + sg.cc[PLAN_E * lane];
                p.uo_val[pos] = sg.v[PLAN_E * lane];
                p.uo_lr[pos] = (uint8_t) key;
            }
            cnt += add;
            continue;
        }
        for(int k = 0; k < nblk; ++k) {
            const int fc = sg.off[k], w = sg.ext[k];
            if(sg.small[k]) {
                for(int e = 0; e < sg.nnz[k]; ++e) {
                    if(sg.rr[PLAN_E * k + e] == lane) {
                        if(FILL) { p.uo_col[base + cnt] = fc + sg.cc[PLAN_E * k + e]; p.uo_val[base + cnt] = sg.v[PLAN_E * k + e]; p.uo_lr[base + cnt] = (uint8_t) lane; }
                        ++cnt;
                    }
                }
            } else if(lane < h) {
                const VT* v = val_out + sg.vp[k];
                if(sg.fmt[k] == FMT_DENSE) {
                    for(int c = 0; c < w; ++c) {
                        const VT x = v[lane * w + c];
                        if(x != (VT) 0) {
                            if(FILL) { p.uo_col[base + cnt] = fc + c; p.uo_val[base + cnt] = x; p.uo_lr[base + cnt] = (uint8_t) lane; }
                            ++cnt;
                        }
                    }
                } else {
                    const int32_t* ix = t.ix_out + sg.cp[k];
                    for(int e = 0; e < sg.nnz[k]; ++e) {
                        const int q = ix[e];
                        const int r = q / w;
                        if(r == lane) {
                            if(FILL) { p.uo_col[base + cnt] = fc + (q - r * w); p.uo_val[base + cnt] = v[e]; p.uo_lr[base + cnt] = (uint8_t) lane; }
                            ++cnt;
                        }
                    }
                }
            }
        }
    }
    if(!FILL && lane < h) p.uo_ptr[fr + lane + 1] = cnt;
}

template<typename VT, bool FILL>
__global__ __launch_bounds__(64) void k_plan_rows(SolveTopo t, const VT* val_out, SolvePlan<VT> p)
{
    extern __shared__ double plan_x[];
    plan_rows_body<VT, FILL>(t, val_out, p, blockIdx.x, plan_x, 32);
}

template<typename VT, bool FILL>
__device__ __forceinline__ void plan_cols_body(const SolveTopo& t, const VT* val_out,
    const SolvePlan<VT>& p, int bid_)
{
    __shared__ PlanStage<VT> sg;
    const int lane = lane_id();
    const int j = bid_;
    const int fc = t.bs[j];
    const int w = t.bs[j + 1] - fc;
    int cnt = 0;
    const int32_t base = (FILL && lane < w) ? p.to_ptr[fc + lane] : 0;
    // blocks of column j above the diagonal, ascending block row, staged 64
    // at a time as in k_plan_rows
    const int q0 = t.csc_off[j], q1 = t.csc_off[j + 1] - 1;
    for(int c0 = q0; c0 < q1; c0 += 64) {
        const int nblk = min(64, q1 - c0);
        wave_sync();
        if(lane < nblk) {
            const BlockMeta m = t.bout[t.csc_idx[c0 + lane]];
            plan_stage(sg, lane, m, t.bs[m.block_row], t.bs[m.block_row + 1] - t.bs[m.block_row], w,
                t.ix_out, val_out);
        }
        wave_sync();
        if(plan_one_entry_round(sg, nblk)) {
            // as in k_plan_rows, by column
            int key, rank, add;
            plan_rank(sg, nblk, w, false, key, rank, add);
            const int prev = __shfl(cnt, key & 63, 64);
            const int32_t cbase = __shfl(base, key & 63, 64);
            if(FILL && key < 64) {
                const int32_t pos = cbase + prev + rank;
                p.to_col[pos] = sg.off[lane] + sg.rr[PLAN_E * lane];
                p.to_val[pos] = sg.v[PLAN_E * lane];
                p.to_lr[pos] = (uint8_t) key;
            }
            cnt += add;
            continue;
        }
        for(int k = 0; k < nblk; ++k) {
            const int fr = sg.off[k], h = sg.ext[k];
            if(sg.small[k]) {
                for(int e = 0; e < sg.nnz[k]; ++e) {
                    if(sg.cc[PLAN_E * k + e] == lane) {
                        if(FILL) { p.to_col[base + cnt] = fr + sg.rr[PLAN_E * k + e]; p.to_val[base + cnt] = sg.v[PLAN_E * k + e]; p.to_lr[base + cnt] = (uint8_t) lane; }
                        ++cnt;
                    }
                }
            } else if(lane < w) {
                const VT* v = val_out + sg.vp[k];
                if(sg.fmt[k] == FMT_DENSE) {
                    for(int r = 0; r < h; ++r) {
                        const VT x = v[r * w + lane];
                        if(x != (VT) 0) {
                            if(FILL) { p.to_col[base + cnt] = fr + r; p.to_val[base + cnt] = x; p.to_lr[base + cnt] = (uint8_t) lane; }
                            ++cnt;
                        }
                    }
                } else {
                    const int32_t* ix = t.ix_out + sg.cp[k];
                    for(int e = 0; e < sg.nnz[k]; ++e) {
                        const int qq = ix[e];
                        const int r = qq / w;
                        if(qq - r * w == lane) {
                            if(FILL) { p.to_col[base + cnt] = fr + r; p.to_val[base + cnt] = v[e]; p.to_lr[base + cnt] = (uint8_t) lane; }
                            ++cnt;
                        }
                    }
                }
            }
        }
    }
    if(!FILL && lane < w) p.to_ptr[fc + lane + 1] = cnt;
}

template<typename VT, bool FILL>
__global__ __launch_bounds__(64) void k_plan_cols(SolveTopo t, const VT* val_out, SolvePlan<VT> p)
{
    plan_cols_body<VT, FILL>(t, val_out, p, blockIdx.x);
}

// inclusive scan of a[1..n] in place (a[0] = 0), one workgroup, in chunks
// prefix sums of a[1..n] in place, in chunks of PS_TPB * PS_IPT entries:
// chunk totals first, then each chunk adds the totals before it
constexpr int PS_TPB = 256;
constexpr int PS_IPT = 8;
constexpr int PS_CHUNK = PS_TPB * PS_IPT;

struct ISum {
    int32_t s;
    __device__ static ISum shfl_up(const ISum& v, int o) { return {__shfl_up(v.s, o, WAVE)}; }
};
struct ISumOp {
    __device__ ISum operator()(const ISum& p, const ISum& q) const { return {p.s + q.s}; }
};

__device__ __forceinline__ void chunk_sums_body(const int32_t* a, int32_t n, int32_t* tmp, int bid_)
{
    __shared__ ISum sh[PS_TPB / WAVE];
    int32_t v = 0;
#pragma unroll
    for(int k = 0; k < PS_IPT; ++k) {
        const int64_t i = 1 + (int64_t) bid_ * PS_CHUNK + (int64_t) threadIdx.x * PS_IPT + k;
        if(i <= n) v += a[i];
    }
    ISum tot;
    (void) block_exclusive_scan<ISum, ISumOp, PS_TPB / WAVE>(ISum{v}, ISumOp(), ISum{0}, sh, &tot);
    if(threadIdx.x == 0) tmp[bid_] = tot.s;
}

__global__ __launch_bounds__(PS_TPB) void k_chunk_sums(const int32_t* a, int32_t n, int32_t* tmp)
{
    chunk_sums_body(a, n, tmp, blockIdx.x);
}

__device__ __forceinline__ void chunk_scan_body(int32_t* a, int32_t n, const int32_t* tmp,
    int32_t* total, int bid_, int nblk_)
{
    __shared__ ISum sh[PS_TPB / WAVE];
    const int c = bid_;
    // totals of the chunks before this one
    int32_t pre = 0;
    for(int q = threadIdx.x; q < c; q += PS_TPB) pre += tmp[q];
    ISum ptot;
    (void) block_exclusive_scan<ISum, ISumOp, PS_TPB / WAVE>(ISum{pre}, ISumOp(), ISum{0}, sh, &ptot);
    int32_t v[PS_IPT];
    int32_t s = 0;
    const int64_t i0 = 1 + (int64_t) c * PS_CHUNK + (int64_t) threadIdx.x * PS_IPT;
#pragma unroll
    for(int k = 0; k < PS_IPT; ++k) {
        v[k] = (i0 + k <= n) ? a[i0 + k] : 0;
        s += v[k];
    }
    ISum tot;
    const ISum ex = block_exclusive_scan<ISum, ISumOp, PS_TPB / WAVE>(ISum{s}, ISumOp(), ISum{0}, sh, &tot);
    int32_t run = ptot.s + ex.s;
#pragma unroll
    for(int k = 0; k < PS_IPT; ++k) {
        run += v[k];
        if(i0 + k <= n) a[i0 + k] = run;
    }
    if(threadIdx.x == 0) {
        if(c == 0) a[0] = 0;
        if(c == nblk_ - 1 && total) *total = ptot.s + tot.s;
    }
}

__global__ __launch_bounds__(PS_TPB) void k_chunk_scan(int32_t* a, int32_t n, const int32_t* tmp,
    int32_t* total)
{
    chunk_scan_body(a, n, tmp, total, blockIdx.x, gridDim.x);
}

// task descriptors of the block rows in level order, two int4 each:
// {fr, h, off-diagonal begin, end}, {inverse begin, end, 0, 0}
__device__ __forceinline__ void plan_tasks_body(int32_t nb, const int32_t* lvl_rows,
    const int32_t* bs, const int32_t* o_ptr, const int32_t* i_ptr, int4* task, int bid_)
{
    const int t = bid_ * blockDim.x + threadIdx.x;
    if(t >= nb) return;
    const int i = lvl_rows[t];
    const int fr = bs[i], h = bs[i + 1] - fr;
    task[2 * t] = make_int4(fr, h, o_ptr[fr], o_ptr[fr + h]);
    task[2 * t + 1] = make_int4(i_ptr[fr], i_ptr[fr + h], 0, 0);
}

__global__ void k_plan_tasks(int32_t nb, const int32_t* lvl_rows, const int32_t* bs,
    const int32_t* o_ptr, const int32_t* i_ptr, int4* task)
{
    plan_tasks_body(nb, lvl_rows, bs, o_ptr, i_ptr, task, blockIdx.x);
}

// res = U_ii^-1 y for one block row (y, res: LDS, res zeroed; all lanes
// active), from the
// row-ordered entries ib..ie of U_ii^-1; ic0 / iv0 / il0: this lane's entry
// of the first 64. Returns res[lane] (lanes < h). A diagonal U_ii^-1 (one
// entry per row, on the diagonal) is applied lane by lane; the
// general form sums each row's products by segmented scans. Both give
// 0 + (products in row order), the same bits.
template<typename VT>
__device__ __forceinline__ double apply_inverse(int fr, int h, int ib, int ie, int ic0, VT iv0,
    int il0, const int32_t* i_col, const VT* i_val, const uint8_t* i_lr, const double* y,
    double* res)
{
    const int lane = lane_id();
    if(ie - ib == h && wave_ballot(lane < h && (il0 != lane || ic0 != fr + lane)) == 0)
        return (lane < h) ? fma((double) iv0, y[lane], 0.0) : 0.0;
    const bool in0 = ib + lane < ie;
    int icol = in0 ? ic0 : fr;
    VT ival = in0 ? iv0 : (VT) 0;
    int isg = in0 ? il0 : 64;
    for(int e0 = ib; e0 < ie; e0 += 64) {
        if(e0 != ib) {
            const int e = e0 + lane;
            isg = 64;
            if(e < ie) { icol = i_col[e]; ival = i_val[e]; isg = i_lr[e]; }
        }
        double p = seg_scan_runs_fma((isg < 64) ? (double) ival : 0.0, y[icol - fr], isg);
        const int nsg = next_lane(isg, -1);
        if(isg < 64 && (lane == 63 || nsg != isg)) res[isg] += p;
        wave_sync();
    }
    return (lane < h) ? res[lane] : 0.0;
}

constexpr int TRSV_PF = 4;   // 64-entry chunks whose loads are issued together

// one level: one wave per block row; y = x_i - sum(off-diag) ; x_i = Uinv y.
// All entry loads of up to TRSV_PF chunks are issued before the x gathers;
// the chunks' segmented sums are then applied to y in chunk order.
// One block row of U x = b: x_i = U_ii^-1 (b_i - sum_j U_ij x_j) with FP64
// sums, the off-diagonal products summed in entry order by segmented scans
// over 64-entry chunks, then the rows of U_ii^-1. Every load that does not
// depend on another block row's result is issued before the first wait (b_i,
// the first 64 entries of U_ii^-1, the first off-diagonal chunk and, once its
// columns are in, its gathers of x), so a level costs two dependent memory
// round trips.
template<typename VT>
__device__ __forceinline__ void block_row_solve(int4 t0, int4 t1, bool offdiag,
    const int32_t* o_col, const VT* o_val, const uint8_t* o_lr, const int32_t* i_col,
    const VT* i_val, const uint8_t* i_lr, const double* b, const double* xsrc, double* xout,
    double* y, double* res)
{
    const int lane = lane_id();
    const int fr = t0.x, h = t0.y, ob = t0.z, oe = offdiag ? t0.w : t0.z, ib = t1.x, ie = t1.y;
    // loads run from clamped (in-range) indices without branches, and the
    // lanes past the end are masked afterwards, so the compiler issues them
    // all before the first wait
    const double bo_raw = b[fr + min(lane, h - 1)];
    const double bo = (lane < h) ? bo_raw : 0.0;
    // the first 64 entries of U_ii^-1 (every block row has at least its
    // diagonal there), selected only where they are used
    const int ei = ib + min(lane, max(ie - ib - 1, 0));
    const int ic0 = i_col[ei];
    const VT iv0 = i_val[ei];
    const int il0 = i_lr[ei];
    int c[TRSV_PF], sg[TRSV_PF];
    VT v[TRSV_PF];
    double xv[TRSV_PF];
    auto load = [&](int e0) {
        int cr[TRSV_PF], sr[TRSV_PF];
        VT vr[TRSV_PF];
#pragma unroll
        for(int k = 0; k < TRSV_PF; ++k) {
            const int e = min(e0 + 64 * k + lane, oe - 1);
            cr[k] = o_col[e];
            vr[k] = o_val[e];
            sr[k] = o_lr[e];
        }
        double xr[TRSV_PF];
#pragma unroll
        for(int k = 0; k < TRSV_PF; ++k) xr[k] = xsrc[cr[k]];
#pragma unroll
        for(int k = 0; k < TRSV_PF; ++k) {
            const bool in = e0 + 64 * k + lane < oe;
            c[k] = in ? cr[k] : 0;
            v[k] = in ? vr[k] : (VT) 0;
            sg[k] = in ? sr[k] : 64;
            xv[k] = in ? xr[k] : 0.0;
        }
    };
    if(ob < oe) load(ob);
    if(lane < h) { y[lane] = bo; res[lane] = 0; }
    for(int e0 = ob; e0 < oe; e0 += 64 * TRSV_PF) {
        if(e0 != ob) load(e0);
        wave_sync();
#pragma unroll
        for(int k = 0; k < TRSV_PF; ++k) {
            if(e0 + 64 * k >= oe) break;
            double p = seg_scan_runs_fma((double) v[k], xv[k], sg[k]);
            const int nsg = next_lane(sg[k], -1);
            if(sg[k] < 64 && (lane == 63 || nsg != sg[k])) y[sg[k]] -= p;
            wave_sync();
        }
    }
    wave_sync();
    // every lane takes part (the general form scans across the wave)
    const double r = apply_inverse<VT>(fr, h, ib, ie, ic0, iv0, il0, i_col, i_val, i_lr, y, res);
    if(lane < h) xout[fr + lane] = r;
}

template<typename VT>
__global__ __launch_bounds__(64) void k_trsv_level(const int4* task, const int32_t* o_col,
    const VT* o_val, const uint8_t* o_lr, const int32_t* i_col, const VT* i_val,
    const uint8_t* i_lr, double* x)
{
    __shared__ double y[33];     // FP64 accumulation for both factor precisions
    __shared__ double res[33];
    block_row_solve<VT>(task[2 * blockIdx.x], task[2 * blockIdx.x + 1], true, o_col, o_val,
        o_lr, i_col, i_val, i_lr, x, x, x, y, res);
}

// One block-Jacobi sweep over all block rows (solve_U_jacobi /
// solve_Ut_jacobi, block_ildlt.impl.cuh:2454-2493, 2637-2676, with
// k_solve_block_stripe_level): x_out_i = U_ii^-1 (b_i - sum_j U_ij x_in_j).
// x_in == nullptr is the first sweep from x_0 = 0. The per-row sums run in
// the same order as in k_trsv_level, so once x_in is the exact solution the
// sweep reproduces the level-scheduled solve bit for bit.
template<typename VT>
__global__ __launch_bounds__(64) void k_jacobi_sweep(const int4* task, const int32_t* o_col,
    const VT* o_val, const uint8_t* o_lr, const int32_t* i_col, const VT* i_val,
    const uint8_t* i_lr, const double* b, const double* xin, double* xout)
{
    __shared__ double y[33];
    __shared__ double res[33];
    block_row_solve<VT>(task[2 * blockIdx.x], task[2 * blockIdx.x + 1], xin != nullptr, o_col,
        o_val, o_lr, i_col, i_val, i_lr, b, xin, xout, y, res);
}

// Whole triangular solve in one launch of G resident waves:
// wave g takes the block rows at processing indices g, g + G, g + 2G, ... of
// the level order (reverse level order for the U solve), so every block row
// it depends on has a smaller index. The lowest unfinished index always has
// its inputs published and every wave is resident, so the solve drains.
// Results are published value by value: before the solve every entry of pub
// holds CHAIN_EMPTY, a signalling NaN that no arithmetic produces, and a
// block row writes its x values there with 8-byte agent-coherent stores (the
// L2 of each XCD is separate). A reader loads the values it needs with
// agent-coherent loads and takes each one once it differs from CHAIN_EMPTY,
// reloading the missing ones until all are there; no flag and no
// store-completion wait sits between a producer and its readers. While a
// block row waits and computes, the wave already loads the task and first
// entries of its next one. Each wave also presets its rows of next_pub, the
// buffer of the next single-launch solve. Every wait is bounded: past
// SPIN_LIMIT polls the wave records an error in *err and goes on.
constexpr int SPIN_LIMIT = 1 << 22;
constexpr unsigned long long CHAIN_EMPTY = 0x7ff4c0de5eed1234ull;

__device__ __forceinline__ double pub_load(const double* p)
{
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool published(double v)
{
    return (unsigned long long) __double_as_longlong(v) != CHAIN_EMPTY;
}

// waits until every lane's value xv[k] (column c[k]) is published; rdy[k]
// on entry: xv[k] already usable
__device__ void wait_published(const double* pub, const int (&c)[TRSV_PF], double (&xv)[TRSV_PF],
    bool (&rdy)[TRSV_PF], int32_t* err, int& spins)
{
    bool all = true;
#pragma unroll
    for(int k = 0; k < TRSV_PF; ++k) all = all && rdy[k];
    while(wave_ballot(!all)) {
        __builtin_amdgcn_s_sleep(1);
        if(++spins > SPIN_LIMIT) {
            // err is pinned host memory: the host reads it without a copy
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            return;
        }
        all = true;
#pragma unroll
        for(int k = 0; k < TRSV_PF; ++k) {
            if(!rdy[k]) {
                xv[k] = pub_load(pub + c[k]);
                rdy[k] = published(xv[k]);
            }
            all = all && rdy[k];
        }
    }
}

// what a block row loads before it can wait: its first TRSV_PF chunks of
// off-diagonal entries (clamped in range), its own right-hand side and the
// first 64 entries of U_ii^-1
template<typename VT>
struct ChainHead {
    int c[TRSV_PF], sg[TRSV_PF];
    VT v[TRSV_PF];
    double bo;
    int ic, il;
    VT iv;
};

// The right-hand side of one block row as K src (rows fr .. fr + h - 1 of a
// CSR matrix), formed exactly as k_csr_spmv forms those rows: eight lanes
// per row, lane sub summing entries s + sub, s + sub + 8, ... by fma in that
// order, then the xor tree over the eight lanes. Four rounds of eight rows
// issue their loads together. Returns row fr + lane's value (lanes < h).
struct CsrRhs {
    const int32_t* rp;
    const int32_t* ci;
    const double* cv;
    const double* src;
};

constexpr int RHS_U = 3;   // entries per lane and row whose loads go together

__device__ double block_rhs(const CsrRhs& k, int fr, int h)
{
    const int lane = lane_id();
    const int sub = lane & 7, g = lane >> 3;
    int s[4], e[4];
#pragma unroll
    for(int r = 0; r < 4; ++r) {
        const int row = fr + min(8 * r + g, h - 1);
        s[r] = k.rp[row];
        e[r] = k.rp[row + 1];
    }
    double acc[4];
#pragma unroll
    for(int r = 0; r < 4; ++r) acc[r] = 0.0;
    for(int q0 = sub;; q0 += 8 * RHS_U) {
        int c[4][RHS_U];
        double v[4][RHS_U], xv[4][RHS_U];
#pragma unroll
        for(int r = 0; r < 4; ++r)
#pragma unroll
            for(int u = 0; u < RHS_U; ++u) {
                const int q = max(min(s[r] + q0 + 8 * u, e[r] - 1), 0);
                const int cq = k.ci[q];
                c[r][u] = (s[r] + q0 + 8 * u < e[r]) ? cq : 0;
                v[r][u] = k.cv[q];
            }
#pragma unroll
        for(int r = 0; r < 4; ++r)
#pragma unroll
            for(int u = 0; u < RHS_U; ++u) xv[r][u] = k.src[c[r][u]];
        bool more = false;
#pragma unroll
        for(int r = 0; r < 4; ++r) {
#pragma unroll
            for(int u = 0; u < RHS_U; ++u)
                if(s[r] + q0 + 8 * u < e[r]) acc[r] = fma(v[r][u], xv[r][u], acc[r]);
            more = more || s[r] + q0 + 8 * RHS_U < e[r];
        }
        if(wave_ballot(more) == 0) break;
    }
    double mine = 0.0;
#pragma unroll
    for(int r = 0; r < 4; ++r) {
        double a = acc[r];
#pragma unroll
        for(int o = 4; o > 0; o >>= 1) a += __shfl_xor(a, o, 8);
        const double row_val = 1.0 * a;          // y = alpha (A x), alpha = 1
        const double got = __shfl(row_val, 8 * (lane & 7), 64);
        if((lane >> 3) == r) mine = got;
    }
    return mine;
}

template<typename VT>
__device__ __forceinline__ void chain_fetch(int4 t0, int4 t1, const int32_t* o_col, const VT* o_val,
    const uint8_t* o_lr, const int32_t* i_col, const VT* i_val, const uint8_t* i_lr,
    const double* x, bool rhs_fused, ChainHead<VT>& f)
{
    const int lane = lane_id();
    const int fr = t0.x, h = t0.y, ob = t0.z, oe = t0.w, ib = t1.x, ie = t1.y;
#pragma unroll
    for(int k = 0; k < TRSV_PF; ++k) {
        const int e = max(min(ob + 64 * k + lane, oe - 1), 0);
        f.c[k] = o_col[e];
        f.v[k] = o_val[e];
        f.sg[k] = o_lr[e];
    }
    // this block row's own right-hand side: no other wave writes x (with a
    // fused product it is formed after the current block row instead)
    f.bo = rhs_fused ? 0.0 : x[fr + min(lane, h - 1)];
    const int ei = ib + min(lane, max(ie - ib - 1, 0));
    f.ic = i_col[ei];
    f.iv = i_val[ei];
    f.il = i_lr[ei];
}

template<typename VT, bool FUSED>
__global__ __launch_bounds__(64) void k_trsv_chain(const int4* task, int32_t nb, int32_t reverse,
    const int32_t* o_col, const VT* o_val, const uint8_t* o_lr, const int32_t* i_col,
    const VT* i_val, const uint8_t* i_lr, double* pub, double* next_pub, int32_t* err, double* x,
    CsrRhs rhs)
{
    __shared__ double y[33];     // FP64 accumulation for both factor precisions
    __shared__ double res[33];
    const int lane = lane_id();
    const int G = gridDim.x;
    constexpr bool fused = FUSED;
    const double empty = __longlong_as_double((long long) CHAIN_EMPTY);
    auto task_at = [&](int w, int4& a, int4& b) {
        const int q = min(w, nb - 1);
        const int pos = reverse ? (nb - 1 - q) : q;
        a = task[2 * pos];
        b = task[2 * pos + 1];
    };
    int spins = 0;
    int w = blockIdx.x;
    int4 t0, t1;
    task_at(w, t0, t1);
    ChainHead<VT> cur;
    chain_fetch<VT>(t0, t1, o_col, o_val, o_lr, i_col, i_val, i_lr, x, fused, cur);
    if(fused) cur.bo = block_rhs(rhs, t0.x, t0.y);
    for(;;) {
        int4 n0, n1;
        task_at(w + G, n0, n1);
        ChainHead<VT> nxt;
        bool fetched = false;
        const int fr = t0.x, h = t0.y, ob = t0.z, oe = t0.w, ib = t1.x, ie = t1.y;
        if(lane < h) next_pub[fr + lane] = empty;
        wave_sync();
        if(lane < h) { y[lane] = cur.bo; res[lane] = 0; }
        for(int e0 = ob; e0 < oe; e0 += 64 * TRSV_PF) {
            int c[TRSV_PF], sg[TRSV_PF];
            VT v[TRSV_PF];
            if(e0 == ob) {
#pragma unroll
                for(int k = 0; k < TRSV_PF; ++k) { c[k] = cur.c[k]; v[k] = cur.v[k]; sg[k] = cur.sg[k]; }
            } else {
#pragma unroll
                for(int k = 0; k < TRSV_PF; ++k) {
                    const int e = min(e0 + 64 * k + lane, oe - 1);
                    c[k] = o_col[e];
                    v[k] = o_val[e];
                    sg[k] = o_lr[e];
                }
            }
            double xv[TRSV_PF];
            bool rdy[TRSV_PF];
#pragma unroll
            for(int k = 0; k < TRSV_PF; ++k) xv[k] = pub_load(pub + c[k]);
#pragma unroll
            for(int k = 0; k < TRSV_PF; ++k) {
                const bool in = e0 + 64 * k + lane < oe;
                v[k] = in ? v[k] : (VT) 0;
                sg[k] = in ? sg[k] : 64;
                rdy[k] = !in || published(xv[k]);
            }
            wait_published(pub, c, xv, rdy, err, spins);
            if(!fetched) {
                chain_fetch<VT>(n0, n1, o_col, o_val, o_lr, i_col, i_val, i_lr, x, fused, nxt);
                fetched = true;
            }
            wave_sync();
#pragma unroll
            for(int k = 0; k < TRSV_PF; ++k) {
                if(e0 + 64 * k >= oe) break;
                double p = seg_scan_runs_fma((double) v[k], sg[k] < 64 ? xv[k] : 0.0, sg[k]);
                const int nsg = next_lane(sg[k], -1);
                if(sg[k] < 64 && (lane == 63 || nsg != sg[k])) y[sg[k]] -= p;
                wave_sync();
            }
        }
        if(!fetched) chain_fetch<VT>(n0, n1, o_col, o_val, o_lr, i_col, i_val, i_lr, x, fused, nxt);
        wave_sync();
        const double r = apply_inverse<VT>(fr, h, ib, ie, cur.ic, cur.iv, cur.il, i_col, i_val, i_lr, y,
            res);
        if(lane < h) {
            __hip_atomic_store(pub + fr + lane, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            x[fr + lane] = r;
        }
        w += G;
        if(w >= nb) break;
        // the next block row's K src: its loads overlap the wait for its
        // inputs, which rarely have all arrived by now
        if(fused) nxt.bo = block_rhs(rhs, n0.x, n0.y);
        t0 = n0;
        t1 = n1;
        cur = nxt;
    }
}

// The first block-Jacobi sweep with b = K src: each block row forms its rows
// of K src as block_rhs does (the values k_csr_spmv gives), keeps them in b
// for the later sweeps and solves from x_0 = 0 as k_jacobi_sweep does.
template<typename VT>
__global__ __launch_bounds__(64) void k_jacobi_first_product(const int4* task,
    const int32_t* o_col, const VT* o_val, const uint8_t* o_lr, const int32_t* i_col,
    const VT* i_val, const uint8_t* i_lr, CsrRhs rhs, double* b, double* xout)
{
    __shared__ double y[33];
    __shared__ double res[33];
    const int4 t0 = task[2 * blockIdx.x];
    const double bo = block_rhs(rhs, t0.x, t0.y);
    // each lane reads back only the row it wrote (lanes >= h are masked)
    if(lane_id() < t0.y) b[t0.x + lane_id()] = bo;
    block_row_solve<VT>(t0, task[2 * blockIdx.x + 1], false, o_col, o_val, o_lr, i_col, i_val,
        i_lr, b, nullptr, xout, y, res);
}

// Batched block-Jacobi sweeps. Item {s, c} is block rows JACOBI_ROWS c ..
// JACOBI_ROWS c + JACOBI_ROWS - 1 of system s, one after the other in one
// wave; each block row is computed exactly as block_row_solve computes it
// (the workgroup of k_jacobi_sweep for that block row): the same masked
// 64-entry chunks, segmented sums in chunk order and U_ii^-1 rows. What a
// block row reads before its first wait (its first chunk, the x values of
// that chunk, its b and its first 64 entries of U_ii^-1) is loaded while the
// block row before it computes; the task records of all the item's block
// rows come in with one load. tr selects the layout (0 the U solve, 1 the
// U' solve). b may be the output vector itself: a block row reads its own
// rows of b before it writes its own rows of the output, and no other block
// row reads them.
constexpr int JACOBI_ROWS = JACOBI_BATCH_ROWS;

template<typename VT>
struct JacobiHead {
    int c, sg;
    VT v;
    double xv;
    double bo;
    int ic, il;
    VT iv;
};

template<typename VT>
__global__ __launch_bounds__(64) void bk_jacobi_sweep(const int2* items, const BatchSys* sys,
    int32_t tr, int32_t b_id, int32_t in, int32_t out)
{
    __shared__ double y[33];
    __shared__ double res[33];
    const int lane = lane_id();
    const int2 it = items[blockIdx.x];
    if(it.x < 0) return;   // a hole of an XCD-aware list
    const BatchSys& S = sys[it.x];
    const int4* task = S.task[tr];
    const int32_t* o_col = S.o_col[tr];
    const VT* o_val = (const VT*) S.o_val[tr];
    const uint8_t* o_lr = S.o_lr[tr];
    const int32_t* i_col = S.i_col[tr];
    const VT* i_val = (const VT*) S.i_val[tr];
    const uint8_t* i_lr = S.i_lr[tr];
    const double* xsrc = in == BV_NONE ? nullptr : batch_vec(S, in);
    const double* b = batch_vec(S, b_id);
    double* xout = batch_vec(S, out);
    const bool offdiag = xsrc != nullptr;
    const int r0 = it.y * JACOBI_ROWS;
    const int cnt = min(JACOBI_ROWS, S.nb - r0);
    // lane 2 j holds block row r0 + j's first task record, lane 2 j + 1 its
    // second
    const int4 tk = task[2 * r0 + min(lane, 2 * cnt - 1)];
    auto rec = [&](int j, int4& t0, int4& t1) {
        t0.x = __shfl(tk.x, 2 * j);
        t0.y = __shfl(tk.y, 2 * j);
        t0.z = __shfl(tk.z, 2 * j);
        t0.w = __shfl(tk.w, 2 * j);
        t1.x = __shfl(tk.x, 2 * j + 1);
        t1.y = __shfl(tk.y, 2 * j + 1);
    };
    // the loads of block row j that do not depend on other rows' results
    auto head = [&](int j, JacobiHead<VT>& H) {
        int4 t0, t1;
        rec(j, t0, t1);
        const int fr = t0.x, h = t0.y, ob = t0.z, oe = offdiag ? t0.w : t0.z;
        const int ib = t1.x, ie = t1.y;
        H.bo = b[fr + min(lane, h - 1)];
        const int ei = ib + min(lane, max(ie - ib - 1, 0));
        H.ic = i_col[ei];
        H.iv = i_val[ei];
        H.il = i_lr[ei];
        H.c = 0;
        H.v = (VT) 0;
        H.sg = 64;
        if(ob < oe) {
            const int e = min(ob + lane, oe - 1);
            H.c = o_col[e];
            H.v = o_val[e];
            H.sg = o_lr[e];
        }
    };
    auto gather = [&](int j, JacobiHead<VT>& H) {
        int4 t0, t1;
        rec(j, t0, t1);
        const int ob = t0.z, oe = offdiag ? t0.w : t0.z;
        H.xv = 0.0;
        if(ob < oe) {
            const double xr = xsrc[H.c];
            const bool in0 = ob + lane < oe;
            H.c = in0 ? H.c : 0;
            H.v = in0 ? H.v : (VT) 0;
            H.sg = in0 ? H.sg : 64;
            H.xv = in0 ? xr : 0.0;
        }
    };
    auto apply_chunk = [&](int c, VT v, int sg, double xv) {
        (void) c;
        const double p = seg_scan_runs_fma((double) v, xv, sg);
        const int nsg = next_lane(sg, -1);
        if(sg < 64 && (lane == 63 || nsg != sg)) y[sg] -= p;
        wave_sync();
    };
    auto compute = [&](int j, const JacobiHead<VT>& H) {
        int4 t0, t1;
        rec(j, t0, t1);
        const int fr = t0.x, h = t0.y, ob = t0.z, oe = offdiag ? t0.w : t0.z;
        const int ib = t1.x, ie = t1.y;
        const double bo = (lane < h) ? H.bo : 0.0;
        if(lane < h) { y[lane] = bo; res[lane] = 0; }
        wave_sync();
        if(ob < oe) {
            apply_chunk(H.c, H.v, H.sg, H.xv);
            for(int e0 = ob + 64; e0 < oe; e0 += 64) {
                const int e = min(e0 + lane, oe - 1);
                const int cr = o_col[e];
                const VT vr = o_val[e];
                const int sr = o_lr[e];
                const double xr = xsrc[cr];
                const bool in0 = e0 + lane < oe;
                apply_chunk(in0 ? cr : 0, in0 ? vr : (VT) 0, in0 ? sr : 64, in0 ? xr : 0.0);
            }
        }
        wave_sync();
        const double r = apply_inverse<VT>(fr, h, ib, ie, H.ic, H.iv, H.il, i_col, i_val, i_lr,
            y, res);
        if(lane < h) xout[fr + lane] = r;
        wave_sync();
    };
    JacobiHead<VT> A, B;
    head(0, A);
    gather(0, A);
    for(int j = 0; j < cnt; j += 2) {
        if(j + 1 < cnt) head(j + 1, B);
        compute(j, A);
        if(j + 1 >= cnt) break;
        gather(j + 1, B);
        if(j + 2 < cnt) head(j + 2, A);
        compute(j + 1, B);
        if(j + 2 < cnt) gather(j + 2, A);
    }
}

// Batched layout build: item {s, q} is workgroup q of the single-system
// launch on system s.
template<typename VT, bool FILL>
__global__ __launch_bounds__(64) void bk_plan_rows(const PlanSys<VT>* ps, const int2* items,
    int32_t hm)
{
    extern __shared__ double plan_x[];
    const int2 it = items[blockIdx.x];
    const PlanSys<VT>& P = ps[it.x];
    plan_rows_body<VT, FILL>(P.t, P.val_out, P.p, it.y, plan_x, hm);
}

template<typename VT, bool FILL>
__global__ __launch_bounds__(64) void bk_plan_cols(const PlanSys<VT>* ps, const int2* items)
{
    const int2 it = items[blockIdx.x];
    const PlanSys<VT>& P = ps[it.x];
    plan_cols_body<VT, FILL>(P.t, P.val_out, P.p, it.y);
}

template<typename VT>
__device__ __forceinline__ int32_t* plan_ptr(const SolvePlan<VT>& p, int q)
{
    return q == 0 ? p.uo_ptr : q == 1 ? p.ui_ptr : q == 2 ? p.to_ptr : p.ti_ptr;
}

template<typename VT>
__global__ __launch_bounds__(PS_TPB) void bk_chunk_sums(const PlanSys<VT>* ps, const int2* items,
    int32_t q)
{
    const int2 it = items[blockIdx.x];
    const PlanSys<VT>& P = ps[it.x];
    chunk_sums_body(plan_ptr(P.p, q), P.t.n, P.scan_tmp, it.y);
}

template<typename VT>
__global__ __launch_bounds__(PS_TPB) void bk_chunk_scan(const PlanSys<VT>* ps, const int2* items,
    int32_t q)
{
    const int2 it = items[blockIdx.x];
    const PlanSys<VT>& P = ps[it.x];
    chunk_scan_body(plan_ptr(P.p, q), P.t.n, P.scan_tmp, P.totals + q, it.y,
        (P.t.n + PS_CHUNK - 1) / PS_CHUNK);
}

template<typename VT>
__global__ void bk_plan_tasks(const PlanSys<VT>* ps, const int2* items, int32_t transpose)
{
    const int2 it = items[blockIdx.x];
    const PlanSys<VT>& P = ps[it.x];
    if(transpose)
        plan_tasks_body(P.t.nb, P.t.lvl_rows, P.t.bs, P.p.to_ptr, P.p.ti_ptr, P.p.task_t, it.y);
    else
        plan_tasks_body(P.t.nb, P.t.lvl_rows, P.t.bs, P.p.uo_ptr, P.p.ui_ptr, P.p.task_u, it.y);
}

__global__ void k_fill_bits(double* p, int64_t n, unsigned long long bits)
{
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if(i < n) p[i] = __longlong_as_double((long long) bits);
}

__global__ void k_solve_D(int32_t n, const double* tri, const int32_t* is_piv, const double* b,
    double* x)
{
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if(g >= n) return;
    const int this_piv = is_piv[g];
    const int next_piv = (g < n - 1) ? is_piv[g + 1] : 1;
    const double p11 = tri[3 * (int64_t) g + 1];
    const double p12 = tri[3 * (int64_t) g + 2];
    const double p22 = !next_piv ? tri[3 * (int64_t) (g + 1) + 1] : 1.0;
    const double r1 = b[g];
    const double r2 = (g < n - 1) ? b[g + 1] : 0.0;
    const double gr = sqrt(fma(p11, p11, p12 * p12));
    const double c = p11 / gr, s = p12 / gr;
    double g1 = fma(c, r1, s * r2);
    double g2 = fma(-s, r1, c * r2);
    const double t12 = fma(c, p12, s * p22);
    const double t22 = fma(-s, p12, c * p22);
    g2 = g2 / t22;
    g1 = fma(-t12, g2, g1) / gr;
    if(this_piv) x[g] = g1;
    if(!next_piv) x[g + 1] = g2;
}

}  // namespace

template<typename VT>
void solve_plan_count(const SolveTopo& t, const VT* val_out, VT*, SolvePlan<VT>& p, void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    k_plan_rows<VT, false><<<t.nb, 64, sizeof(double) * 1024, st>>>(t, val_out, p);
    k_plan_cols<VT, false><<<t.nb, 64, 0, st>>>(t, val_out, p);
}

template<typename VT>
void solve_plan_fill(const SolveTopo& t, const VT* val_out, const VT*, SolvePlan<VT>& p,
    void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    k_plan_rows<VT, true><<<t.nb, 64, sizeof(double) * 1024, st>>>(t, val_out, p);
    k_plan_cols<VT, true><<<t.nb, 64, 0, st>>>(t, val_out, p);
    const unsigned g = (unsigned) ((t.nb + 255) / 256);
    k_plan_tasks<<<g, 256, 0, st>>>(t.nb, t.lvl_rows, t.bs, p.uo_ptr, p.ui_ptr, p.task_u);
    k_plan_tasks<<<g, 256, 0, st>>>(t.nb, t.lvl_rows, t.bs, p.to_ptr, p.ti_ptr, p.task_t);
}

size_t scan_tmp_ints(int32_t n) { return (size_t) ((n + PS_CHUNK - 1) / PS_CHUNK) + 1; }

void scan_counts(int32_t* ptr, int32_t n, int32_t* total_dev, int32_t* tmp, void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    const unsigned nch = (unsigned) ((n + PS_CHUNK - 1) / PS_CHUNK);
    if(nch == 0) {
        k_chunk_scan<<<1, PS_TPB, 0, st>>>(ptr, n, tmp, total_dev);
        return;
    }
    k_chunk_sums<<<nch, PS_TPB, 0, st>>>(ptr, n, tmp);
    k_chunk_scan<<<nch, PS_TPB, 0, st>>>(ptr, n, tmp, total_dev);
}

template<typename VT>
void trsv_upper(const SolveTopo& t, const SolvePlan<VT>& p, double* x, void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    for(int l = t.num_levels - 1; l >= 0; --l) {
        const int lo = t.lvl_off_h[l], cnt = t.lvl_off_h[l + 1] - lo;
        if(cnt > 0)
            k_trsv_level<VT><<<cnt, 64, 0, st>>>(p.task_u + 2 * lo, p.uo_col, p.uo_val, p.uo_lr,
                p.ui_col, p.ui_val, p.ui_lr, x);
    }
}

template<typename VT>
void trsv_upper_t(const SolveTopo& t, const SolvePlan<VT>& p, double* x, void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    for(int l = 0; l < t.num_levels; ++l) {
        const int lo = t.lvl_off_h[l], cnt = t.lvl_off_h[l + 1] - lo;
        if(cnt > 0)
            k_trsv_level<VT><<<cnt, 64, 0, st>>>(p.task_t + 2 * lo, p.to_col, p.to_val, p.to_lr,
                p.ti_col, p.ti_val, p.ti_lr, x);
    }
}

__global__ void k_perm_pull(int32_t n, const int32_t* perm, const double* in, double* out)
{
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if(g < n) out[g] = in[perm[g]];
}

__global__ void k_perm_push(int32_t n, const int32_t* perm, const double* in, double* out)
{
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if(g < n) out[perm[g]] = in[g];
}

void permute_pull(int32_t n, const int32_t* perm, const double* in, double* out, void* stream)
{
    if(n > 0) k_perm_pull<<<(n + 255) / 256, 256, 0, (hipStream_t) stream>>>(n, perm, in, out);
}

void permute_push(int32_t n, const int32_t* perm, const double* in, double* out, void* stream)
{
    if(n > 0) k_perm_push<<<(n + 255) / 256, 256, 0, (hipStream_t) stream>>>(n, perm, in, out);
}

void solve_D(int32_t n, const double* tri, const int32_t* is_piv, const double* b, double* x,
    void* stream)
{
    k_solve_D<<<(n + 255) / 256, 256, 0, (hipStream_t) stream>>>(n, tri, is_piv, b, x);
}

template void solve_plan_count<double>(const SolveTopo&, const double*, double*, SolvePlan<double>&, void*);
template void solve_plan_count<float>(const SolveTopo&, const float*, float*, SolvePlan<float>&, void*);
template void solve_plan_fill<double>(const SolveTopo&, const double*, const double*, SolvePlan<double>&, void*);
template void solve_plan_fill<float>(const SolveTopo&, const float*, const float*, SolvePlan<float>&, void*);
// Waves of k_trsv_chain per CU. A solve needs all of its waves resident.
// The kernel's registers let a CU hold 16 of its waves, so with 4 per CU
// TRSV_SINGLE_MAX_CONCURRENT = 4 such solves fit on the GPU at the same
// time; other kernels end on their own.
constexpr int CHAIN_WAVES_PER_CU = 4;

// grid of k_trsv_chain: CHAIN_WAVES_PER_CU waves per CU, fewer when the
// kernel's occupancy would not hold TRSV_SINGLE_MAX_CONCURRENT solves (CU
// count and occupancy queried once per device)
template<typename VT>
int chain_grid(int32_t nb)
{
    static int cus[64] = {0}, occ[64] = {0};
    int dev = 0;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    if(cus[dev] == 0) {
        hipDeviceProp_t pr;
        int o = 0;
        if(hipGetDeviceProperties(&pr, dev) != hipSuccess) return 0;
        if(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_trsv_chain<VT, true>, 64, 0) != hipSuccess)
            return 0;
        occ[dev] = o;
        cus[dev] = pr.multiProcessorCount;
    }
    const int per = max(1, min(CHAIN_WAVES_PER_CU, occ[dev] / TRSV_SINGLE_MAX_CONCURRENT));
    const int g = cus[dev] * per;
    return min(g, (int) nb);
}

template<typename VT>
hipError_t trsv_chain(const int4* task, int32_t nb, int32_t reverse, const int32_t* o_col,
    const VT* o_val, const uint8_t* o_lr, const int32_t* i_col, const VT* i_val,
    const uint8_t* i_lr, double* pub, double* next_pub, int32_t* err, double* x, CsrRhs rhs,
    hipStream_t st)
{
    const int g = chain_grid<VT>(nb);
    if(g <= 0) return hipErrorInvalidConfiguration;
    if(rhs.rp)
        k_trsv_chain<VT, true><<<g, 64, 0, st>>>(task, nb, reverse, o_col, o_val, o_lr, i_col, i_val,
            i_lr, pub, next_pub, err, x, rhs);
    else
        k_trsv_chain<VT, false><<<g, 64, 0, st>>>(task, nb, reverse, o_col, o_val, o_lr, i_col,
            i_val, i_lr, pub, next_pub, err, x, rhs);
    return hipGetLastError();
}

template<typename VT>
int trsv_upper_single(const SolveTopo& t, const SolvePlan<VT>& p, double* x, int32_t parity,
    void* stream)
{
    if(t.nb <= 0) return 0;
    double* pub = t.chain_pub + (parity ? t.n : 0);
    double* next = t.chain_pub + (parity ? 0 : t.n);
    return (int) trsv_chain<VT>(p.task_u, t.nb, 1, p.uo_col, p.uo_val, p.uo_lr, p.ui_col, p.ui_val,
        p.ui_lr, pub, next, t.err, x, CsrRhs{nullptr, nullptr, nullptr, nullptr},
        (hipStream_t) stream);
}

template<typename VT>
int trsv_upper_t_single(const SolveTopo& t, const SolvePlan<VT>& p, double* x, int32_t parity,
    void* stream)
{
    return trsv_upper_t_single_product<VT>(t, p, nullptr, nullptr, x, parity, stream);
}

template<typename VT>
int trsv_upper_t_single_product(const SolveTopo& t, const SolvePlan<VT>& p, const DevCsr* K,
    const double* src, double* x, int32_t parity, void* stream)
{
    if(t.nb <= 0) return 0;
    double* pub = t.chain_pub + (parity ? t.n : 0);
    double* next = t.chain_pub + (parity ? 0 : t.n);
    const CsrRhs rhs = K ? CsrRhs{K->row_ptr, K->col, K->val, src}
                         : CsrRhs{nullptr, nullptr, nullptr, nullptr};
    return (int) trsv_chain<VT>(p.task_t, t.nb, 0, p.to_col, p.to_val, p.to_lr, p.ti_col, p.ti_val,
        p.ti_lr, pub, next, t.err, x, rhs, (hipStream_t) stream);
}

void chain_reset(double* pub, int64_t count, void* stream)
{
    if(count <= 0) return;
    k_fill_bits<<<(unsigned) ((count + 255) / 256), 256, 0, (hipStream_t) stream>>>(pub, count,
        CHAIN_EMPTY);
}

template<typename VT>
void jacobi_sweeps(const SolveTopo& t, const SolvePlan<VT>& p, bool transpose, double* x,
    int32_t sweeps, double* b_tmp, double* x_a, double* x_b, void* stream)
{
    jacobi_sweeps_product<VT>(t, p, transpose, nullptr, nullptr, x, sweeps, b_tmp, x_a, x_b, stream);
}

template<typename VT>
void jacobi_sweeps_product(const SolveTopo& t, const SolvePlan<VT>& p, bool transpose,
    const DevCsr* K, const double* src, double* x, int32_t sweeps, double* b_tmp, double* x_a,
    double* x_b, void* stream)
{
    if(t.nb <= 0 || sweeps <= 0) return;
    hipStream_t st = (hipStream_t) stream;
    const int4* task = transpose ? p.task_t : p.task_u;
    const int32_t* oc = transpose ? p.to_col : p.uo_col;
    const VT* ov = transpose ? p.to_val : p.uo_val;
    const uint8_t* ol = transpose ? p.to_lr : p.uo_lr;
    const int32_t* ic = transpose ? p.ti_col : p.ui_col;
    const VT* iv = transpose ? p.ti_val : p.ui_val;
    const uint8_t* il = transpose ? p.ti_lr : p.ui_lr;
    const double* in = nullptr;
    for(int32_t s = 0; s < sweeps; ++s) {
        double* out = (s == sweeps - 1) ? x : ((s & 1) ? x_b : x_a);
        if(s == 0 && K)
            k_jacobi_first_product<VT><<<t.nb, 64, 0, st>>>(task, oc, ov, ol, ic, iv, il,
                CsrRhs{K->row_ptr, K->col, K->val, src}, b_tmp, out);
        else
            k_jacobi_sweep<VT><<<t.nb, 64, 0, st>>>(task, oc, ov, ol, ic, iv, il, b_tmp, in, out);
        in = out;
    }
}

template void jacobi_sweeps<double>(const SolveTopo&, const SolvePlan<double>&, bool, double*, int32_t, double*, double*, double*, void*);
template void jacobi_sweeps<float>(const SolveTopo&, const SolvePlan<float>&, bool, double*, int32_t, double*, double*, double*, void*);
template void jacobi_sweeps_product<double>(const SolveTopo&, const SolvePlan<double>&, bool, const DevCsr*, const double*, double*, int32_t, double*, double*, double*, void*);
template void jacobi_sweeps_product<float>(const SolveTopo&, const SolvePlan<float>&, bool, const DevCsr*, const double*, double*, int32_t, double*, double*, double*, void*);
template int trsv_upper_single<double>(const SolveTopo&, const SolvePlan<double>&, double*, int32_t, void*);
template int trsv_upper_single<float>(const SolveTopo&, const SolvePlan<float>&, double*, int32_t, void*);
template int trsv_upper_t_single<double>(const SolveTopo&, const SolvePlan<double>&, double*, int32_t, void*);
template int trsv_upper_t_single<float>(const SolveTopo&, const SolvePlan<float>&, double*, int32_t, void*);
template int trsv_upper_t_single_product<double>(const SolveTopo&, const SolvePlan<double>&, const DevCsr*, const double*, double*, int32_t, void*);
template int trsv_upper_t_single_product<float>(const SolveTopo&, const SolvePlan<float>&, const DevCsr*, const double*, double*, int32_t, void*);
template void trsv_upper<double>(const SolveTopo&, const SolvePlan<double>&, double*, void*);
template void trsv_upper<float>(const SolveTopo&, const SolvePlan<float>&, double*, void*);
template void trsv_upper_t<double>(const SolveTopo&, const SolvePlan<double>&, double*, void*);
template void trsv_upper_t<float>(const SolveTopo&, const SolvePlan<float>&, double*, void*);


// the batched form of Bildlt::jacobi (copy, then jacobi_sweeps) and of
// left_product (K src, then the sweeps): the same sweeps and values. The
// product is the batched CSR product, whose rows equal those the fused
// first sweep forms; the sweeps read b from the target itself instead of a
// copy.
template<typename VT>
void batch_jacobi(const BatchLists& L, const BatchSys* sys, bool transpose, int32_t sweeps,
    int32_t target, int32_t product_src, hipStream_t st)
{
    if(L.blk.count <= 0 || sweeps <= 0) return;
    if(product_src != BV_NONE) batch_csr_spmv(L, sys, product_src, target, st);
    const int32_t tr = transpose ? 1 : 0;
    int32_t in = BV_NONE;
    for(int32_t s = 0; s < sweeps; ++s) {
        const int32_t out = (s == sweeps - 1) ? target : ((s & 1) ? BV_JB2 : BV_JB1);
        bk_jacobi_sweep<VT><<<(unsigned) L.blk.count, 64, 0, st>>>(L.blk.items, sys, tr, target,
            in, out);
        in = out;
    }
}

template void batch_jacobi<double>(const BatchLists&, const BatchSys*, bool, int32_t, int32_t, int32_t, hipStream_t);
template void batch_jacobi<float>(const BatchLists&, const BatchSys*, bool, int32_t, int32_t, int32_t, hipStream_t);


int32_t scan_chunks(int32_t n) { return (n + PS_CHUNK - 1) / PS_CHUNK; }

template<typename VT>
void plan_batch(const PlanSys<VT>* ps, FactorItems blk, FactorItems scan, FactorItems task,
    int32_t hm, hipStream_t st)
{
    if(blk.count <= 0) return;
    const unsigned nb = (unsigned) blk.count;
    const size_t lds = sizeof(double) * (size_t) hm * (size_t) hm;
    bk_plan_rows<VT, false><<<nb, 64, lds, st>>>(ps, blk.items, hm);
    bk_plan_cols<VT, false><<<nb, 64, 0, st>>>(ps, blk.items);
    for(int32_t q = 0; q < 4; ++q) {
        bk_chunk_sums<VT><<<(unsigned) scan.count, PS_TPB, 0, st>>>(ps, scan.items, q);
        bk_chunk_scan<VT><<<(unsigned) scan.count, PS_TPB, 0, st>>>(ps, scan.items, q);
    }
    bk_plan_rows<VT, true><<<nb, 64, lds, st>>>(ps, blk.items, hm);
    bk_plan_cols<VT, true><<<nb, 64, 0, st>>>(ps, blk.items);
    bk_plan_tasks<VT><<<(unsigned) task.count, 256, 0, st>>>(ps, task.items, 0);
    bk_plan_tasks<VT><<<(unsigned) task.count, 256, 0, st>>>(ps, task.items, 1);
}

template void plan_batch<double>(const PlanSys<double>*, FactorItems, FactorItems, FactorItems, int32_t, hipStream_t);
template void plan_batch<float>(const PlanSys<float>*, FactorItems, FactorItems, FactorItems, int32_t, hipStream_t);

}  // namespace culip
