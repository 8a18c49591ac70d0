// BlockiLDLt object: symbolic phase on the host, pattern upload, numeric
// factor and solves on the device (block_ildlt.impl.cuh:44-246, 1284-1390).
#include "objects.hpp"

#include <algorithm>
#include <cstdio>

#include "../device/scale.hpp"
#include <cstring>

namespace culip {

// Schur pairs up to which an off-diagonal block takes the one-lane factor
// kernel. At a level with few off-diagonal blocks the one-lane kernel's time
// is set by its longest list, and a block with more than TINY_PAIRS pairs is
// faster in the wave form, which fetches its pairs 64 at a time. At a level
// with TINY_WIDE_LEVEL or more the one-lane kernel is bound by throughput and
// the wave form by the number of waves the chip holds, so lists up to
// TINY_PAIRS_WIDE stay on the one-lane kernel.
constexpr int32_t TINY_PAIRS = 8;
constexpr int32_t TINY_PAIRS_WIDE = 16;
constexpr int32_t TINY_WIDE_LEVEL = 8192;

void hip_check(hipError_t e, const char* what)
{
    if(e != hipSuccess) fail(CULIP_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void Bildlt::symbolic(int32_t fill_level, double fill_factor)
{
    release_graphs();
    symbolic_analyze(sym, A, piv_starts, block_starts, fill_level, fill_factor, f32());
    have_sym = true;
    pat->uploaded = false;
    uploaded = false;
    factored = false;
    host_int.clear();
    host_real.clear();
}

int64_t BildltPattern::device_bytes() const
{
    int64_t b = 0;
    for(const DevBuf* d : {&d_bs, &d_is_piv, &d_bin, &d_csr_off, &d_csc_off, &d_csc_idx, &d_upd_off,
            &d_upd_ki, &d_upd_kj, &d_lvlb_idx, &d_block_in_lvl, &d_lvl_rows, &d_ix_in,
            &d_value_slot, &d_rfo, &d_lq_cap, &d_lq_fr, &d_off_task, &d_upd4})
        b += (int64_t) d->bytes;
    return b;
}

void BildltPattern::upload_static()
{
    const Symbolic& s = sym;
    upload(d_bs, s.block_starts);
    upload(d_is_piv, s.is_piv);
    upload(d_bin, s.blocks);
    upload(d_csr_off, s.csr_off);
    upload(d_csc_off, s.csc_off);
    upload(d_csc_idx, s.csc_idx);
    upload(d_upd_off, s.upd_off);
    upload(d_upd_ki, s.upd_ki);
    upload(d_upd_kj, s.upd_kj);
    upload(d_lvlb_idx, s.lvlb_idx);
    upload(d_block_in_lvl, s.block_in_lvl);
    upload(d_lvl_rows, s.lvl_rows);
    upload(d_rfo, s.row_first_block_offset);
    upload(d_ix_in, s.ix_store);
    upload(d_value_slot, s.value_slot);

    const int32_t K = s.coarse_nnz;
    // per level-list position: sparse capacity or -1 - rows * cols (dense),
    // and the first fine row of the block row
    {
        std::vector<int32_t> cap(K), fr(K);
        for(int32_t p = 0; p < K; ++p) {
            const BlockMeta& m = s.blocks[s.lvlb_idx[p]];
            const int32_t h = s.block_starts[m.block_row + 1] - s.block_starts[m.block_row];
            const int32_t w = s.block_starts[m.block_col + 1] - s.block_starts[m.block_col];
            cap[p] = (m.format == FMT_SPARSE) ? m.max_nnz : -1 - h * w;
            fr[p] = s.block_starts[m.block_row];
        }
        upload(d_lq_cap, cap);
        upload(d_lq_fr, fr);
    }
    // off-diagonal task descriptors (12 int32 per level-list position) and
    // Schur pairs with the rows of their block k
    {
        std::vector<int32_t> diag_id(s.nb, -1);
        for(int32_t b = 0; b < K; ++b)
            if(s.blocks[b].block_row == s.blocks[b].block_col) diag_id[s.blocks[b].block_row] = b;
        std::vector<int32_t> task((size_t) 12 * K, 0);
        for(int32_t l = 0; l < s.num_levels; ++l) {
            const int32_t lo = s.lvlb_off[l], hi = s.lvlb_off[l + 1];
            const int32_t npiv = s.lvl_off[l + 1] - s.lvl_off[l];
            const int32_t lim = (hi - lo - npiv >= TINY_WIDE_LEVEL) ? TINY_PAIRS_WIDE : TINY_PAIRS;
            for(int32_t p = lo + npiv; p < hi; ++p) {
                const int32_t b = s.lvlb_idx[p];
                const BlockMeta& m = s.blocks[b];
                const int32_t i = m.block_row, jb = m.block_col;
                const int32_t fr = s.block_starts[i];
                const int32_t h = s.block_starts[i + 1] - fr;
                const int32_t w = s.block_starts[jb + 1] - s.block_starts[jb];
                bool all1 = true;
                for(int32_t r = fr; r < fr + h; ++r) all1 = all1 && s.is_piv[r] != 0;
                CULIP_REQUIRE(diag_id[i] >= 0, "factor: block row without a pivot block");
                const int32_t pq = s.block_in_lvl[diag_id[i]] - lo;
                CULIP_REQUIRE(pq >= 0 && pq < npiv, "factor: pivot block outside its level");
                int32_t* t = &task[(size_t) 12 * p];
                t[0] = b; t[1] = fr; t[2] = h | (w << 8) | ((all1 ? 1 : 0) << 16); t[3] = pq;
                t[4] = m.format; t[5] = m.nnz; t[6] = m.col_ptr; t[7] = m.val_ptr;
                // first pair as two 32-bit halves (t[8] low, t[11] high)
                const int64_t first = s.upd_off[b], cnt = s.upd_off[b + 1] - first;
                CULIP_REQUIRE(cnt < (int64_t) INT32_MAX, "factor: too many Schur pairs in one block");
                t[8] = (int32_t) (uint32_t) ((uint64_t) first & 0xffffffffu);
                t[9] = (int32_t) cnt;
                t[11] = (int32_t) (uint32_t) ((uint64_t) first >> 32);
                // one-lane kernel: sparse input of at most 8 entries, a
                // short Schur pair list, 1x1 pivots in the block row
                t[10] = (m.format == FMT_SPARSE && m.nnz <= 8 && t[9] <= lim && all1) ? 1 : 0;
            }
        }
        upload(d_off_task, task);
        const size_t npairs = s.upd_ki.size();
        std::vector<int32_t> u4(4 * npairs);
        for(size_t u = 0; u < npairs; ++u) {
            const int32_t k = s.blocks[s.upd_ki[u]].block_row;
            u4[4 * u + 0] = s.upd_ki[u];
            u4[4 * u + 1] = s.upd_kj[u];
            u4[4 * u + 2] = s.block_starts[k];
            u4[4 * u + 3] = s.block_starts[k + 1] - s.block_starts[k];
        }
        upload(d_upd4, u4);
    }
    uploaded = true;
}

void Bildlt::upload_pattern(hipStream_t st)
{
    release_graphs();
    if(!pat->uploaded) pat->upload_static();
    const Symbolic& s = sym;
    const int32_t n = s.n, K = s.coarse_nnz;
    const int64_t mlb = std::max<int64_t>(1, s.max_level_blocks);
    d_keep.ensure(sizeof(int32_t) * mlb);
    d_granted.ensure(sizeof(int32_t) * mlb);
    // 32 row masks per pivot position, then one "any" word per position
    d_piv_umask.ensure(sizeof(uint32_t) * 33 * mlb);
    d_defer_list.ensure(sizeof(int32_t) * mlb);
    d_defer_cnt.ensure(sizeof(int32_t) * std::max(1, s.num_levels));
    d_wave_list.ensure(sizeof(int32_t) * mlb);
    d_wave_cnt.ensure(sizeof(int32_t) * std::max(1, s.num_levels));
    d_store_list.ensure(sizeof(int32_t) * mlb);
    d_store_cnt.ensure(sizeof(int32_t) * std::max(1, s.num_levels));
    // per scan chunk 4 int64, then per level 2 int32 flags per chunk and
    // a chunk ticket counter
    {
        const size_t maxch = (size_t) ((mlb + SCAN_CHUNK - 1) / SCAN_CHUNK);
        d_chunk.ensure(sizeof(int64_t) * 4 * maxch +
            sizeof(int32_t) * (2 * maxch + 1) * (size_t) std::max(1, s.num_levels));
    }
    d_vals.ensure(sizeof(double) * std::max<int64_t>(1, A.nnz()));
    d_val_in.ensure(vt() * std::max<int64_t>(1, s.val_space));
    d_bout.ensure(sizeof(BlockMeta) * std::max(1, K));
    d_ix_out.ensure(sizeof(int32_t) * std::max<int64_t>(1, s.ix_space));
    d_val_out.ensure(vt() * std::max<int64_t>(1, s.val_space));
    d_tri.ensure(sizeof(double) * 3 * (size_t) n);
    d_rn.ensure(vt() * n);
    d_if_val.ensure(vt() * 1024 * mlb);
    d_if_idx.ensure(sizeof(uint16_t) * 1024 * mlb);
    d_if_cnt.ensure(sizeof(int32_t) * mlb);
    d_rn_part.ensure(vt() * 32 * mlb);
    d_drop_cnt.ensure(sizeof(int32_t) * mlb);
    d_overflow.ensure(mlb);
    d_piv_loc.ensure(sizeof(int32_t) * s.nb);
    d_carry.ensure(sizeof(int32_t) * 8);
    for(int q = 0; q < 4; ++q) d_ptr[q].ensure(sizeof(int32_t) * (size_t) (n + 1));
    for(int q = 0; q < 2; ++q) d_task[q].ensure(sizeof(int32_t) * 8 * (size_t) s.nb);
    d_totals.ensure(sizeof(int32_t) * 4);
    // solve layout at its upper bounds, so that a refactor needs no host
    // round trip to size it: the off-diagonal rows hold at most every
    // stored entry (leftover borrowing moves capacity, never adds to it),
    // the pivot-inverse rows at most h^2 entries per block row; 64 entries
    // of slack for the solve kernels' clamped reads
    {
        int64_t inv = 0;
        for(int32_t i = 0; i < s.nb; ++i) {
            const int64_t h = s.block_starts[i + 1] - s.block_starts[i];
            inv += h * h;
        }
        const int64_t cap[4] = {s.val_space, inv, s.val_space, inv};
        for(int q = 0; q < 4; ++q) {
            const size_t cnt = (size_t) std::max<int64_t>(1, cap[q]) + 64;
            d_col[q].ensure(sizeof(int32_t) * cnt);
            d_val[q].ensure(vt() * cnt);
            d_lr[q].ensure(cnt);
        }
    }
    d_scan_tmp.ensure(sizeof(int32_t) * scan_tmp_ints(n));
    d_fis_piv.ensure(sizeof(int32_t) * (size_t) std::max(1, n));
    d_perm.ensure(sizeof(int32_t) * (size_t) std::max(1, n));
    d_perm_tmp.ensure(sizeof(double) * (size_t) std::max(1, n));
    d_chain_pub.ensure(sizeof(double) * 2 * (size_t) std::max(1, n));
    h_trsv_err.ensure();
    // the presets on the caller's stream: no device-wide synchronisation,
    // which would break graph captures other threads run meanwhile
    chain_reset(d_chain_pub.as<double>(), 2 * (int64_t) n, st);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
    chain_parity = 0;
    (void) h_trsv_err.take();
    uploaded = true;
}

template<typename VT>
FactorArgs<VT> Bildlt::factor_args() const
{
    FactorArgs<VT> a;
    a.n = sym.n;
    a.nb = sym.nb;
    a.bs = d_bs.as<int32_t>();
    a.is_piv = d_fis_piv.as<int32_t>();
    a.perm = d_perm.as<int32_t>();
    a.pivoting = (int32_t) pivoting;
    a.piv64 = (f32() && pivot_precision == CULIP_PIVOT2_F64) ? 1 : 0;
    a.bin = d_bin.as<BlockMeta>();
    a.bout = d_bout.as<BlockMeta>();
    a.csr_off = d_csr_off.as<int32_t>();
    a.csc_off = d_csc_off.as<int32_t>();
    a.csc_idx = d_csc_idx.as<int32_t>();
    a.upd_off = d_upd_off.as<int64_t>();
    a.upd_ki = d_upd_ki.as<int32_t>();
    a.upd_kj = d_upd_kj.as<int32_t>();
    a.lvlb_idx = d_lvlb_idx.as<int32_t>();
    a.block_in_lvl = d_block_in_lvl.as<int32_t>();
    a.lvl_rows = d_lvl_rows.as<int32_t>();
    a.rfo = d_rfo.as<int32_t>();
    a.ix_in = d_ix_in.as<int32_t>();
    a.val_in = d_val_in.as<VT>();
    a.ix_out = d_ix_out.as<int32_t>();
    a.val_out = d_val_out.as<VT>();
    a.tri = d_tri.as<double>();
    a.rn = d_rn.as<VT>();
    a.if_val = d_if_val.as<VT>();
    a.if_idx = d_if_idx.as<uint16_t>();
    a.if_cnt = d_if_cnt.as<int32_t>();
    a.rn_part = d_rn_part.as<VT>();
    a.drop_cnt = d_drop_cnt.as<int32_t>();
    a.keep = d_keep.as<int32_t>();
    a.granted = d_granted.as<int32_t>();
    a.overflow = d_overflow.as<uint8_t>();
    a.piv_loc = d_piv_loc.as<int32_t>();
    a.lq_cap = d_lq_cap.as<int32_t>();
    a.lq_fr = d_lq_fr.as<int32_t>();
    a.chunk = d_chunk.as<int64_t>();
    {
        const int64_t mlb = std::max<int64_t>(1, sym.max_level_blocks);
        a.scan_maxch = (int32_t) ((mlb + SCAN_CHUNK - 1) / SCAN_CHUNK);
        a.scan_flags = (int32_t*) (a.chunk + 4 * (int64_t) a.scan_maxch);
    }
    a.err = h_trsv_err.p;
    a.off_task = d_off_task.as<int32_t>();
    a.upd4 = d_upd4.as<int32_t>();
    a.piv_umask = d_piv_umask.as<uint32_t>();
    a.piv_any = a.piv_umask + 32 * std::max<int64_t>(1, sym.max_level_blocks);
    a.defer_list = d_defer_list.as<int32_t>();
    a.defer_cnt = d_defer_cnt.as<int32_t>();
    a.wave_list = d_wave_list.as<int32_t>();
    a.wave_cnt = d_wave_cnt.as<int32_t>();
    a.store_list = d_store_list.as<int32_t>();
    a.store_cnt = d_store_cnt.as<int32_t>();
    a.num_levels = sym.num_levels;
    a.carry = d_carry.as<int32_t>();
    a.threshold = (VT) threshold;
    return a;
}

template<typename VT>
SolvePlan<VT> Bildlt::plan() const
{
    SolvePlan<VT> p;
    int32_t** ptrs[4] = {&p.uo_ptr, &p.ui_ptr, &p.to_ptr, &p.ti_ptr};
    int32_t** cols[4] = {&p.uo_col, &p.ui_col, &p.to_col, &p.ti_col};
    VT** vals[4] = {&p.uo_val, &p.ui_val, &p.to_val, &p.ti_val};
    uint8_t** lrs[4] = {&p.uo_lr, &p.ui_lr, &p.to_lr, &p.ti_lr};
    for(int q = 0; q < 4; ++q) {
        *ptrs[q] = d_ptr[q].as<int32_t>();
        *cols[q] = d_col[q].as<int32_t>();
        *vals[q] = d_val[q].as<VT>();
        *lrs[q] = d_lr[q].as<uint8_t>();
    }
    p.task_u = d_task[0].as<int4>();
    p.task_t = d_task[1].as<int4>();
    return p;
}

template SolvePlan<float> Bildlt::plan<float>() const;
template FactorArgs<float> Bildlt::factor_args<float>() const;
template FactorArgs<double> Bildlt::factor_args<double>() const;
template SolvePlan<double> Bildlt::plan<double>() const;

SolveTopo Bildlt::topo() const
{
    SolveTopo t;
    t.n = sym.n;
    t.nb = sym.nb;
    t.num_levels = sym.num_levels;
    t.bs = d_bs.as<int32_t>();
    t.lvl_rows = d_lvl_rows.as<int32_t>();
    t.lvl_off_h = sym.lvl_off.data();
    t.csr_off = d_csr_off.as<int32_t>();
    t.csc_off = d_csc_off.as<int32_t>();
    t.csc_idx = d_csc_idx.as<int32_t>();
    t.bout = d_bout.as<BlockMeta>();
    t.ix_out = d_ix_out.as<int32_t>();
    t.tri = d_tri.as<double>();
    t.is_piv = d_fis_piv.as<int32_t>();
    t.chain_pub = d_chain_pub.as<double>();
    t.err = h_trsv_err.p;
    return t;
}

template<typename VT>
void Bildlt::factor_impl(hipStream_t st)
{
    const Symbolic& s = sym;
    FactorArgs<VT> a = factor_args<VT>();
    factor_reset_pivots(s.n, d_is_piv.as<int32_t>(), a.is_piv, a.perm, st);
    gather_values<VT>(d_vals.as<double>(), d_value_slot.as<int64_t>(), A.nnz(), d_val_in.as<VT>(),
        s.val_space, st);
    factor_begin<VT>(a, st);
    for(int32_t l = 0; l < s.num_levels; ++l) {
        LevelRange r;
        r.lo = s.lvlb_off[l];
        r.size = s.lvlb_off[l + 1] - r.lo;
        r.npiv = s.lvl_off[l + 1] - s.lvl_off[l];
        r.row_lo = s.lvl_off[l];
        r.level = l;
        factor_level<VT>(a, r, st);
    }
    HIP_CHECK(hipGetLastError());

    // solve layout: count, scan, size, fill
    SolveTopo t = topo();
    SolvePlan<VT> p = plan<VT>();
    solve_plan_count<VT>(t, a.val_out, nullptr, p, st);
    int32_t* tot = d_totals.as<int32_t>();
    for(int q = 0; q < 4; ++q)
        scan_counts(d_ptr[q].as<int32_t>(), s.n, tot + q, d_scan_tmp.as<int32_t>(), st);
    // the layout buffers hold the upper bounds (upload_pattern)
    solve_plan_fill<VT>(t, a.val_out, nullptr, p, st);
    HIP_CHECK(hipGetLastError());
}

void Bildlt::release_graphs()
{
    if(factor_exec) (void) hipGraphExecDestroy(factor_exec);
    factor_exec = nullptr;
}

bool Bildlt::factor_graph(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    // graphs are captured on (and replayed from) explicit streams only; a
    // stream some caller is capturing gets the kernels themselves
    if(!use_graphs || st == nullptr || hipStreamIsCapturing(st, &cs) != hipSuccess ||
        cs != hipStreamCaptureStatusNone)
        return false;
    if(factor_exec && graph_threshold != threshold) release_graphs();
    if(!factor_exec) {
        HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            if(f32()) factor_impl<float>(st);
            else factor_impl<double>(st);
        } catch(...) {
            hipGraph_t g = nullptr;
            (void) hipStreamEndCapture(st, &g);
            if(g) (void) hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        HIP_CHECK(hipStreamEndCapture(st, &g));
        const hipError_t e = hipGraphInstantiate(&factor_exec, g, nullptr, nullptr, 0);
        (void) hipGraphDestroy(g);
        hip_check(e, "factor graph instantiation");
        graph_threshold = threshold;
    }
    HIP_CHECK(hipGraphLaunch(factor_exec, st));
    return true;
}

void Bildlt::factor(const double* vals, bool on_device, double thr, hipStream_t st)
{
    if(!have_sym) fail(CULIP_ERR_STATE, "factor: symbolic phase has not run");
    take_async_error();
    if(!uploaded) upload_pattern(st);
    threshold = thr;
    const size_t bytes = sizeof(double) * (size_t) A.nnz();
    if(bytes) {
        if(vals == nullptr) vals = A.val.data(), on_device = false;
        // values already in the factor's own buffer need no copy
        if(vals != d_vals.p)
            HIP_CHECK(hipMemcpyAsync(d_vals.p, vals, bytes,
                on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    }
    if(!factor_graph(st)) {
        if(f32()) factor_impl<float>(st);
        else factor_impl<double>(st);
    }
    HIP_CHECK(hipStreamSynchronize(st));
    after_factor();
}

void Bildlt::after_factor()
{
    // a level's dropping scan whose bounded wait gave up leaves the object
    // unfactored
    factored = false;
    take_async_error();
    factored = true;
    plan_known = false;
    host_int.erase("fblocks");
    host_int.erase("perm");
    host_int.erase("fis_piv");
    host_int.erase("fix_store");
    host_real.erase("fval_store");
    host_real.erase("tri");
    host_real.erase("row_norms");
}

const int32_t* Bildlt::plan_sizes()
{
    if(!plan_known) {
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipMemcpy(plan_tot, d_totals.p, sizeof(plan_tot), hipMemcpyDeviceToHost));
        plan_known = true;
    }
    return plan_tot;
}

void Bildlt::set_solve(culip_solve_algorithm a, int32_t sweeps)
{
    if(a != CULIP_TRSV_LEVELS && a != CULIP_TRSV_SINGLE && a != CULIP_TRSV_JACOBI)
        fail(CULIP_ERR_INVALID, "bildlt: unknown solve algorithm");
    if(a == CULIP_TRSV_JACOBI && sweeps < 1)
        fail(CULIP_ERR_INVALID, "bildlt: block-Jacobi solves need at least one sweep");
    solve_algo = a;
    jacobi_sweeps = sweeps;
    if(a == CULIP_TRSV_JACOBI)
        for(int q = 0; q < 3; ++q) d_jac[q].ensure(sizeof(double) * (size_t) std::max(1, n()));
}

void Bildlt::jacobi(double* x, bool transpose, hipStream_t st) const
{
    const size_t bytes = sizeof(double) * (size_t) std::max(1, n());
    DevBuf* jb = const_cast<DevBuf*>(d_jac);
    for(int q = 0; q < 3; ++q) jb[q].ensure(bytes);
    HIP_CHECK(hipMemcpyAsync(jb[0].p, x, sizeof(double) * (size_t) n(), hipMemcpyDeviceToDevice, st));
    if(f32())
        ::culip::jacobi_sweeps<float>(topo(), plan<float>(), transpose, x, jacobi_sweeps,
            jb[0].as<double>(), jb[1].as<double>(), jb[2].as<double>(), st);
    else
        ::culip::jacobi_sweeps<double>(topo(), plan<double>(), transpose, x, jacobi_sweeps,
            jb[0].as<double>(), jb[1].as<double>(), jb[2].as<double>(), st);
}

void Bildlt::perm_pull(double* x, hipStream_t st) const
{
    if(pivoting == CULIP_PIVOT_STATIC) return;
    const size_t bytes = sizeof(double) * (size_t) n();
    HIP_CHECK(hipMemcpyAsync(d_perm_tmp.p, x, bytes, hipMemcpyDeviceToDevice, st));
    permute_pull(n(), d_perm.as<int32_t>(), d_perm_tmp.as<double>(), x, st);
}

void Bildlt::perm_push(double* x, hipStream_t st) const
{
    if(pivoting == CULIP_PIVOT_STATIC) return;
    const size_t bytes = sizeof(double) * (size_t) n();
    HIP_CHECK(hipMemcpyAsync(d_perm_tmp.p, x, bytes, hipMemcpyDeviceToDevice, st));
    permute_push(n(), d_perm.as<int32_t>(), d_perm_tmp.as<double>(), x, st);
}

void Bildlt::trsv_u(double* x, hipStream_t st) const
{
    trsv_u_unpermuted(x, st);
    perm_push(x, st);
}

void Bildlt::trsv_ut(double* x, hipStream_t st) const
{
    perm_pull(x, st);
    trsv_ut_unpermuted(x, st);
}

void Bildlt::trsv_u_unpermuted(double* x, hipStream_t st) const
{
    if(solve_algo == CULIP_TRSV_JACOBI) {
        jacobi(x, false, st);
        return;
    }
    if(solve_algo == CULIP_TRSV_SINGLE && !solve_levels &&
        chain_admitted(st, [&] {
            const int32_t par = chain_parity;
            chain_parity ^= 1;
            return (hipError_t) (f32() ? trsv_upper_single<float>(topo(), plan<float>(), x, par, st)
                                       : trsv_upper_single<double>(topo(), plan<double>(), x, par, st));
        }))
        return;
    // level launches (also when a single launch was not admitted: same bits)
    if(f32()) trsv_upper<float>(topo(), plan<float>(), x, st);
    else trsv_upper<double>(topo(), plan<double>(), x, st);
}

void Bildlt::trsv_ut_unpermuted(double* x, hipStream_t st) const
{
    if(solve_algo == CULIP_TRSV_JACOBI) {
        jacobi(x, true, st);
        return;
    }
    if(solve_algo == CULIP_TRSV_SINGLE && !solve_levels &&
        chain_admitted(st, [&] {
            const int32_t par = chain_parity;
            chain_parity ^= 1;
            return (hipError_t) (f32() ? trsv_upper_t_single<float>(topo(), plan<float>(), x, par, st)
                                       : trsv_upper_t_single<double>(topo(), plan<double>(), x, par, st));
        }))
        return;
    if(f32()) trsv_upper_t<float>(topo(), plan<float>(), x, st);
    else trsv_upper_t<double>(topo(), plan<double>(), x, st);
}

bool Bildlt::left_product(const DevCsr& K, const double* src, double* x, hipStream_t st) const
{
    // the fused forms read K's rows in the factor's row order: not with P
    if(K.m != n() || K.n != n() || pivoting != CULIP_PIVOT_STATIC) return false;
    if(solve_algo == CULIP_TRSV_JACOBI) {
        // the first U' sweep forms each block row's K src itself
        const size_t bytes = sizeof(double) * (size_t) std::max(1, n());
        DevBuf* jb = const_cast<DevBuf*>(d_jac);
        for(int q = 0; q < 3; ++q) jb[q].ensure(bytes);
        if(f32())
            jacobi_sweeps_product<float>(topo(), plan<float>(), true, &K, src, x, jacobi_sweeps,
                jb[0].as<double>(), jb[1].as<double>(), jb[2].as<double>(), st);
        else
            jacobi_sweeps_product<double>(topo(), plan<double>(), true, &K, src, x, jacobi_sweeps,
                jb[0].as<double>(), jb[1].as<double>(), jb[2].as<double>(), st);
        HIP_CHECK(hipGetLastError());
        return true;
    }
    // the single-launch U' solve forms each block row's K src itself; when
    // it is not admitted the caller multiplies and calls left() (same bits)
    if(solve_algo != CULIP_TRSV_SINGLE || solve_levels) return false;
    return chain_admitted(st, [&] {
        const int32_t par = chain_parity;
        chain_parity ^= 1;
        return (hipError_t) (f32()
                ? trsv_upper_t_single_product<float>(topo(), plan<float>(), &K, src, x, par, st)
                : trsv_upper_t_single_product<double>(topo(), plan<double>(), &K, src, x, par, st));
    });
}

void Bildlt::check_solve_error(hipStream_t st) const
{
    if(!uploaded) return;
    HIP_CHECK(hipStreamSynchronize(st));
    take_async_error();
}

void Bildlt::take_async_error() const
{
    const int32_t e = h_trsv_err.take();
    if(e == 2)
        fail(CULIP_ERR_HIP, "factor: a chunk wait of a level's dropping scan timed out");
    if(e)
        fail(CULIP_ERR_HIP, "triangular solve: a dependency wait of a single-launch solve timed "
                            "out (its x holds NaN)");
}

void Bildlt::solve(culip_side side, bool transpose, const double* b, double* x,
    hipStream_t st) const
{
    if(!factored) fail(CULIP_ERR_STATE, "solve: preconditioner has not been computed");
    take_async_error();
    const size_t bytes = sizeof(double) * (size_t) n();
    if(side == CULIP_SOLVE_MIDDLE) {
        solve_D(n(), d_tri.as<double>(), d_fis_piv.as<int32_t>(), b, x, st);
        return;
    }
    // transpose swaps left and right (block_ildlt.impl.cuh:192-201, 236-245)
    const bool right = (side == CULIP_SOLVE_RIGHT) != transpose;
    if(b != x) HIP_CHECK(hipMemcpyAsync(x, b, bytes, hipMemcpyDeviceToDevice, st));
    // P is the identity for static pivoting (k_permute_P / k_permute_Pt)
    if(right) trsv_u(x, st);
    else trsv_ut(x, st);
}

void Bildlt::solve_noalloc(culip_side side, bool transpose, const double* b, double* x,
    hipStream_t st) const
{
    if(n() > NOALLOC_MAX_ROWS)
        fail(CULIP_ERR_INVALID, "solve_noalloc: " + std::to_string(n()) + " rows exceed the "
            "static scratch of " + std::to_string(NOALLOC_MAX_ROWS) + " rows");
    if(!factored) fail(CULIP_ERR_STATE, "solve: preconditioner has not been computed");
    // the scratch array's address on this device (looked up on the first
    // call per device)
    double* s0 = noalloc_scratch();
    if(!s0) fail(CULIP_ERR_HIP, "solve_noalloc: the scratch array's device address is unknown");
    take_async_error();
    if(side == CULIP_SOLVE_MIDDLE) {
        solve_D(n(), d_tri.as<double>(), d_fis_piv.as<int32_t>(), b, x, st);
        return;
    }
    const bool right = (side == CULIP_SOLVE_RIGHT) != transpose;
    const size_t bytes = sizeof(double) * (size_t) n();
    if(b != x) HIP_CHECK(hipMemcpyAsync(x, b, bytes, hipMemcpyDeviceToDevice, st));
    const bool piv = pivoting != CULIP_PIVOT_STATIC;
    // P' before a left solve (the scratch holds the copy it reads)
    if(piv && !right) {
        HIP_CHECK(hipMemcpyAsync(s0, x, bytes, hipMemcpyDeviceToDevice, st));
        permute_pull(n(), d_perm.as<int32_t>(), s0, x, st);
    }
    if(solve_algo != CULIP_TRSV_JACOBI) {
        // the exact solve as level launches: no published-value buffer
        if(right) {
            if(f32()) trsv_upper<float>(topo(), plan<float>(), x, st);
            else trsv_upper<double>(topo(), plan<double>(), x, st);
        } else {
            if(f32()) trsv_upper_t<float>(topo(), plan<float>(), x, st);
            else trsv_upper_t<double>(topo(), plan<double>(), x, st);
        }
    } else {
        double* s1 = s0 + NOALLOC_MAX_ROWS;
        double* s2 = s1 + NOALLOC_MAX_ROWS;
        HIP_CHECK(hipMemcpyAsync(s0, x, bytes, hipMemcpyDeviceToDevice, st));
        if(f32())
            ::culip::jacobi_sweeps<float>(topo(), plan<float>(), !right, x, jacobi_sweeps, s0, s1,
                s2, st);
        else
            ::culip::jacobi_sweeps<double>(topo(), plan<double>(), !right, x, jacobi_sweeps, s0, s1,
                s2, st);
    }
    // P after a right solve
    if(piv && right) {
        HIP_CHECK(hipMemcpyAsync(s0, x, bytes, hipMemcpyDeviceToDevice, st));
        permute_push(n(), d_perm.as<int32_t>(), s0, x, st);
    }
}

int64_t Bildlt::numeric_bytes() const
{
    int64_t b = 0;
    for(const DevBuf* d : {&d_vals, &d_val_in, &d_bout, &d_ix_out, &d_val_out, &d_tri, &d_rn,
            &d_if_val, &d_if_idx, &d_if_cnt, &d_rn_part, &d_drop_cnt, &d_overflow, &d_piv_loc,
            &d_carry, &d_totals, &d_task[0], &d_task[1], &d_keep, &d_granted, &d_chunk,
            &d_scan_tmp, &d_piv_umask, &d_defer_list, &d_defer_cnt, &d_wave_list, &d_wave_cnt,
            &d_store_list, &d_store_cnt,
            &d_chain_pub, &d_jac[0], &d_jac[1], &d_jac[2]})
        b += (int64_t) d->bytes;
    for(int q = 0; q < 4; ++q)
        b += (int64_t) (d_ptr[q].bytes + d_col[q].bytes + d_val[q].bytes + d_lr[q].bytes);
    return b;
}

int64_t Bildlt::device_bytes() const { return pat->device_bytes() + numeric_bytes(); }

int64_t Bildlt::numeric_bytes_estimate(bool with_layout) const
{
    const Symbolic& s = sym;
    const int64_t n = s.n, nb = s.nb, K = s.coarse_nnz, v = (int64_t) vt();
    const int64_t mlb = std::max<int64_t>(1, s.max_level_blocks);
    auto r16 = [](int64_t b) { return std::max<int64_t>(16, b); };
    int64_t b = 0;
    b += r16(4 * mlb) * 2 + r16(128 * mlb) + r16(4 * mlb) + r16(4 * std::max(1, s.num_levels));
    b += r16(32 * ((mlb + SCAN_CHUNK - 1) / SCAN_CHUNK));
    b += r16(8 * (int64_t) s.value_slot.size()) + 2 * r16(v * s.val_space) + r16(36 * K);
    b += r16(4 * s.ix_space) + 24 * n + v * n;
    b += v * 1024 * mlb + 2048 * mlb + 4 * mlb + v * 32 * mlb + 4 * mlb + mlb;
    b += 4 * nb + 32 + 4 * 4 * (n + 1) + 2 * 32 * nb + 16 + 4 * ((int64_t) scan_tmp_ints(s.n));
    b += 16 * std::max<int64_t>(1, n) + 4;   // single-launch solves' published values, error
    b += 2 * r16(4 * mlb) + 2 * r16(4 * std::max(1, s.num_levels));   // wave / store lists
    if(with_layout) {
        // U and U' rows: each stored entry once per layout, plus the pivot
        // block inverses (at most rows^2 per block row, twice)
        int64_t inv = 0;
        for(int32_t i = 0; i < s.nb; ++i) {
            const int64_t h = s.block_starts[i + 1] - s.block_starts[i];
            inv += h * h;
        }
        b += (2 * s.val_space + 2 * inv + 4 * 64) * (4 + v + 1);
    }
    if(solve_algo == CULIP_TRSV_JACOBI) b += 3 * 8 * n;   // block-Jacobi vectors
    return b;
}

void Bildlt::download_factor()
{
    if(!factored) fail(CULIP_ERR_STATE, "factor has not been computed");
    if(host_int.count("fblocks")) return;
    HIP_CHECK(hipDeviceSynchronize());
    // final {pool, out_ix, out_val} set after the last level
    int32_t carry[4];
    HIP_CHECK(hipMemcpy(carry, d_carry.as<int32_t>() + 4 * (sym.num_levels & 1), sizeof(carry),
        hipMemcpyDeviceToHost));
    const int32_t K = sym.coarse_nnz;
    std::vector<int32_t> fb((size_t) K * 9);
    HIP_CHECK(hipMemcpy(fb.data(), d_bout.p, sizeof(BlockMeta) * K, hipMemcpyDeviceToHost));
    host_int["fblocks"] = std::move(fb);
    std::vector<int32_t> fix(carry[1]);
    if(carry[1]) HIP_CHECK(hipMemcpy(fix.data(), d_ix_out.p, sizeof(int32_t) * carry[1], hipMemcpyDeviceToHost));
    host_int["fix_store"] = std::move(fix);
    host_int["carry"] = std::vector<int32_t>(carry, carry + 4);
    std::vector<double> fv(carry[2]);
    std::vector<double> rn(n());
    if(f32()) {
        std::vector<float> t(std::max(carry[2], n()));
        if(carry[2]) HIP_CHECK(hipMemcpy(t.data(), d_val_out.p, sizeof(float) * carry[2], hipMemcpyDeviceToHost));
        for(int32_t q = 0; q < carry[2]; ++q) fv[q] = t[q];
        HIP_CHECK(hipMemcpy(t.data(), d_rn.p, sizeof(float) * n(), hipMemcpyDeviceToHost));
        for(int32_t q = 0; q < n(); ++q) rn[q] = t[q];
    } else {
        if(carry[2]) HIP_CHECK(hipMemcpy(fv.data(), d_val_out.p, sizeof(double) * carry[2], hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(rn.data(), d_rn.p, sizeof(double) * n(), hipMemcpyDeviceToHost));
    }
    host_real["fval_store"] = std::move(fv);
    host_real["row_norms"] = std::move(rn);
    std::vector<double> tri(3 * (size_t) n());
    HIP_CHECK(hipMemcpy(tri.data(), d_tri.p, sizeof(double) * tri.size(), hipMemcpyDeviceToHost));
    host_real["tri"] = std::move(tri);
    std::vector<int32_t> perm(n()), fpiv(n());
    if(n()) {
        HIP_CHECK(hipMemcpy(perm.data(), d_perm.p, sizeof(int32_t) * n(), hipMemcpyDeviceToHost));
        HIP_CHECK(hipMemcpy(fpiv.data(), d_fis_piv.p, sizeof(int32_t) * n(), hipMemcpyDeviceToHost));
    }
    host_int["perm"] = std::move(perm);
    host_int["fis_piv"] = std::move(fpiv);
}

// export_factorized (block_ildlt.impl.cuh:1700-1916): U as fine CSR (the
// nonzero entries of every factored block, columns sorted per row), the
// tridiagonal D and the permutation (identity for static pivots), each
// written as a MatrixMarket coordinate file (test.impl.h:480-545).
static void write_mm(const std::string& path, int32_t m, const std::vector<int32_t>& rp,
    const std::vector<int32_t>& ci, const std::vector<double>& cv)
{
    FILE* f = std::fopen(path.c_str(), "w");
    if(!f) fail(CULIP_ERR_INVALID, "export: cannot write " + path);
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%d %d %d\n", m, m, rp[m]);
    for(int32_t i = 0; i < m; ++i)
        for(int32_t q = rp[i]; q < rp[i + 1]; ++q)
            std::fprintf(f, "%d %d %.12g\n", i + 1, ci[q] + 1, cv[q]);
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    if(!ok) fail(CULIP_ERR_INVALID, "export: write error on " + path);
}

double Bildlt::export_factorized(const std::string& folder)
{
    download_factor();
    const int32_t N = n();
    const std::vector<int32_t>& fb = host_int["fblocks"];
    const std::vector<int32_t>& fix = host_int["fix_store"];
    const std::vector<double>& fv = host_real["fval_store"];
    const std::vector<double>& tri = host_real["tri"];
    const std::vector<int32_t>& bs = sym.block_starts;   // nb + 1, with the sentinel N
    const int32_t K = (int32_t) (fb.size() / 9);
    // U: count, offsets, fill, sort
    std::vector<int32_t> rp(N + 1, 0);
    auto each = [&](auto&& visit) {
        for(int32_t b = 0; b < K; ++b) {
            const BlockMeta& m = *reinterpret_cast<const BlockMeta*>(&fb[9 * (size_t) b]);
            const int32_t fi = bs[m.block_row], fj = bs[m.block_col];
            const int32_t h = bs[m.block_row + 1] - fi, w = bs[m.block_col + 1] - fj;
            if(m.format == FMT_DENSE) {
                for(int32_t r = 0; r < h; ++r)
                    for(int32_t c = 0; c < w; ++c) {
                        const double v = fv[(size_t) m.val_ptr + r * w + c];
                        if(v != 0.0) visit(fi + r, fj + c, v);
                    }
            } else {
                for(int32_t e = 0; e < m.nnz; ++e) {
                    const int32_t ix = fix[(size_t) m.col_ptr + e];
                    const double v = fv[(size_t) m.val_ptr + e];
                    if(v != 0.0) visit(fi + ix / w, fj + ix % w, v);
                }
            }
        }
    };
    each([&](int32_t r, int32_t, double) { ++rp[r + 1]; });
    for(int32_t i = 0; i < N; ++i) rp[i + 1] += rp[i];
    std::vector<int32_t> ci(rp[N]), next(rp.begin(), rp.end() - 1);
    std::vector<double> cv(rp[N]);
    each([&](int32_t r, int32_t c, double v) {
        const int32_t q = next[r]++;
        ci[q] = c;
        cv[q] = v;
    });
    std::vector<std::pair<int32_t, double>> row;
    for(int32_t i = 0; i < N; ++i) {
        row.clear();
        for(int32_t q = rp[i]; q < rp[i + 1]; ++q) row.emplace_back(ci[q], cv[q]);
        std::sort(row.begin(), row.end(),
            [](const std::pair<int32_t, double>& x, const std::pair<int32_t, double>& y) {
                return x.first < y.first;
            });
        for(size_t k = 0; k < row.size(); ++k) {
            ci[rp[i] + k] = row[k].first;
            cv[rp[i] + k] = row[k].second;
        }
    }
    write_mm(folder + "/out_U.mtx", N, rp, ci, cv);
    const int64_t nnz_u = rp[N];
    // D: per row its sub-diagonal neighbour above, diagonal, neighbour below,
    // the off-diagonal ones only when nonzero; rows in order, columns sorted
    std::vector<int32_t> drp(N + 1, 0), dci;
    std::vector<double> dcv;
    for(int32_t i = 0; i < N; ++i) {
        // row i holds D(i, i-1) = tri[3 (i-1) + 2], D(i, i), D(i, i+1) = tri[3 (i+1)]
        if(i > 0 && tri[3 * (size_t) (i - 1) + 2] != 0.0) {
            dci.push_back(i - 1);
            dcv.push_back(tri[3 * (size_t) (i - 1) + 2]);
        }
        dci.push_back(i);
        dcv.push_back(tri[3 * (size_t) i + 1]);
        if(i < N - 1 && tri[3 * (size_t) (i + 1)] != 0.0) {
            dci.push_back(i + 1);
            dcv.push_back(tri[3 * (size_t) (i + 1)]);
        }
        drp[i + 1] = (int32_t) dci.size();
    }
    write_mm(folder + "/out_D.mtx", N, drp, dci, dcv);
    // P: row i holds a one in column perm[i] (new -> old; the identity for
    // static pivots), as the reference writes it (:1902-1910)
    const std::vector<int32_t>& perm = host_int.at("perm");
    std::vector<int32_t> prp(N + 1), pci(N);
    for(int32_t i = 0; i <= N; ++i) prp[i] = i;
    for(int32_t i = 0; i < N; ++i) pci[i] = perm[i];
    write_mm(folder + "/out_P.mtx", N, prp, pci, std::vector<double>(N, 1.0));
    return (2.0 * (double) nnz_u + (double) dci.size()) / (double) A.nnz();
}

// ---------------------------------------------------------------------------

void Spmv::multiply(const double* x, double* y, double alpha, double beta, bool transpose,
    hipStream_t st) const
{
    if(kkt) { kkt_spmv(K, x, y, alpha, beta, st); return; }
    // N is symmetric: the transpose flag is ignored (spmv.impl.cuh:230)
    if(normal) { normal_spmv(N, x, y, alpha, beta, st); return; }
    if(transpose && have_at) csr_spmv(At, x, y, alpha, beta, st);
    else csr_spmv(A, x, y, alpha, beta, st);
}

void Spmv::equilibrate(int32_t sweeps, double* d, hipStream_t st)
{
    if(kkt || normal) fail(CULIP_ERR_UNSUPPORTED, "equilibrate: only CSR operators");
    if(have_at) fail(CULIP_ERR_UNSUPPORTED, "equilibrate: the operator must be symmetric");
    if(A.m != A.n) fail(CULIP_ERR_INVALID, "equilibrate: the operator must be square");
    if(sweeps < 1) fail(CULIP_ERR_INVALID, "equilibrate: at least one sweep");
    b_eq.ensure(sizeof(double) * (size_t) std::max(1, A.m));
    equilibrate_csr(A, sweeps, d, b_eq.as<double>(), st);
    HIP_CHECK(hipGetLastError());
}

void upload_csr(DevCsr& d, DevBuf& br, DevBuf& bc, DevBuf& bv, const HostCsr& h)
{
    upload(br, h.row_ptr);
    upload(bc, h.col);
    upload(bv, h.val);
    d.m = h.m;
    d.n = h.n;
    d.nnz = h.nnz();
    d.row_ptr = br.as<int32_t>();
    d.col = bc.as<int32_t>();
    d.val = bv.as<double>();
}

Sqmr::~Sqmr()
{
    if(host_flags) (void) hipHostFree(host_flags);
    if(host_scal) (void) hipHostFree(host_scal);
}

void Sqmr::init(const Spmv* op, const Precond* p, const culip_sqmr_opts& o)
{
    A = op;
    prec = p;
    opts = o;
    const int32_t n = op->m;
    work.ensure(sqmr_work_bytes(n));
    if(!host_flags)
        HIP_CHECK(hipHostMalloc((void**) &host_flags, 8 * sizeof(int32_t), hipHostMallocDefault));
    if(!host_scal)
        HIP_CHECK(hipHostMalloc((void**) &host_scal, 64 * sizeof(double), hipHostMallocDefault));
    sqmr_work_bind(w, n, work.as<char>(), host_flags, host_scal);
}

void Sqmr::run(const double* b, double* x, culip_sqmr_result& res, hipStream_t st)
{
    if(prec && !prec->ready()) fail(CULIP_ERR_STATE, "sqmr: preconditioner has not been computed");
    SqmrOps ops;
    ops.n = A->m;
    const Spmv* op = A;
    ops.spmv = [op](const double* xx, double* yy, void* s) {
        op->multiply(xx, yy, 1.0, 0.0, false, (hipStream_t) s);
    };
    if(prec) {
        const Precond* P = prec;
        ops.left = [P](double* xx, void* s) { P->left(xx, (hipStream_t) s); };
        ops.right = [P](double* xx, void* s) { P->right(xx, (hipStream_t) s); };
        if(!op->kkt && !op->normal && opts.separate_product == 0)
            ops.spmv_left = [P, op](const double* xx, double* yy, void* s) {
                return P->left_product(op->A, xx, yy, (hipStream_t) s);
            };
        ops.tri = P->tri();
        ops.is_piv = P->is_piv();
    }
    ops.res_scale = res_scale;
    sqmr_solve(ops, w, b, x, opts, res, st, graphs.get());
    if(prec) prec->after_solve(st);
}

}  // namespace culip
