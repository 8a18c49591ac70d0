// Batched refactor + SQMR over independent systems (batch_solver.hpp).
//
// The per-system control flow is sqmr_solve / sqmr_run with a true-residual
// period P > 0 (sqmr.hip): a run iterates, every P iterations its solution
// x = Pr \ (D \ xt) and true residual are formed and the run ends on the
// FP64 gate, a breakdown, a stall or drift; then refinement runs solve for
// the residual while the true residual misses the tolerance. Here every
// system keeps those locals on the host and all running systems advance
// together: one launch per step of an iteration for all of them, one check
// (launches plus one read-back) for all systems at a check, and the ends of
// runs handled in groups. Runs start at checks or at the beginning, so the
// systems' checks fall on the same iterations.
#include "batch_solver.hpp"

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <numeric>

namespace culip {

namespace {

constexpr int SCAL = 64;    // doubles per system in the scalar pool
constexpr int FLAGS = 8;    // ints per system in the flag pool

size_t align256(size_t v) { return (v + 255) & ~(size_t) 255; }
int32_t cdiv(int32_t a, int32_t b) { return (a + b - 1) / b; }

}  // namespace

BatchSolver::ListSlot::~ListSlot()
{
    if(h_ent) (void) hipHostFree(h_ent);
    if(ev) (void) hipEventDestroy(ev);
}

BatchSolver::BatchSolver(const BatchSolverOpts& o) : o_(o)
{
    CULIP_REQUIRE(o.prec == CULIP_PREC_F64 || o.prec == CULIP_PREC_F32, "batch: bad precision");
    CULIP_REQUIRE(o.jacobi_sweeps >= 1, "batch: block-Jacobi solves need at least one sweep");
    CULIP_REQUIRE(o.sqmr.true_residual_period > 0,
        "batch: the batched SQMR needs a true-residual period > 0");
    CULIP_REQUIRE(o.sqmr.max_it >= 0 && o.sqmr.max_restarts >= 0 && o.sqmr.tol >= 0,
        "batch: bad SQMR options");
    HIP_CHECK(hipGetDevice(&device_));
}

BatchSolver::~BatchSolver()
{
    if(seg_exec_) (void) hipGraphExecDestroy(seg_exec_);
    if(sq_) (void) hipStreamDestroy(sq_);
    if(h_scal_) (void) hipHostFree(h_scal_);
    if(h_flags_) (void) hipHostFree(h_flags_);
    if(h_sys_pinned_) (void) hipHostFree(h_sys_pinned_);
}

int32_t BatchSolver::add(const HostCsr& K, const std::vector<int32_t>& piv_starts,
    const std::vector<int32_t>& block_starts)
{
    CULIP_REQUIRE(K.m == K.n && K.m > 0, "batch: a system is a non-empty square matrix");
    {
        std::lock_guard<std::mutex> g(mu_);
        if(finalized_) fail(CULIP_ERR_STATE, "batch: systems are added before the first solve");
    }
    HIP_CHECK(hipSetDevice(device_));
    auto e = std::make_unique<Entry>();
    e->pat = std::make_shared<BildltPattern>();
    e->pat->A = K;
    e->pat->piv_starts = piv_starts;
    e->pat->block_starts = block_starts;
    e->pat->prec = o_.prec;
    e->f = std::make_unique<Bildlt>(e->pat);
    e->f->symbolic(o_.fill_level, o_.fill_factor);
    e->pat->upload_static();
    upload(e->k_row, K.row_ptr);
    upload(e->k_col, K.col);
    e->f->set_solve(CULIP_TRSV_JACOBI, o_.jacobi_sweeps);
    e->f->upload_pattern(nullptr);
    for(int q = 0; q < 3; ++q) e->f->d_jac[q].ensure(sizeof(double) * (size_t) K.m);
    e->n = K.m;
    e->nnz = K.nnz();
    HIP_CHECK(hipDeviceSynchronize());
    std::lock_guard<std::mutex> g(mu_);
    if(finalized_) fail(CULIP_ERR_STATE, "batch: systems are added before the first solve");
    sys_.push_back(std::move(e));
    return (int32_t) sys_.size() - 1;
}

int32_t BatchSolver::count() const
{
    std::lock_guard<std::mutex> g(mu_);
    return (int32_t) sys_.size();
}

int32_t BatchSolver::system_n(int32_t s) const
{
    std::lock_guard<std::mutex> g(mu_);
    CULIP_REQUIRE(s >= 0 && s < (int32_t) sys_.size(), "batch: unknown system");
    return sys_[s]->n;
}

int64_t BatchSolver::system_nnz(int32_t s) const
{
    std::lock_guard<std::mutex> g(mu_);
    CULIP_REQUIRE(s >= 0 && s < (int32_t) sys_.size(), "batch: unknown system");
    return sys_[s]->nnz;
}

int64_t BatchSolver::device_bytes() const
{
    std::lock_guard<std::mutex> g(mu_);
    int64_t b = (int64_t) (d_work_.bytes + d_scal_.bytes + d_flags_.bytes + d_iter_.bytes +
        d_sys_.bytes);
    for(const DevBuf* d : {&d_fsys_, &d_psys_, &d_lr_, &d_chunks_, &d_blk_, &d_scan_, &d_task_})
        b += (int64_t) d->bytes;
    for(const DevBuf& d : d_fit_) b += (int64_t) d.bytes;
    for(const ListSlot& l : slot_) {
        b += (int64_t) l.ent.bytes;
        for(const DevBuf& d : l.items) b += (int64_t) d.bytes;
    }
    for(const auto& e : sys_)
        b += e->f->device_bytes() + (int64_t) (e->k_row.bytes + e->k_col.bytes);
    return b;
}

// pooled SQMR work, scalar and flag pools, the device system tables, work
// lists and the batch's stream, once all systems are known (the first
// solve); a set-up that fails leaves the batch unusable
void BatchSolver::finalize(hipStream_t st)
{
    {
        std::lock_guard<std::mutex> g(mu_);
        if(ready_) return;
        if(finalized_) fail(CULIP_ERR_STATE, "batch: its set-up failed earlier");
        finalized_ = true;
    }
    setup(st);
    ready_ = true;
}

void BatchSolver::setup(hipStream_t st)
{
    const int32_t S = (int32_t) sys_.size();
    CULIP_REQUIRE(S > 0, "batch: no systems");
    std::vector<size_t> off(S);
    size_t total = 0;
    for(int32_t s = 0; s < S; ++s) {
        off[s] = total;
        const size_t vec = align256(sizeof(double) * (size_t) sys_[s]->n);
        total += align256(sqmr_work_bytes(sys_[s]->n)) + 2 * vec;   // work, b, x
    }
    d_work_.ensure(total);
    d_scal_.ensure(sizeof(double) * SCAL * (size_t) S);
    d_flags_.ensure(sizeof(int32_t) * FLAGS * (size_t) S);
    d_iter_.ensure(sizeof(int32_t) * (size_t) S);
    d_sys_.ensure(sizeof(BatchSys) * (size_t) S);
    HIP_CHECK(hipHostMalloc((void**) &h_scal_, sizeof(double) * SCAL * (size_t) S,
        hipHostMallocDefault));
    HIP_CHECK(hipHostMalloc((void**) &h_flags_, sizeof(int32_t) * FLAGS * (size_t) S,
        hipHostMallocDefault));
    HIP_CHECK(hipHostMalloc((void**) &h_sys_pinned_, sizeof(BatchSys) * (size_t) S,
        hipHostMallocDefault));
    // the iteration counters start at zero (each run's start sets them again)
    HIP_CHECK(hipMemsetAsync(d_iter_.p, 0, sizeof(int32_t) * (size_t) S, st));
    work_.assign(S, SqmrWork{});
    h_sys_.assign(S, BatchSys{});
    int64_t tot[5] = {0, 0, 0, 0, 0};
    int64_t maxc[5] = {0, 0, 0, 0, 0};
    for(int32_t s = 0; s < S; ++s) {
        Entry& e = *sys_[s];
        Bildlt& f = *e.f;
        char* base = d_work_.as<char>() + off[s];
        SqmrWork& w = work_[s];
        sqmr_work_bind(w, e.n, base, nullptr, nullptr);
        w.scal = d_scal_.as<double>() + (size_t) SCAL * s;
        w.flags = d_flags_.as<int32_t>() + (size_t) FLAGS * s;
        BatchSys& B = h_sys_[s];
        B.n = e.n;
        B.G = w.grid;
        B.nb = f.topo().nb;
        for(int q = 0; q < 3; ++q) {
            B.v[q] = w.v[q];
            B.w[q] = w.w[q];
            B.p[q] = w.p[q];
            B.jb[q] = f.d_jac[q].as<double>();
        }
        B.z = w.z;
        B.tv = w.tv;
        B.ev = w.ev;
        B.av[0] = w.av[0];
        B.av[1] = w.av[1];
        B.xt = w.xt;
        B.rt = w.rt;
        B.xacc = w.xacc;
        B.rhs = w.rhs;
        B.part = w.part;
        B.scal = w.scal;
        B.flags = w.flags;
        B.iter = d_iter_.as<int32_t>() + s;
        B.tri = f.tri();
        B.ip = f.is_piv();
        B.k_rp = e.k_row.as<int32_t>();
        B.k_col = e.k_col.as<int32_t>();
        B.k_val = f.d_vals.as<double>();
        const size_t vec = align256(sizeof(double) * (size_t) e.n);
        double* bvec = (double*) (base + align256(sqmr_work_bytes(e.n)));
        B.b = bvec;
        B.x = (double*) ((char*) bvec + vec);
        auto set_plan = [&](const auto& p) {
            B.task[0] = p.task_u;
            B.task[1] = p.task_t;
            B.o_col[0] = p.uo_col;
            B.o_val[0] = p.uo_val;
            B.o_lr[0] = p.uo_lr;
            B.i_col[0] = p.ui_col;
            B.i_val[0] = p.ui_val;
            B.i_lr[0] = p.ui_lr;
            B.o_col[1] = p.to_col;
            B.o_val[1] = p.to_val;
            B.o_lr[1] = p.to_lr;
            B.i_col[1] = p.ti_col;
            B.i_val[1] = p.ti_val;
            B.i_lr[1] = p.ti_lr;
        };
        if(f.f32()) set_plan(f.plan<float>());
        else set_plan(f.plan<double>());
        const int64_t c[5] = {B.G, cdiv(B.n, 256), cdiv(B.nb, JACOBI_BATCH_ROWS),
            cdiv(B.n, BATCH_SPMV_ROWS), 1};
        for(int k = 0; k < 5; ++k) {
            tot[k] += c[k];
            maxc[k] = std::max(maxc[k], c[k]);
        }
    }
    // an XCD-aware list is at most BATCH_XCDS times its largest group, and
    // a group (filled largest system first into the lightest group) exceeds
    // the mean by at most one system's items
    for(int k = 0; k < 5; ++k) tot[k] += (int64_t) BATCH_XCDS * maxc[k];
    for(int k = 0; k < 5; ++k)
        CULIP_REQUIRE(tot[k] < (int64_t) INT32_MAX, "batch: too many work items for one launch");
    for(ListSlot& l : slot_) {
        l.ent.ensure(sizeof(int4) * 5 * (size_t) S);
        for(int k = 0; k < 5; ++k) l.items[k].ensure(sizeof(int2) * (size_t) std::max<int64_t>(1, tot[k]));
        HIP_CHECK(hipHostMalloc((void**) &l.h_ent, sizeof(int4) * 5 * (size_t) S,
            hipHostMallocDefault));
        HIP_CHECK(hipEventCreateWithFlags(&l.ev, hipEventDisableTiming));
    }
    if(o_.prec == CULIP_PREC_F32) build_factor<float>();
    else build_factor<double>();
    HIP_CHECK(hipStreamCreateWithFlags(&sq_, hipStreamNonBlocking));
    st_.assign(S, State{});
    upload_sys(st);
}

void BatchSolver::upload_sys(hipStream_t st)
{
    // the pinned copy may still be read by the previous upload
    HIP_CHECK(hipStreamSynchronize(st));
    std::memcpy(h_sys_pinned_, h_sys_.data(), sizeof(BatchSys) * h_sys_.size());
    HIP_CHECK(hipMemcpyAsync(d_sys_.p, h_sys_pinned_, sizeof(BatchSys) * h_sys_.size(),
        hipMemcpyHostToDevice, st));
}

void BatchSolver::read_back(hipStream_t st)
{
    const size_t S = sys_.size();
    HIP_CHECK(hipMemcpyAsync(h_scal_, d_scal_.p, sizeof(double) * SCAL * S,
        hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipMemcpyAsync(h_flags_, d_flags_.p, sizeof(int32_t) * FLAGS * S,
        hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
}

// the five item lists of the systems `ids` (expanded on the device from one
// entry per system and kind); an unchanged set reuses its lists
const BatchLists& BatchSolver::lists(ListSlot& slot, const std::vector<int32_t>& ids,
    hipStream_t st)
{
    if(slot.valid && slot.ids == ids) return slot.L;
    const int32_t m = (int32_t) ids.size();
    // the staging buffer may still be read by the slot's previous upload
    if(slot.ev_used) HIP_CHECK(hipEventSynchronize(slot.ev));
    int32_t total[5] = {0, 0, 0, 0, 0};
    for(int32_t k = 0; k < m; ++k) {
        const BatchSys& B = h_sys_[ids[k]];
        const int32_t cnt[5] = {B.G, cdiv(B.n, 256), cdiv(B.nb, JACOBI_BATCH_ROWS),
            cdiv(B.n, BATCH_SPMV_ROWS), 1};
        for(int q = 0; q < 5; ++q) {
            slot.h_ent[q * m + k] = make_int4(ids[k], cnt[q], total[q], 0);
            total[q] += cnt[q];
        }
    }
    // the product's and the sweeps' lists XCD-aware (device/batch.hpp)
    int32_t inter[5] = {0, 0, 0, 0, 0};
    if(m >= BATCH_XCD_MIN_SYSTEMS) {
        std::vector<int32_t> order(m);
        for(int q : {2, 3}) {
            int4* e = slot.h_ent + (size_t) q * m;
            for(int32_t k = 0; k < m; ++k) order[k] = k;
            std::stable_sort(order.begin(), order.end(),
                [&](int32_t a, int32_t b) { return e[a].y > e[b].y; });
            int64_t load[BATCH_XCDS] = {};
            for(int32_t k : order) {
                int g = 0;
                for(int x = 1; x < BATCH_XCDS; ++x)
                    if(load[x] < load[g]) g = x;
                e[k].z = (int32_t) load[g];
                e[k].w = g;
                load[g] += e[k].y;
            }
            const int64_t mx = *std::max_element(load, load + BATCH_XCDS);
            const int64_t t = (int64_t) BATCH_XCDS * mx;
            CULIP_REQUIRE(t <= (int64_t) (slot.items[q].bytes / sizeof(int2)),
                "batch: XCD-aware list exceeds its buffer");
            inter[q] = (int32_t) t;
            total[q] = (int32_t) t;
        }
    }
    if(m > 0) {
        HIP_CHECK(hipMemcpyAsync(slot.ent.p, slot.h_ent, sizeof(int4) * 5 * (size_t) m,
            hipMemcpyHostToDevice, st));
        HIP_CHECK(hipEventRecord(slot.ev, st));
        slot.ev_used = true;
        for(int q = 0; q < 5; ++q)
            batch_expand(slot.ent.as<int4>() + (size_t) q * m, m, slot.items[q].as<int2>(),
                inter[q], st);
    }
    BatchList* dst[5] = {&slot.L.vec, &slot.L.row, &slot.L.blk, &slot.L.spmv, &slot.L.sys};
    slot.L.one_round = true;
    for(int32_t k = 0; k < m; ++k) {
        const BatchSys& B = h_sys_[ids[k]];
        if((int64_t) B.n > (int64_t) B.G * 256) slot.L.one_round = false;
    }
    for(int q = 0; q < 5; ++q) {
        dst[q]->items = slot.items[q].as<int2>();
        dst[q]->count = total[q];
    }
    slot.ids = ids;
    slot.valid = true;
    return slot.L;
}

void BatchSolver::jacobi(const BatchLists& L, bool transpose, int32_t target,
    int32_t product_src, hipStream_t st)
{
    const BatchSys* d = d_sys_.as<BatchSys>();
    if(o_.prec == CULIP_PREC_F32)
        batch_jacobi<float>(L, d, transpose, o_.jacobi_sweeps, target, product_src, st);
    else
        batch_jacobi<double>(L, d, transpose, o_.jacobi_sweeps, target, product_src, st);
    HIP_CHECK(hipGetLastError());
}

// the start of sqmr_run: zero vectors, v0 = Pl \ b normalised, w0, and
// av[0] = Pl \ K Pr \ ev (iteration counters to 0)
void BatchSolver::start_runs(const BatchLists& L, hipStream_t st)
{
    const BatchSys* d = d_sys_.as<BatchSys>();
    batch_step(BS_INIT_FILL, L, d, 0, 0, 0, 0, st);
    batch_step(BS_COPY, L, d, BV_BRUN, BV_TV, 0, 0, st);
    jacobi(L, true, BV_TV, BV_NONE, st);
    batch_step(BS_SUMSQ, L, d, BV_TV, 0, 0, 0, st);
    batch_step(BS_INIT_V, L, d, 0, 0, 0, 0, st);
    batch_step(BS_INIT_W, L, d, 0, 0, 0, 0, st);
    jacobi(L, false, BV_EV, BV_NONE, st);
    jacobi(L, true, BV_AVCUR, BV_EV, st);
}

// launch_iteration of sqmr_run for every listed system
void BatchSolver::iteration(const BatchLists& L, hipStream_t st)
{
    const BatchSys* d = d_sys_.as<BatchSys>();
    batch_step(BS_ALPHA, L, d, 0, 0, 0, 0, st);
    batch_step(BS_VNEXT, L, d, 0, 0, 0, 0, st);
    batch_step(BS_VSCALE, L, d, 0, 0, 0, 0, st);
    jacobi(L, false, BV_EV, BV_NONE, st);
    jacobi(L, true, BV_AVNXT, BV_EV, st);
    batch_step(BS_BETA, L, d, 0, 0, 0, 0, st);
    batch_step(BS_PUPDATE, L, d, 0, 0, 0, 0, st);
    batch_step(BS_SCALARS, L, d, 0, 0, o_.sqmr.tol, SQMR_BREAKDOWN_REL, st);
}

void BatchSolver::run_solution(const BatchLists& L, hipStream_t st)
{
    batch_step(BS_DSOLVE, L, d_sys_.as<BatchSys>(), BV_XT, BV_X, 0, 0, st);
    jacobi(L, false, BV_X, BV_NONE, st);
}

// residual_norms of sqmr.hip: rt = bvec - K x, scal[40] = ||rt||,
// scal[41] = ||bvec||
void BatchSolver::residual_norms(const BatchLists& L, int32_t bvec, hipStream_t st)
{
    const BatchSys* d = d_sys_.as<BatchSys>();
    batch_csr_spmv(L, d, BV_X, BV_RT, st);
    batch_step(BS_RESIDUAL, L, d, bvec, BV_RT, 0, 0, st);
    batch_step(BS_NORM2_PAIR, L, d, BV_RT, bvec, 0, 0, st);
    batch_step(BS_NORM_FINAL, L, d, 0, 0, 0, 0, st);
}

// the tables and work items of the batched refactor (static for the batch:
// they follow from the symbolic analyses)
template<typename VT>
void BatchSolver::build_factor()
{
    const int32_t S = (int32_t) sys_.size();
    std::vector<LevelRange> lrs;
    std::vector<size_t> lr_off(S);
    int32_t max_levels = 0;
    for(int32_t s = 0; s < S; ++s) {
        const Symbolic& y = sys_[s]->f->sym;
        for(int32_t b = 0; b < y.nb; ++b)
            max_h_ = std::max(max_h_, y.block_starts[b + 1] - y.block_starts[b]);
        lr_off[s] = lrs.size();
        for(int32_t l = 0; l < y.num_levels; ++l) {
            LevelRange r;
            r.lo = y.lvlb_off[l];
            r.size = y.lvlb_off[l + 1] - r.lo;
            r.npiv = y.lvl_off[l + 1] - y.lvl_off[l];
            r.row_lo = y.lvl_off[l];
            r.level = l;
            lrs.push_back(r);
        }
        max_levels = std::max(max_levels, y.num_levels);
    }
    // the solve layout's LDS tile of a pivot-block inverse is max_h_ squared
    CULIP_REQUIRE(max_h_ >= 1 && max_h_ <= 32, "batch: block height outside 1..32");
    upload(d_lr_, lrs);
    std::vector<FactorSys<VT>> F(S);
    std::vector<PlanSys<VT>> P(S);
    std::vector<int2> chunks, blk, scan, task;
    for(int32_t s = 0; s < S; ++s) {
        Bildlt& f = *sys_[s]->f;
        f.threshold = o_.threshold;
        FactorSys<VT>& q = F[s];
        q.a = f.factor_args<VT>();
        q.lr = d_lr_.as<LevelRange>() + lr_off[s];
        q.vals = f.d_vals.as<double>();
        q.value_slot = f.d_value_slot.as<int64_t>();
        q.nnz = f.A.nnz();
        q.val_space = f.sym.val_space;
        q.val_in = f.d_val_in.as<VT>();
        q.static_is_piv = f.d_is_piv.as<int32_t>();
        PlanSys<VT>& p = P[s];
        p.t = f.topo();
        p.p = f.plan<VT>();
        p.val_out = q.a.val_out;
        p.totals = f.d_totals.as<int32_t>();
        p.scan_tmp = f.d_scan_tmp.as<int32_t>();
        const int64_t nch = factor_batch_chunks(f.n(), q.nnz, q.val_space, f.sym.num_levels,
            q.a.scan_maxch);
        for(int64_t c = 0; c < nch; ++c) chunks.push_back(make_int2(s, (int) c));
        for(int32_t r = 0; r < p.t.nb; ++r) blk.push_back(make_int2(s, r));
        for(int32_t c = 0; c < scan_chunks(f.n()); ++c) scan.push_back(make_int2(s, c));
        for(int32_t c = 0; c < (p.t.nb + 255) / 256; ++c) task.push_back(make_int2(s, c));
    }
    d_fsys_.ensure(sizeof(FactorSys<VT>) * (size_t) S);
    HIP_CHECK(hipMemcpy(d_fsys_.p, F.data(), sizeof(FactorSys<VT>) * (size_t) S,
        hipMemcpyHostToDevice));
    d_psys_.ensure(sizeof(PlanSys<VT>) * (size_t) S);
    HIP_CHECK(hipMemcpy(d_psys_.p, P.data(), sizeof(PlanSys<VT>) * (size_t) S,
        hipMemcpyHostToDevice));
    auto put = [](DevBuf& d, const std::vector<int2>& v) {
        CULIP_REQUIRE(v.size() < (size_t) INT32_MAX, "batch: too many work items for one launch");
        upload(d, v);
        return FactorItems{d.as<int2>(), (int32_t) v.size()};
    };
    chunks_ = put(d_chunks_, chunks);
    blk_ = put(d_blk_, blk);
    scan_ = put(d_scan_, scan);
    task_ = put(d_task_, task);
    // per level and launch: the workgroups of every system that has the level
    std::vector<int2> it[7];
    std::vector<std::array<size_t, 7>> off(max_levels + 1);
    for(int32_t l = 0; l < max_levels; ++l) {
        for(int k = 0; k < 7; ++k) off[l][k] = it[k].size();
        for(int32_t s = 0; s < S; ++s) {
            if(l >= sys_[s]->f->sym.num_levels) continue;
            const LevelRange& r = lrs[lr_off[s] + l];
            const int32_t noff = r.size - r.npiv;
            const int32_t cnt[7] = {fl_grid_pivots_pairs(r), fl_tiny_blocks(r),
                fl_grid_tasks(noff, BATCH_TASK_GRID), r.npiv, fl_grid_level_scan(r),
                fl_grid_store_tiny(r), fl_grid_tasks(r.size, BATCH_TASK_GRID)};
            for(int k = 0; k < 7; ++k)
                for(int32_t q = 0; q < cnt[k]; ++q) it[k].push_back(make_int2(s, q));
        }
    }
    for(int k = 0; k < 7; ++k) off[max_levels][k] = it[k].size();
    FactorItems all[7];
    for(int k = 0; k < 7; ++k) all[k] = put(d_fit_[k], it[k]);
    flevels_.assign(max_levels, FactorBatchLevel{});
    for(int32_t l = 0; l < max_levels; ++l) {
        FactorItems* dst[7] = {&flevels_[l].pp, &flevels_[l].ts, &flevels_[l].off,
            &flevels_[l].rn, &flevels_[l].sc, &flevels_[l].stt, &flevels_[l].st};
        for(int k = 0; k < 7; ++k) {
            dst[k]->items = all[k].items + off[l][k];
            dst[k]->count = (int32_t) (off[l + 1][k] - off[l][k]);
        }
    }
}

// factor_impl of every system: one launch per step over all systems (the
// values are already in each factor's value buffer)
template<typename VT>
void BatchSolver::factor_batched(hipStream_t st)
{
    const FactorSys<VT>* fs = d_fsys_.as<FactorSys<VT>>();
    factor_batch_begin<VT>(fs, chunks_, st);
    for(int32_t l = 0; l < (int32_t) flevels_.size(); ++l)
        factor_batch_level<VT>(fs, flevels_[l], l, st);
    plan_batch<VT>(d_psys_.as<PlanSys<VT>>(), blk_, scan_, task_, max_h_, st);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
    for(auto& e : sys_) e->f->after_factor();
}

void BatchSolver::capture_segment(const BatchLists& L, int32_t period,
    const std::vector<int32_t>& ids, hipStream_t st)
{
    if(seg_exec_) (void) hipGraphExecDestroy(seg_exec_);
    seg_exec_ = nullptr;
    seg_ids_.clear();
    HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        for(int32_t j = 0; j < period; ++j) iteration(L, st);
    } catch(...) {
        hipGraph_t bad = nullptr;
        (void) hipStreamEndCapture(st, &bad);
        if(bad) (void) hipGraphDestroy(bad);
        throw;
    }
    hipGraph_t g = nullptr;
    HIP_CHECK(hipStreamEndCapture(st, &g));
    const hipError_t e = hipGraphInstantiate(&seg_exec_, g, nullptr, nullptr, 0);
    (void) hipGraphDestroy(g);
    hip_check(e, "batch: segment graph instantiation");
    seg_ids_ = ids;
}

void BatchSolver::solve(const double* const* vals, const double* const* rhs, double* const* x,
    culip_sqmr_result* res, bool device_arrays, hipStream_t user)
{
    CULIP_REQUIRE(vals && rhs && x, "batch: null argument");
    HIP_CHECK(hipSetDevice(device_));
    finalize(user);
    // the caller's stream is done with everything before the call; the
    // batch runs on its own stream and ends synchronised
    HIP_CHECK(hipStreamSynchronize(user));
    hipStream_t st = sq_;
    seg_prev_.clear();
    const int32_t S = (int32_t) sys_.size();
    for(int32_t s = 0; s < S; ++s) CULIP_REQUIRE(vals[s] && rhs[s] && x[s], "batch: null array");
    const culip_sqmr_opts& O = o_.sqmr;
    const int32_t period = O.true_residual_period;
    const BatchSys* d = d_sys_.as<BatchSys>();

    const hipMemcpyKind in_kind = device_arrays ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    const auto t0 = std::chrono::steady_clock::now();
    for(int32_t s = 0; s < S; ++s)
        HIP_CHECK(hipMemcpyAsync(sys_[s]->f->d_vals.p, vals[s], sizeof(double) * (size_t) sys_[s]->nnz,
            in_kind, st));
    if(o_.prec == CULIP_PREC_F32) factor_batched<float>(st);
    else factor_batched<double>(st);
    const auto t1 = std::chrono::steady_clock::now();
    for(int32_t s = 0; s < S; ++s) {
        HIP_CHECK(hipMemcpyAsync(const_cast<double*>(h_sys_[s].b), rhs[s],
            sizeof(double) * (size_t) sys_[s]->n, in_kind, st));
        h_sys_[s].brun_rhs = 0;
        st_[s] = State{};
    }
    upload_sys(st);
    std::vector<int32_t> all(S);
    std::iota(all.begin(), all.end(), 0);
    // gate: tol * ||b|| (sqmr_solve)
    {
        const BatchLists& L = lists(slot_[0], all, st);
        batch_step(BS_NORM2_PAIR, L, d, BV_B, BV_B, 0, 0, st);
        batch_step(BS_NORM_FINAL, L, d, 0, 0, 0, 0, st);
        read_back(st);
    }
    auto start_state = [&](State& q, bool restart) {
        q.running = true;
        q.restart_run = restart;
        q.it = 0;
        q.run_max = O.max_it - q.total_it;
        q.gate_met = false;
        q.at_check = false;
        q.best_nr = -1.0;
        q.gam_at_best = 0.0;
        q.stall = 0;
    };
    for(int32_t s = 0; s < S; ++s) {
        State& q = st_[s];
        q.gate_abs = O.tol * h_scal_[(size_t) SCAL * s + 40];
        start_state(q, false);
    }
    start_runs(lists(slot_[0], all, st), st);

    auto scal = [&](int32_t s, int k) { return h_scal_[(size_t) SCAL * s + k]; };
    auto flag = [&](int32_t s, int k) { return h_flags_[(size_t) FLAGS * s + k]; };
    std::vector<int32_t> act, chk, ended, need, first, refined, zeroed, reverted, again;
    std::vector<char> is_chk(S, 0), is_end(S, 0);
    for(;;) {
        act.clear();
        for(int32_t s = 0; s < S; ++s)
            if(st_[s].running && st_[s].it < st_[s].run_max) act.push_back(s);
        ended.clear();
        std::fill(is_chk.begin(), is_chk.end(), 0);
        std::fill(is_end.begin(), is_end.end(), 0);
        if(!act.empty()) {
            const BatchLists& L = lists(slot_[1], act, st);
            // a whole segment (every running system at a check boundary with
            // `period` iterations left) replays the captured graph when the
            // same systems ran the previous segment too
            bool aligned = period > 1;
            for(int32_t s : act)
                if(st_[s].it % period != 0 || st_[s].it + period > st_[s].run_max) {
                    aligned = false;
                    break;
                }
            int32_t steps = 1;
            if(aligned && act == seg_prev_) {
                if(!seg_exec_ || seg_ids_ != act) capture_segment(L, period, act, st);
                HIP_CHECK(hipGraphLaunch(seg_exec_, st));
                steps = period;
            } else {
                iteration(L, st);
            }
            if(aligned) seg_prev_ = act;
            for(int32_t s : act) st_[s].it += steps - 1;
            chk.clear();
            for(int32_t s : act)
                if((st_[s].it + 1) % period == 0) {
                    chk.push_back(s);
                    is_chk[s] = 1;
                }
            if(!chk.empty()) {
                const BatchLists& LC = chk == act ? L : lists(slot_[2], chk, st);
                run_solution(LC, st);
                residual_norms(LC, BV_BRUN, st);
                read_back(st);
                for(int32_t s : chk) {
                    State& q = st_[s];
                    const double nr = scal(s, 40);
                    if(nr <= q.gate_abs) {
                        q.gate_met = true;
                        is_end[s] = 1;
                        continue;
                    }
                    const double gam = std::fabs(scal(s, SQMR_S_GAMMA));
                    if(q.best_nr < 0 || nr < 0.5 * q.best_nr) {
                        q.best_nr = nr;
                        q.gam_at_best = gam;
                        q.stall = 0;
                    } else {
                        ++q.stall;
                    }
                    const bool drifted = q.stall > 0 && gam < 1e-3 * q.gam_at_best;
                    if(flag(s, 0) || q.stall >= SQMR_STALL_PERIODS || drifted) is_end[s] = 1;
                }
            }
            for(int32_t s : act) {
                State& q = st_[s];
                ++q.it;
                if(q.it >= q.run_max) is_end[s] = 1;
                if(is_end[s]) q.at_check = is_chk[s] != 0;
            }
        }
        for(int32_t s = 0; s < S; ++s)
            if(st_[s].running && (is_end[s] || st_[s].it >= st_[s].run_max)) ended.push_back(s);
        if(ended.empty()) {
            if(act.empty()) break;
            continue;
        }

        // the end of sqmr_run: x = Pr \ (D \ xt) unless a check just formed it
        need.clear();
        for(int32_t s : ended)
            if(!st_[s].at_check) need.push_back(s);
        if(!need.empty()) run_solution(lists(slot_[2], need, st), st);
        read_back(st);
        first.clear();
        refined.clear();
        for(int32_t s : ended) {
            State& q = st_[s];
            q.running = false;
            q.total_it += q.it;
            q.run_res = culip_sqmr_result{};
            q.run_res.residual = scal(s, SQMR_S_GAMMA) / scal(s, SQMR_S_NORMPB);
            q.run_res.converged = q.gate_met || flag(s, 0) || flag(s, 1);
            (q.restart_run ? refined : first).push_back(s);
        }
        // sqmr_solve after the first run: its true residual against b (the
        // check's, when the run ended at one: the same x and b)
        need.clear();
        for(int32_t s : first)
            if(!st_[s].at_check) need.push_back(s);
        if(!need.empty()) {
            residual_norms(lists(slot_[2], need, st), BV_B, st);
            read_back(st);
        }
        zeroed.clear();
        again.clear();
        for(int32_t s : first) {
            State& q = st_[s];
            q.res = q.run_res;
            const double nr = scal(s, 40), nb = scal(s, 41);
            q.rr = nb > 0 ? nr / nb : nr;
            if(nb > 0 && !(q.rr <= 1.0)) {
                zeroed.push_back(s);
                q.rr = 1.0;
            }
            again.push_back(s);
        }
        if(!zeroed.empty()) {
            // worse than the zero start: x = 0, residual b
            const BatchLists& LZ = lists(slot_[2], zeroed, st);
            batch_step(BS_ZERO, LZ, d, BV_X, 0, 0, 0, st);
            batch_step(BS_COPY, LZ, d, BV_B, BV_RT, 0, 0, st);
        }
        // after a refinement run: x += x_acc, its true residual against b;
        // a correction that does not lower it is undone
        if(!refined.empty()) {
            const BatchLists& LR = lists(slot_[3], refined, st);
            batch_step(BS_AXPY1, LR, d, BV_XACC, BV_X, 0, 0, st);
            residual_norms(LR, BV_B, st);
            read_back(st);
            reverted.clear();
            for(int32_t s : refined) {
                State& q = st_[s];
                const double nr = scal(s, 40), nb = scal(s, 41);
                const double rr_new = nb > 0 ? nr / nb : nr;
                ++q.restarts;
                if(!(rr_new < q.rr)) {
                    reverted.push_back(s);
                    q.done = true;
                    continue;
                }
                q.rr = rr_new;
                q.res.residual = q.run_res.residual;
                again.push_back(s);
            }
            if(!reverted.empty())
                batch_step(BS_COPY, lists(slot_[2], reverted, st), d, BV_XACC, BV_X, 0, 0, st);
        }
        // the refinement loop's condition
        need.clear();
        for(int32_t s : again) {
            State& q = st_[s];
            if(q.rr > O.tol && q.restarts < O.max_restarts && q.total_it < O.max_it)
                need.push_back(s);
            else
                q.done = true;
        }
        if(!need.empty()) {
            std::sort(need.begin(), need.end());
            const BatchLists& LP = lists(slot_[2], need, st);
            batch_step(BS_COPY, LP, d, BV_X, BV_XACC, 0, 0, st);
            batch_step(BS_COPY, LP, d, BV_RT, BV_RHS, 0, 0, st);
            for(int32_t s : need) {
                h_sys_[s].brun_rhs = 1;
                start_state(st_[s], true);
            }
            upload_sys(st);
            start_runs(LP, st);
        }
    }

    for(int32_t s = 0; s < S; ++s) {
        HIP_CHECK(hipMemcpyAsync(x[s], h_sys_[s].x, sizeof(double) * (size_t) sys_[s]->n,
            device_arrays ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
        if(res) {
            State& q = st_[s];
            culip_sqmr_result r = q.res;
            r.iterations = q.total_it;
            r.restarts = q.restarts;
            r.true_rel_residual = q.rr;
            r.converged = q.rr <= O.tol;
            res[s] = r;
        }
    }
    HIP_CHECK(hipStreamSynchronize(st));
    const auto t2 = std::chrono::steady_clock::now();
    sec_[0] = std::chrono::duration<double>(t1 - t0).count();
    sec_[1] = std::chrono::duration<double>(t2 - t1).count();
}

}  // namespace culip
