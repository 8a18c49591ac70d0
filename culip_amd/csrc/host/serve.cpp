// Request loop: resident patterns, per-slot numeric instances, admission by
// device-memory budget, overlapping solves on one stream per slot.
#include "serve.hpp"

#include <algorithm>

#include "../device/scale.hpp"

namespace culip {

Server::Server(const culip_server_opts& o) : o_(o)
{
    CULIP_REQUIRE(o.slots >= 1 && o.slots <= 64, "server: slots must be in [1, 64]");
    CULIP_REQUIRE(o.prec == CULIP_PREC_F64 || o.prec == CULIP_PREC_F32, "server: unknown precision");
    CULIP_REQUIRE(o.fill_level >= 0 && o.fill_factor >= 1.0, "server: bad fill parameters");
    CULIP_REQUIRE(o.sqmr.max_it >= 0 && o.sqmr.tol >= 0, "server: bad SQMR options");
    CULIP_REQUIRE(o.equilibrate_sweeps >= 0 && o.equilibrate_sweeps <= 64,
        "server: equilibrate_sweeps must be in [0, 64]");
    CULIP_REQUIRE(o.solve_algorithm == CULIP_TRSV_LEVELS || o.solve_algorithm == CULIP_TRSV_SINGLE ||
            (o.solve_algorithm == CULIP_TRSV_JACOBI && o.jacobi_sweeps >= 1),
        "server: bad solve algorithm");
    HIP_CHECK(hipGetDevice(&device_));
    size_t fr = 0, tot = 0;
    HIP_CHECK(hipMemGetInfo(&fr, &tot));
    budget_ = o.hbm_budget > 0 ? o.hbm_budget : (int64_t) (0.9 * (double) fr);
    t0_ = std::chrono::steady_clock::now();
    for(int32_t i = 0; i < o.slots; ++i) {
        auto s = std::make_unique<Slot>();
        s->index = i;
        HIP_CHECK(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
        slots_.push_back(std::move(s));
    }
    for(auto& s : slots_) s->th = std::thread(&Server::worker, this, s.get());
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    cv_mem_.notify_all();
    for(auto& s : slots_)
        if(s->th.joinable()) s->th.join();
    (void) hipSetDevice(device_);
    for(auto& s : slots_) {
        s->cache.clear();
        if(s->st) (void) hipStreamDestroy(s->st);
    }
}

double Server::now() const
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

int32_t Server::add_pattern(const HostCsr& K, const std::vector<int32_t>& piv_starts,
    const std::vector<int32_t>& block_starts)
{
    HIP_CHECK(hipSetDevice(device_));
    auto P = std::make_unique<ServePattern>();
    P->pat = std::make_shared<BildltPattern>();
    P->pat->A = K;
    P->pat->piv_starts = piv_starts;
    P->pat->block_starts = block_starts;
    P->pat->prec = o_.prec;
    P->algo = o_.solve_algorithm;
    P->sweeps = o_.jacobi_sweeps;
    {
        Bildlt tmp(P->pat);
        tmp.symbolic(o_.fill_level, o_.fill_factor);
        estimate(*P);
    }
    P->pat->upload_static();
    upload(P->k_row, K.row_ptr);
    upload(P->k_col, K.col);
    P->N = K.m;
    P->nnz = K.nnz();
    P->static_bytes = P->pat->device_bytes() + (int64_t) (P->k_row.bytes + P->k_col.bytes);
    std::lock_guard<std::mutex> g(mu_);
    if(used_ + P->static_bytes > budget_)
        fail(CULIP_ERR_BUDGET, "server: pattern does not fit the device memory budget");
    used_ += P->static_bytes;
    peak_ = std::max(peak_, used_);
    patterns_.push_back(std::move(P));
    return (int32_t) patterns_.size() - 1;
}

int32_t Server::pattern_n(int32_t p) const
{
    std::lock_guard<std::mutex> g(mu_);
    CULIP_REQUIRE(p >= 0 && p < (int32_t) patterns_.size(), "server: unknown pattern");
    return patterns_[p]->N;
}

// device bytes of one instance of P: an upper bound, and the part without
// the solve layout (the pattern's symbolic analysis must exist)
void Server::estimate(ServePattern& P) const
{
    Bildlt tmp(P.pat);
    tmp.solve_algo = P.algo;
    const int64_t nvec = o_.equilibrate_sweeps > 0 ? 4 : 2;   // rhs, x (+ scale, work)
    const int64_t vec = (int64_t) sqmr_work_bytes(P.pat->A.m) + nvec * 8 * (int64_t) P.pat->A.m + 512;
    P.est_max = tmp.numeric_bytes_estimate(true) + vec;
    P.est_min = tmp.numeric_bytes_estimate(false) + vec;
}

void Server::set_pattern_solve(int32_t pattern, culip_solve_algorithm a, int32_t sweeps)
{
    CULIP_REQUIRE(a == CULIP_TRSV_LEVELS || a == CULIP_TRSV_SINGLE ||
            (a == CULIP_TRSV_JACOBI && sweeps >= 1),
        "server: bad solve algorithm");
    std::lock_guard<std::mutex> g(mu_);
    CULIP_REQUIRE(pattern >= 0 && pattern < (int32_t) patterns_.size(), "server: unknown pattern");
    ServePattern& P = *patterns_[pattern];
    if(P.submitted > 0)
        fail(CULIP_ERR_STATE, "server: the solve schedule is fixed once a request was submitted");
    P.algo = a;
    P.sweeps = sweeps;
    estimate(P);
}

int64_t Server::submit(int32_t pattern, const double* vals, const double* rhs)
{
    CULIP_REQUIRE(vals && rhs, "server: null argument");
    auto r = std::make_unique<ServeRequest>();
    int64_t nnz = 0, N = 0;
    {
        std::lock_guard<std::mutex> g(mu_);
        CULIP_REQUIRE(!stop_, "server: stopped");
        CULIP_REQUIRE(pattern >= 0 && pattern < (int32_t) patterns_.size(), "server: unknown pattern");
        nnz = patterns_[pattern]->nnz;
        N = patterns_[pattern]->N;
    }
    r->pattern = pattern;
    r->vals.assign(vals, vals + nnz);
    r->rhs.assign(rhs, rhs + N);
    std::lock_guard<std::mutex> g(mu_);
    ++patterns_[pattern]->submitted;
    r->id = next_id_++;
    r->t_submit = now();
    const int64_t id = r->id;
    queue_.push_back(id);
    reqs_[id] = std::move(r);
    cv_work_.notify_one();
    return id;
}

void Server::wait(int64_t id, double* x, culip_sqmr_result* res, double* timing, int32_t* slot)
{
    std::unique_ptr<ServeRequest> r;
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(id);
        CULIP_REQUIRE(it != reqs_.end(), "server: unknown request");
        ServeRequest* p = it->second.get();
        cv_done_.wait(lk, [p] { return p->done; });
        r = std::move(it->second);
        reqs_.erase(it);
    }
    if(r->status != CULIP_OK) fail(r->status, r->err);
    if(x) std::copy(r->x.begin(), r->x.end(), x);
    if(res) *res = r->res;
    if(timing) {
        timing[0] = r->t_start - r->t_submit;   // queued
        timing[1] = r->t_end - r->t_start;      // service
        timing[2] = r->t_end - r->t_submit;     // latency
        timing[3] = r->t_submit;
    }
    if(slot) *slot = r->slot;
}

void Server::stats(int64_t* out) const
{
    std::lock_guard<std::mutex> g(mu_);
    out[0] = used_;
    out[1] = peak_;
    out[2] = budget_;
    out[3] = instances_;
    out[4] = completed_;
}

// Instance of `pattern` for slot s, created under the memory budget; idle
// instances of any slot are released, least recently used first, to make
// room. Called and returns with lk held.
ServeInstance* Server::instance(Slot* s, int32_t pattern, std::unique_lock<std::mutex>& lk)
{
    auto it = s->cache.find(pattern);
    if(it != s->cache.end()) {
        it->second->last_use = ++tick_;
        it->second->busy = true;
        return it->second.get();
    }
    ServePattern& P = *patterns_[pattern];
    // the measured footprint once an instance of the pattern has run, an
    // upper bound before; a first instance whose bound does not fit is
    // still tried when its allocations without the solve layout fit
    int64_t need = P.inst_bytes > 0 ? P.inst_bytes : P.est_max;
    while(used_ + need > budget_) {
        Slot* vs = nullptr;
        int32_t vp = -1;
        uint64_t oldest = ~0ull;
        bool any_busy = false;
        for(auto& t : slots_)
            for(auto& kv : t->cache) {
                if(kv.second->busy) { any_busy = true; continue; }
                if(kv.second->last_use < oldest) {
                    oldest = kv.second->last_use;
                    vs = t.get();
                    vp = kv.first;
                }
            }
        if(vs) {
            std::unique_ptr<ServeInstance> victim = std::move(vs->cache[vp]);
            vs->cache.erase(vp);
            used_ -= victim->bytes;
            --instances_;
            lk.unlock();
            victim.reset();
            lk.lock();
            continue;
        }
        // instances being created by other slots will become busy or idle
        if(creating_ > 0) any_busy = true;
        if(!any_busy) {
            if(P.inst_bytes == 0 && used_ + P.est_min <= budget_) {
                need = P.est_min;
                break;
            }
            fail(CULIP_ERR_BUDGET, "server: request does not fit the device memory budget");
        }
        ++mem_waiters_;
        cv_mem_.wait(lk);
        --mem_waiters_;
        if(stop_) fail(CULIP_ERR_STATE, "server: stopped");
    }
    used_ += need;
    peak_ = std::max(peak_, used_);
    ++creating_;
    lk.unlock();
    std::unique_ptr<ServeInstance> in;
    try {
        in = std::make_unique<ServeInstance>();
        in->f = std::make_unique<Bildlt>(P.pat);
        in->f->upload_pattern(s->st);
        in->f->set_solve(P.algo, P.sweeps);
        in->f->use_graphs = o_.graphs != 0;
        in->f->solve_levels = o_.graphs != 0;
        in->op.kkt = false;
        in->op.m = in->op.n = P.N;
        in->op.A.m = in->op.A.n = P.N;
        in->op.A.nnz = P.nnz;
        in->op.A.row_ptr = P.k_row.as<int32_t>();
        in->op.A.col = P.k_col.as<int32_t>();
        in->op.A.val = in->f->d_vals.as<double>();   // the factor's copy of K's values
        in->q.init(&in->op, in->f.get(), o_.sqmr);
        if(o_.graphs) in->q.graphs = std::make_unique<SqmrGraphs>();
        in->rhs.ensure(sizeof(double) * (size_t) std::max(1, P.N));
        in->x.ensure(sizeof(double) * (size_t) std::max(1, P.N));
        if(o_.equilibrate_sweeps > 0) {
            in->scale.ensure(sizeof(double) * (size_t) std::max(1, P.N));
            in->scale_tmp.ensure(sizeof(double) * (size_t) std::max(1, P.N));
        }
    } catch(...) {
        lk.lock();
        used_ -= need;
        --creating_;
        cv_mem_.notify_all();
        throw;
    }
    lk.lock();
    --creating_;
    in->bytes = need;
    in->last_use = ++tick_;
    in->busy = true;
    ServeInstance* raw = in.get();
    s->cache[pattern] = std::move(in);
    ++instances_;
    return raw;
}

void Server::run(ServeInstance* in, ServeRequest* r, hipStream_t st)
{
    Bildlt& f = *in->f;
    const size_t vb = sizeof(double) * r->rhs.size();
    const int32_t N = (int32_t) r->rhs.size();
    if(o_.equilibrate_sweeps > 0) {
        // K's values into the factor's value buffer, scaled there to D K D
        // (the instance's SpMV reads the same buffer); b -> D b
        HIP_CHECK(hipMemcpyAsync(f.d_vals.p, r->vals.data(), sizeof(double) * r->vals.size(),
            hipMemcpyHostToDevice, st));
        equilibrate_csr(in->op.A, o_.equilibrate_sweeps, in->scale.as<double>(),
            in->scale_tmp.as<double>(), st);
        f.factor(f.d_vals.as<double>(), true, o_.threshold, st);
        HIP_CHECK(hipMemcpyAsync(in->rhs.p, r->rhs.data(), vb, hipMemcpyHostToDevice, st));
        scale_vector(N, in->scale.as<double>(), in->rhs.as<double>(), in->rhs.as<double>(), st);
        in->q.res_scale = in->scale.as<double>();
        in->q.run(in->rhs.as<double>(), in->x.as<double>(), r->res, st);
        scale_vector(N, in->scale.as<double>(), in->x.as<double>(), in->x.as<double>(), st);
    } else {
        f.factor(r->vals.data(), false, o_.threshold, st);
        HIP_CHECK(hipMemcpyAsync(in->rhs.p, r->rhs.data(), vb, hipMemcpyHostToDevice, st));
        in->q.run(in->rhs.as<double>(), in->x.as<double>(), r->res, st);
    }
    r->x.resize(r->rhs.size());
    HIP_CHECK(hipMemcpyAsync(r->x.data(), in->x.p, vb, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
}

void Server::worker(Slot* s)
{
    (void) hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    for(;;) {
        cv_work_.wait(lk, [this] { return stop_ || !queue_.empty(); });
        if(queue_.empty()) break;
        const int64_t id = queue_.front();
        queue_.pop_front();
        ServeRequest* r = reqs_[id].get();
        r->t_start = now();
        r->slot = s->index;
        ServeInstance* in = nullptr;
        try {
            in = instance(s, r->pattern, lk);
            lk.unlock();
            run(in, r, s->st);
            lk.lock();
            // the solve layout is sized by the first factorization: account
            // the instance's actual bytes
            const int64_t actual = in->f->numeric_bytes() + (int64_t) (in->q.work.bytes +
                in->rhs.bytes + in->x.bytes + in->scale.bytes + in->scale_tmp.bytes);
            used_ += actual - in->bytes;
            in->bytes = actual;
            peak_ = std::max(peak_, used_);
            ServePattern& P = *patterns_[r->pattern];
            P.inst_bytes = std::max(P.inst_bytes, actual);
        } catch(const Error& e) {
            if(!lk.owns_lock()) lk.lock();
            r->status = e.status;
            r->err = e.what();
        } catch(const std::exception& e) {
            if(!lk.owns_lock()) lk.lock();
            r->status = CULIP_ERR_INVALID;
            r->err = e.what();
        }
        if(in && r->status != CULIP_OK) {
            // a failed request may leave its instance in any state: the
            // next request of the pattern builds a fresh one
            auto it = s->cache.find(r->pattern);
            if(it != s->cache.end() && it->second.get() == in) {
                std::unique_ptr<ServeInstance> dead = std::move(it->second);
                s->cache.erase(it);
                used_ -= dead->bytes;
                --instances_;
                lk.unlock();
                dead.reset();
                lk.lock();
            }
            in = nullptr;
        }
        if(in) in->busy = false;
        r->t_end = now();
        r->done = true;
        ++completed_;
        cv_done_.notify_all();
        if(mem_waiters_ > 0) cv_mem_.notify_all();
    }
}

}  // namespace culip
