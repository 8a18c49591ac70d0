// A batch of independent KKT systems of uneven size, factored and solved
// together: every step of the refactor (a level of all systems) and of the
// SQMR recurrence is one launch over all systems that take part in it
// (device/batch.hpp, factor_batch_level). Each system follows exactly the control flow
// of the one-system SQMR with the FP64 true-residual gate and refinement
// restarts (sqmr_solve), and ends with the bits it ends with when it is
// solved alone with the same options and block-Jacobi schedule.
#pragma once

#include <memory>
#include <mutex>
#include <vector>

#include "objects.hpp"
#include "../device/batch.hpp"

namespace culip {

struct BatchSolverOpts {
    culip_prec prec = CULIP_PREC_F32;
    int32_t fill_level = 0;
    double fill_factor = 1.0;
    double threshold = 1e-4;
    culip_sqmr_opts sqmr{1e-8, 1000, 20, 10};
    int32_t jacobi_sweeps = 2;   // block-Jacobi sweeps per triangular solve
};

class BatchSolver {
public:
    explicit BatchSolver(const BatchSolverOpts& o);
    ~BatchSolver();
    BatchSolver(const BatchSolver&) = delete;
    BatchSolver& operator=(const BatchSolver&) = delete;

    // registers one system (symbolic analysis, static upload); safe to call
    // from several threads at once; returns its index
    int32_t add(const HostCsr& K, const std::vector<int32_t>& piv_starts,
        const std::vector<int32_t>& block_starts);
    int32_t count() const;
    int32_t system_n(int32_t s) const;
    int64_t system_nnz(int32_t s) const;
    // refactors system s with vals[s] and solves K_s x_s = rhs_s for every
    // system (host arrays, or device arrays when device_arrays); res may be
    // null
    void solve(const double* const* vals, const double* const* rhs, double* const* x,
        culip_sqmr_result* res, bool device_arrays, hipStream_t st);
    int64_t device_bytes() const;
    // wall seconds of the last solve: refactors, SQMR
    void last_seconds(double* out) const { out[0] = sec_[0]; out[1] = sec_[1]; }

private:
    struct Entry {
        std::shared_ptr<BildltPattern> pat;
        std::unique_ptr<Bildlt> f;
        DevBuf k_row, k_col;
        int32_t n = 0;
        int64_t nnz = 0;
    };
    // the device lists of one set of systems
    struct ListSlot {
        DevBuf ent, items[5];
        int4* h_ent = nullptr;   // pinned staging of the entries
        hipEvent_t ev = nullptr; // recorded after the staging's upload
        bool ev_used = false;
        BatchLists L;
        std::vector<int32_t> ids;
        bool valid = false;
        ~ListSlot();
    };
    // per-system state of the solve (the locals of sqmr_solve / sqmr_run)
    struct State {
        bool done = false;
        bool running = false;     // inside a run
        bool restart_run = false; // the run solves for rhs (a refinement)
        int32_t it = 0;           // iterations of the run
        int32_t run_max = 0;      // the run's iteration limit
        int32_t total_it = 0;
        bool gate_met = false;
        bool at_check = false;    // the run ended right after a true-residual check
        double gate_abs = -1.0;
        double best_nr = -1.0, gam_at_best = 0.0;
        int32_t stall = 0;
        double rr = 0.0;
        int32_t restarts = 0;
        culip_sqmr_result res{};
        culip_sqmr_result run_res{};
    };

    void finalize(hipStream_t st);
    void setup(hipStream_t st);
    template<typename VT> void build_factor();
    template<typename VT> void factor_batched(hipStream_t st);
    const BatchLists& lists(ListSlot& slot, const std::vector<int32_t>& ids, hipStream_t st);
    void start_runs(const BatchLists& L, hipStream_t st);
    void iteration(const BatchLists& L, hipStream_t st);
    void jacobi(const BatchLists& L, bool transpose, int32_t target, int32_t product_src,
        hipStream_t st);
    // x of each listed system = right(D \ xt); then its residual against
    // brun (or b) and the norms into scal[40..41]
    void run_solution(const BatchLists& L, hipStream_t st);
    void residual_norms(const BatchLists& L, int32_t bvec, hipStream_t st);
    void read_back(hipStream_t st);
    void upload_sys(hipStream_t st);
    // `period` iterations of the listed systems as a captured graph
    void capture_segment(const BatchLists& L, int32_t period, const std::vector<int32_t>& ids,
        hipStream_t st);

    BatchSolverOpts o_;
    int device_ = 0;
    mutable std::mutex mu_;
    std::vector<std::unique_ptr<Entry>> sys_;
    bool finalized_ = false;   // no more systems (set-up started)
    bool ready_ = false;       // set-up done
    DevBuf d_work_, d_scal_, d_flags_, d_iter_, d_sys_;
    std::vector<SqmrWork> work_;
    std::vector<BatchSys> h_sys_;
    double* h_scal_ = nullptr;      // pinned mirrors of the scalar and flag pools
    int32_t* h_flags_ = nullptr;
    BatchSys* h_sys_pinned_ = nullptr;
    // batched refactor: system tables, level ranges, work items
    DevBuf d_fsys_, d_psys_, d_lr_, d_fit_[7], d_chunks_, d_blk_, d_scan_, d_task_;
    std::vector<FactorBatchLevel> flevels_;
    int32_t max_h_ = 1;   // largest block height of the batch
    FactorItems chunks_, blk_, scan_, task_;
    ListSlot slot_[4];
    std::vector<State> st_;
    // the batch's own stream (graphs cannot be captured on the null stream)
    hipStream_t sq_ = nullptr;
    // a segment of `period` iterations captured for the set of running
    // systems seg_ids_ (replayed while that set runs whole segments)
    hipGraphExec_t seg_exec_ = nullptr;
    std::vector<int32_t> seg_ids_, seg_prev_;
    double sec_[2] = {0.0, 0.0};
};

}  // namespace culip
