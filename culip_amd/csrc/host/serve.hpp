// Request loop for streams of KKT solves whose sparsity patterns are fixed
// and whose values change (interior-point serving), and for batches of
// independent systems.
//
// Patterns are registered once: their symbolic analysis and static device
// structures stay resident. Requests (new values of K, right-hand side) are
// queued and served by `slots` worker threads, each with its own HIP stream
// and its own numeric instances (factor storage, solve layout, SQMR work),
// so requests that arrive while others run start as soon as a slot and
// device memory are free, and their kernels overlap on the device. A slot
// keeps the instances of the patterns it served while the memory budget
// allows; under pressure idle instances are released, least recently used
// first. Every request runs the same kernels in the same order whatever
// else is in flight, so its result does not depend on the mix.
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "objects.hpp"

namespace culip {

struct ServePattern {
    std::shared_ptr<BildltPattern> pat;
    DevBuf k_row, k_col;      // CSR structure of K for the SpMV
    int32_t N = 0;
    int64_t nnz = 0;
    int64_t static_bytes = 0;
    int64_t inst_bytes = 0;   // measured on the first instance
    int64_t est_max = 0;      // before that: upper bound of an instance
    int64_t est_min = 0;      // and its part without the solve layout
    culip_solve_algorithm algo = CULIP_TRSV_LEVELS;
    int32_t sweeps = 0;
    int64_t submitted = 0;    // requests queued for it so far
};

struct ServeInstance {
    std::unique_ptr<Bildlt> f;
    Spmv op;
    Sqmr q;
    DevBuf rhs, x;
    DevBuf scale, scale_tmp;  // equilibration: diag(D) and sweep work
    int64_t bytes = 0;
    uint64_t last_use = 0;
    bool busy = false;        // a request is running on it
};

struct ServeRequest {
    int64_t id = 0;
    int32_t pattern = 0;
    std::vector<double> vals, rhs, x;
    culip_sqmr_result res{};
    double t_submit = 0, t_start = 0, t_end = 0;
    int32_t slot = -1;
    culip_status status = CULIP_OK;
    std::string err;
    bool done = false;
};

class Server {
public:
    explicit Server(const culip_server_opts& o);
    ~Server();
    int32_t add_pattern(const HostCsr& K, const std::vector<int32_t>& piv_starts,
        const std::vector<int32_t>& block_starts);
    int64_t submit(int32_t pattern, const double* vals, const double* rhs);
    void set_pattern_solve(int32_t pattern, culip_solve_algorithm a, int32_t sweeps);
    // blocks until the request is done; copies x (N doubles) if x != nullptr
    void wait(int64_t id, double* x, culip_sqmr_result* res, double* timing, int32_t* slot);
    void stats(int64_t* out) const;
    int32_t pattern_n(int32_t p) const;

private:
    struct Slot {
        int32_t index = 0;
        hipStream_t st = nullptr;
        std::thread th;
        std::map<int32_t, std::unique_ptr<ServeInstance>> cache;
    };

    void worker(Slot* s);
    ServeInstance* instance(Slot* s, int32_t pattern, std::unique_lock<std::mutex>& lk);
    void run(ServeInstance* in, ServeRequest* r, hipStream_t st);
    double now() const;
    void estimate(ServePattern& P) const;

    culip_server_opts o_;
    int device_ = 0;
    mutable std::mutex mu_;
    std::condition_variable cv_work_, cv_done_, cv_mem_;
    std::vector<std::unique_ptr<ServePattern>> patterns_;
    std::deque<int64_t> queue_;
    std::unordered_map<int64_t, std::unique_ptr<ServeRequest>> reqs_;
    int64_t next_id_ = 1;
    int64_t budget_ = 0, used_ = 0, peak_ = 0, completed_ = 0, instances_ = 0;
    int32_t mem_waiters_ = 0;
    int32_t creating_ = 0;     // instances being allocated outside the lock
    uint64_t tick_ = 0;
    bool stop_ = false;
    std::vector<std::unique_ptr<Slot>> slots_;
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace culip
