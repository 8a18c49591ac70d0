// SQMR (libs/la/iterative_solvers/sqmr.impl.cuh:109-355) on the device.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>

#include "culip_amd.h"

namespace culip {

struct SqmrOps {
    int32_t n = 0;
    std::function<void(const double* x, double* y, void* stream)> spmv;   // y = A x
    std::function<void(double* x, void* stream)> left;                    // x = Pl \ x (NULL: I)
    std::function<void(double* x, void* stream)> right;                   // x = Pr \ x (NULL: I)
    // y = Pl \ (A x) in one pass, the same values as spmv then left; returns
    // false when not available (NULL: never)
    std::function<bool(const double* x, double* y, void* stream)> spmv_left;
    const double* tri = nullptr;      // middle D (NULL: identity)
    const int32_t* is_piv = nullptr;
    // true residuals measured as ||(b - A x) ./ s|| / ||b ./ s|| (NULL: s = 1):
    // with A = D K D, b = D b0 and s = diag(D) that is K's own residual
    const double* res_scale = nullptr;
};

// Device work space of one solve (allocated by the caller of sqmr_solve).
struct SqmrWork {
    int32_t n = 0, grid = 0;
    double* v[3] = {nullptr, nullptr, nullptr};
    double* w[3] = {nullptr, nullptr, nullptr};
    double* p[3] = {nullptr, nullptr, nullptr};
    double* z = nullptr;        // D \ v_k
    double* tv = nullptr;       // unscaled v_{k+1}
    double* ev = nullptr;       // preconditioner work
    double* av[2] = {nullptr, nullptr};
    double* xt = nullptr;       // true solution / residual work
    double* rt = nullptr;
    double* xacc = nullptr;     // refinement: accumulated solution
    double* rhs = nullptr;      // refinement: current residual
    double* part = nullptr;     // 8 * grid partial sums
    double* scal = nullptr;     // 32 scalars
    int32_t* flags = nullptr;   // 4 ints: test[0], test[1]
    int32_t* host_flags = nullptr;  // pinned
    double* host_scal = nullptr;    // pinned, 32 doubles
};

// scalar slots of SqmrWork::scal the host reads: the recurrence's gamma and
// ||Pl \ b||; [40] and [41] hold the norms of a residual check
constexpr int SQMR_S_GAMMA = 20, SQMR_S_NORMPB = 21;
// the true-residual gate's restart rules: a run ends when the best true
// residual has not halved for this many checks; Lanczos breakdown is
// |delta| <= SQMR_BREAKDOWN_REL * ||A||est
constexpr int SQMR_STALL_PERIODS = 25;
constexpr double SQMR_BREAKDOWN_REL = 1e-14;

int32_t sqmr_grid(int32_t n);
size_t sqmr_work_bytes(int32_t n);
void sqmr_work_bind(SqmrWork& w, int32_t n, char* dev_base, int32_t* host_flags,
    double* host_scal);

// Captured graphs of the iterations between two true-residual checks, one
// per phase (iteration mod 6). Valid while the operator, the
// preconditioner's buffers and solve schedule, and the work space stay the
// same.
struct SqmrGraphs {
    int32_t period = 0;
    hipGraphExec_t seg[6] = {};
    SqmrGraphs() = default;
    SqmrGraphs(const SqmrGraphs&) = delete;
    SqmrGraphs& operator=(const SqmrGraphs&) = delete;
    ~SqmrGraphs() { release(); }
    void release();
};

// Solves A x = b. b, x device vectors. Throws Error on HIP failures.
// graphs (may be null): replay the iterations between true-residual checks
// from captured graphs (explicit streams only).
void sqmr_solve(const SqmrOps& ops, SqmrWork& w, const double* b, double* x,
    const culip_sqmr_opts& opt, culip_sqmr_result& res, void* stream,
    SqmrGraphs* graphs = nullptr);

// ||b - A x|| / ||b|| in FP64 (synchronises the stream).
double true_rel_residual(const SqmrOps& ops, SqmrWork& w, const double* b, const double* x,
    void* stream);

}  // namespace culip
