/*
 * C++ classes over the C-ABI of culip_amd.h with the names, argument
 * meaning and call order of the reference's C++ interface for the BILDLt +
 * SQMR path, so a caller of the reference changes types and includes, not
 * its call sequence:
 *
 *   BlockiLDLt<T>     libs/la/preconditioner/block_ildlt.cuh:188-311
 *   CSRMatrixSPMV     libs/la/spmv.cuh:36-66
 *   AugmentedMatrixSPMV libs/la/spmv.cuh:116-149
 *   NormalMatrixSPMV  libs/la/spmv.cuh:74-107
 *   SQMR              libs/la/iterative_solvers/sqmr.cuh:21-72
 *   ImportLDUPreconditioner libs/la/preconditioner/import_ldu_preconditioner.cuh:36-125
 *
 * Differences: matrices are culip_csr views of host arrays (the reference's
 * csr_matrix_t on the host), vectors are device arrays of doubles whatever T
 * is (T selects the factor's storage and arithmetic), the GPU handle is a
 * HIP stream, and failures throw culip_amd::Error instead of raising
 * SIGSEGV (libs/utils/types.cuh:32).
 */
#pragma once

#include <stdexcept>
#include <string>
#include <type_traits>

#include "culip_amd.h"

namespace culip_amd {

struct Error : std::runtime_error {
    culip_status status;
    Error(culip_status s, const char* what) : std::runtime_error(what), status(s) {}
};

inline void check(culip_status s)
{
    if(s != CULIP_OK) throw Error(s, culip_last_error());
}

/* block_ildlt.cuh:24-30, same names and values. */
enum SOLVE_ALGORITHM { SOLVE_STRIPE = 0, SOLVE_BLOCK = 1, SOLVE_SCALAR = 2, SOLVE_JACOBI = 3 };

/* Preconditioner<T> with solve_left/middle/right (preconditioner.cuh).
 * BK / ROOK: BlockiLDLt<T, bk, rook> (block_ildlt_instances.cu:54-60),
 * numeric Bunch-Kaufman or rook pivoting inside every pivot block. */
template<typename T, bool BK = false, bool ROOK = false>
class BlockiLDLt {
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
        "BlockiLDLt<T>: T is double or float");
    static_assert(!(BK && ROOK), "BlockiLDLt<T, bk, rook>: one pivoting rule at most");

public:
    /* block_ildlt.impl.cuh:44-79 */
    BlockiLDLt(void* stream, const culip_csr* h_A, int32_t num_pivs, const int32_t* piv_starts,
        int32_t num_blocks, const int32_t* block_starts)
        : stream_(stream)
    {
        check(culip_bildlt_create(h_A, num_pivs, piv_starts, num_blocks, block_starts,
            std::is_same<T, float>::value ? CULIP_PREC_F32 : CULIP_PREC_F64, &h_));
        if(BK || ROOK)
            check(culip_bildlt_set_pivoting(h_,
                BK ? CULIP_PIVOT_BUNCH_KAUFMAN : CULIP_PIVOT_ROOK));
    }
    ~BlockiLDLt() { culip_bildlt_destroy(h_); }
    BlockiLDLt(const BlockiLDLt&) = delete;
    BlockiLDLt& operator=(const BlockiLDLt&) = delete;

    int32_t n() const { return culip_bildlt_n(h_); }

    /* block_ildlt.impl.cuh:103-146; includes solve_analysis(). The
     * reference returns a Manteuffel shift that its static path leaves 0. */
    T compute(int32_t fill_level, T fill_factor, T threshold)
    {
        check(culip_bildlt_compute(h_, fill_level, fill_factor, threshold, stream_));
        return T(0);
    }

    /* New values of the same pattern (nnz doubles in h_A's entry order). */
    void refactor(const double* vals, bool vals_on_device, T threshold)
    {
        check(culip_bildlt_refactor(h_, vals, vals_on_device ? 1 : 0, threshold, stream_));
    }

    /* solve_analysis() is part of compute(); kept so callers compile. */
    void solve_analysis() {}

    /* block_ildlt.impl.cuh:1700-1916; returns the achieved fill factor. */
    double export_factorized(const char* folder)
    {
        double f = 0.0;
        check(culip_bildlt_export_factorized(h_, folder, &f));
        return f;
    }

    /* block_ildlt.impl.cuh:253-260. STRIPE, BLOCK and SCALAR all apply
     * the exact U^-1 (resp. U^-T); here they share the level-scheduled
     * solve. JACOBI runs num_jacobi_sweeps block-Jacobi sweeps from x = 0. */
    void set_solve_algorithm(SOLVE_ALGORITHM algorithm, int32_t num_jacobi_sweeps)
    {
        check(culip_bildlt_set_solve_algorithm(h_,
            algorithm == SOLVE_JACOBI ? CULIP_TRSV_JACOBI : CULIP_TRSV_SINGLE,
            num_jacobi_sweeps));
    }

    /* FP32 factor: 2x2 pivot arithmetic in FP64 (default) or in FP32 as
     * the reference's BlockiLDLt<float> */
    void set_pivot_precision(culip_pivot_precision p)
    {
        check(culip_bildlt_set_pivot_precision(h_, p));
    }

    bool is_left() const { return true; }
    bool is_middle() const { return true; }
    bool is_right() const { return true; }

    /* block_ildlt.impl.cuh:183-246; b, x: n device doubles. */
    void solve_left(const double* b, double* x, bool transpose = false) const
    {
        check(culip_bildlt_solve(h_, CULIP_SOLVE_LEFT, transpose, b, x, stream_));
    }
    void solve_middle(const double* b, double* x, bool transpose = false) const
    {
        check(culip_bildlt_solve(h_, CULIP_SOLVE_MIDDLE, transpose, b, x, stream_));
    }
    void solve_right(const double* b, double* x, bool transpose = false) const
    {
        check(culip_bildlt_solve(h_, CULIP_SOLVE_RIGHT, transpose, b, x, stream_));
    }

    const culip_bildlt* handle() const { return h_; }
    culip_bildlt* handle() { return h_; }

private:
    void* stream_;
    culip_bildlt* h_ = nullptr;
};

/* ImportLDUPreconditioner<T> (import_ldu_preconditioner.cuh:36-125). The
 * three ctors: (L, U), (L) and (L, D, D_is_piv); D is the tridiagonal
 * vector (3n) and its pivot-start flags, host arrays. */
template<typename T>
class ImportLDUPreconditioner {
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
        "ImportLDUPreconditioner<T>: T is double or float");

public:
    ImportLDUPreconditioner(void* stream, const culip_csr* L, const culip_csr* U,
        bool exact_solve = true)
        : stream_(stream)
    {
        check(culip_ldu_create(L, U, nullptr, nullptr, exact_solve ? 1 : 0, prec(), &h_));
    }
    ImportLDUPreconditioner(void* stream, const culip_csr* L, bool exact_solve = true)
        : stream_(stream)
    {
        check(culip_ldu_create(L, nullptr, nullptr, nullptr, exact_solve ? 1 : 0, prec(), &h_));
    }
    ImportLDUPreconditioner(void* stream, const culip_csr* L, const double* D,
        const int32_t* D_is_piv, bool exact_solve = true)
        : stream_(stream)
    {
        check(culip_ldu_create(L, nullptr, D, D_is_piv, exact_solve ? 1 : 0, prec(), &h_));
    }
    ~ImportLDUPreconditioner() { culip_ldu_destroy(h_); }
    ImportLDUPreconditioner(const ImportLDUPreconditioner&) = delete;
    ImportLDUPreconditioner& operator=(const ImportLDUPreconditioner&) = delete;

    /* push convention: entry i moves to P[i] (host arrays) */
    void set_permutation(const int32_t* P_r, const int32_t* P_c)
    {
        check(culip_ldu_set_permutation(h_, P_r, P_c));
    }
    void set_scaling(const double* S_r, const double* S_c)
    {
        check(culip_ldu_set_scaling(h_, S_r, S_c));
    }

    int32_t n() const { return culip_ldu_n(h_); }
    bool is_left() const { return true; }
    bool is_middle() const { return culip_ldu_is_middle(h_) != 0; }
    bool is_right() const { return true; }

    void solve_left(const double* b, double* x, bool transpose = false) const
    {
        check(culip_ldu_solve(h_, CULIP_SOLVE_LEFT, transpose, b, x, stream_));
    }
    void solve_middle(const double* b, double* x, bool transpose = false) const
    {
        check(culip_ldu_solve(h_, CULIP_SOLVE_MIDDLE, transpose, b, x, stream_));
    }
    void solve_right(const double* b, double* x, bool transpose = false) const
    {
        check(culip_ldu_solve(h_, CULIP_SOLVE_RIGHT, transpose, b, x, stream_));
    }

    const culip_ldu* handle() const { return h_; }

private:
    static culip_prec prec() { return std::is_same<T, float>::value ? CULIP_PREC_F32 : CULIP_PREC_F64; }
    void* stream_;
    culip_ldu* h_ = nullptr;
};

/* SPMV<T> (spmv.cuh:18-31): y = alpha op(A) x + beta y on device doubles. */
class SPMV {
public:
    virtual ~SPMV() { culip_spmv_destroy(op_); }
    SPMV(const SPMV&) = delete;
    SPMV& operator=(const SPMV&) = delete;
    int32_t m() const { return culip_spmv_m(op_); }
    int32_t n() const { return culip_spmv_n(op_); }
    void multiply(const double* x, double* y, double alpha = 1.0, double beta = 0.0,
        bool transpose = false) const
    {
        check(culip_spmv_multiply(op_, x, y, alpha, beta, transpose, stream_));
    }
    const culip_spmv* handle() const { return op_; }

protected:
    explicit SPMV(void* stream) : stream_(stream) {}
    void* stream_;
    culip_spmv* op_ = nullptr;
};

/* CSRMatrixSPMV (spmv.cuh:36-66); copies the host CSR to the device. */
class CSRMatrixSPMV : public SPMV {
public:
    CSRMatrixSPMV(void* stream, const culip_csr* h_A) : SPMV(stream)
    {
        check(culip_spmv_create_csr(h_A, &op_));
    }
};

/* AugmentedMatrixSPMV (spmv.cuh:116-149): [diag(d) A'; A delta I(ix)]. */
class AugmentedMatrixSPMV : public SPMV {
public:
    AugmentedMatrixSPMV(void* stream, const culip_csr* h_A, const double* d, double delta,
        int32_t n_ix = 0, const int32_t* ix = nullptr)
        : SPMV(stream)
    {
        check(culip_spmv_create_kkt(h_A, d, delta, n_ix, ix, &op_));
    }
};

/* NormalMatrixSPMV (spmv.cuh:74-107): A diag(d)^-1 A' + delta I(ix). */
class NormalMatrixSPMV : public SPMV {
public:
    NormalMatrixSPMV(void* stream, const culip_csr* h_A, const double* d, double delta = 0.0,
        int32_t n_ix = 0, const int32_t* ix = nullptr)
        : SPMV(stream)
    {
        check(culip_spmv_create_normal(h_A, d, delta, n_ix, ix, &op_));
    }
};

/* SQMR<T> (sqmr.cuh:21-72). The reference's ctor takes (spmv, tol, max_it,
 * preconditioner); true_residual_period = 0 keeps its stopping test. */
class SQMR {
public:
    template<typename T, bool BK, bool ROOK>
    SQMR(void* stream, const SPMV* spmv, double tol, int32_t max_it,
        const BlockiLDLt<T, BK, ROOK>* prec, int32_t true_residual_period = 0,
        int32_t max_restarts = 0)
        : stream_(stream)
    {
        culip_sqmr_opts o{tol, max_it, true_residual_period, max_restarts, 0};
        check(culip_sqmr_create(spmv->handle(), prec->handle(), &o, &s_));
    }
    SQMR(void* stream, const SPMV* spmv, double tol, int32_t max_it)
        : stream_(stream)
    {
        culip_sqmr_opts o{tol, max_it, 0, 0, 0};
        check(culip_sqmr_create(spmv->handle(), nullptr, &o, &s_));
    }
    template<typename T>
    SQMR(void* stream, const SPMV* spmv, double tol, int32_t max_it,
        const ImportLDUPreconditioner<T>* prec, int32_t true_residual_period = 0,
        int32_t max_restarts = 0)
        : stream_(stream)
    {
        culip_sqmr_opts o{tol, max_it, true_residual_period, max_restarts, 0};
        if(prec) check(culip_sqmr_create_ldu(spmv->handle(), prec->handle(), &o, &s_));
        else check(culip_sqmr_create(spmv->handle(), nullptr, &o, &s_));
    }
    ~SQMR() { culip_sqmr_destroy(s_); }
    SQMR(const SQMR&) = delete;
    SQMR& operator=(const SQMR&) = delete;

    /* sqmr.impl.cuh:109-355: residual and iterations as the reference
     * reports them; the full result (true residual, restarts) via last(). */
    void solve(const double* b, double* x, double& residual, int32_t& iterations)
    {
        check(culip_sqmr_solve(s_, b, x, &last_, stream_));
        residual = last_.residual;
        iterations = last_.iterations;
    }
    const culip_sqmr_result& last() const { return last_; }

private:
    void* stream_;
    culip_sqmr* s_ = nullptr;
    culip_sqmr_result last_{};
};

/* Request loop (no reference counterpart): resident patterns, requests with
 * new values of K and a right-hand side, concurrent solves on `slots`
 * streams, admission by device-memory budget. Host arrays throughout. */
class Server {
public:
    explicit Server(const culip_server_opts& o) { check(culip_server_create(&o, &s_)); }
    ~Server() { culip_server_destroy(s_); }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    int32_t add_pattern(const culip_csr* K, int32_t num_pivs, const int32_t* piv_starts,
        int32_t num_blocks, const int32_t* block_starts)
    {
        int32_t p = -1;
        check(culip_server_add_pattern(s_, K, num_pivs, piv_starts, num_blocks, block_starts, &p));
        return p;
    }
    /* before the pattern's first submit */
    void set_pattern_solve(int32_t pattern, culip_solve_algorithm a, int32_t jacobi_sweeps)
    {
        check(culip_server_set_pattern_solve(s_, pattern, a, jacobi_sweeps));
    }
    int64_t submit(int32_t pattern, const double* vals, const double* rhs)
    {
        int64_t r = -1;
        check(culip_server_submit(s_, pattern, vals, rhs, &r));
        return r;
    }
    /* timing: {queued, service, latency, submit time} seconds (may be null) */
    culip_sqmr_result wait(int64_t request, double* x, double* timing = nullptr)
    {
        culip_sqmr_result r{};
        check(culip_server_wait(s_, request, x, &r, timing, nullptr));
        return r;
    }

private:
    culip_server* s_ = nullptr;
};

}  // namespace culip_amd
