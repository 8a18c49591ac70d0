// Library objects behind the C-ABI handles.
#pragma once

#include <hip/hip_runtime_api.h>

#include <map>
#include <memory>
#include <string>
#include <type_traits>

#include "common.hpp"
#include "symbolic.hpp"
#include "../device/factor.hpp"
#include "../device/noalloc.hpp"
#include "../device/solve.hpp"
#include "../device/spmv.hpp"
#include "../device/sqmr.hpp"

namespace culip {

void hip_check(hipError_t e, const char* what);
#define HIP_CHECK(x) ::culip::hip_check((x), #x)

// Owning device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if(p) (void) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t b)
    {
        if(b <= bytes && p) return;
        release();
        const size_t want = b ? b : 16;
        HIP_CHECK(hipMalloc(&p, want));
        bytes = want;
    }
    template<typename T> T* as() const { return (T*) p; }
};

// One int32 of pinned, host-coherent memory that kernels can write and the
// host can read without a copy (error flags).
struct PinnedFlag {
    int32_t* p = nullptr;
    PinnedFlag() = default;
    PinnedFlag(const PinnedFlag&) = delete;
    PinnedFlag& operator=(const PinnedFlag&) = delete;
    ~PinnedFlag()
    {
        if(p) (void) hipHostFree(p);
    }
    void ensure()
    {
        if(p) return;
        HIP_CHECK(hipHostMalloc((void**) &p, sizeof(int32_t), hipHostMallocCoherent));
        *p = 0;
    }
    // the flag's value, cleared (0 when never allocated)
    int32_t take()
    {
        if(!p) return 0;
        const int32_t v = __atomic_exchange_n(p, 0, __ATOMIC_ACQ_REL);
        return v;
    }
};

// Runs `launch` (a single-launch triangular solve on st) if the process-wide
// admission of such solves allows it now; returns false, without calling
// launch, when TRSV_SINGLE_MAX_CONCURRENT other streams of the device still
// have one in flight or st is being captured. chain_guard.hpp explains the
// rule.
bool chain_admitted_launch(hipStream_t st, hipError_t (*launch)(void*), void* ctx);

template<typename F>
bool chain_admitted(hipStream_t st, F&& f)
{
    using Fn = std::remove_reference_t<F>;
    return chain_admitted_launch(
        st, [](void* c) -> hipError_t { return (*static_cast<Fn*>(c))(); }, (void*) &f);
}

template<typename T>
void upload(DevBuf& d, const std::vector<T>& h)
{
    d.ensure(sizeof(T) * h.size());
    if(!h.empty()) HIP_CHECK(hipMemcpy(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
}

// What SQMR needs of a preconditioner M = Pl D Pr (Preconditioner<T>,
// preconditioner.cuh:18-43): in-place left and right solves and D as a
// tridiagonal vector with pivot-start flags (NULL: identity).
struct Precond {
    virtual ~Precond() = default;
    virtual int32_t size() const = 0;
    virtual bool ready() const = 0;
    virtual void left(double* x, hipStream_t st) const = 0;    // x = Pl \ x
    virtual void right(double* x, hipStream_t st) const = 0;   // x = Pr \ x
    virtual const double* tri() const = 0;
    virtual const int32_t* is_piv() const = 0;
    // after a solve: report errors the device recorded (may synchronise st)
    virtual void after_solve(hipStream_t) const {}
    // x = Pl \ (K src) in one pass where the preconditioner can fold the
    // product into its left solve, with the same result as K src followed by
    // left(); false: not done (the caller multiplies, then calls left)
    virtual bool left_product(const DevCsr&, const double*, double*, hipStream_t) const
    {
        return false;
    }
};

// Per sparsity pattern: the matrix as given, its symbolic analysis and the
// static device structures derived from it. Shared by every numeric
// instance of the pattern (one per concurrent factorization).
struct BildltPattern {
    HostCsr A;
    std::vector<int32_t> piv_starts, block_starts;
    Symbolic sym;
    bool have_sym = false;
    bool uploaded = false;
    culip_prec prec = CULIP_PREC_F64;
    DevBuf d_bs, d_is_piv, d_bin, d_csr_off, d_csc_off, d_csc_idx, d_upd_off, d_upd_ki, d_upd_kj,
        d_lvlb_idx, d_block_in_lvl, d_lvl_rows, d_ix_in, d_value_slot, d_rfo;
    DevBuf d_lq_cap, d_lq_fr, d_off_task, d_upd4;
    void upload_static();
    int64_t device_bytes() const;
};

struct Bildlt final : Precond {
    std::shared_ptr<BildltPattern> pat;
    // the pattern's members under their own names
    HostCsr& A;
    std::vector<int32_t>& piv_starts;
    std::vector<int32_t>& block_starts;
    Symbolic& sym;
    bool& have_sym;
    DevBuf &d_bs, &d_is_piv, &d_bin, &d_csr_off, &d_csc_off, &d_csc_idx, &d_upd_off, &d_upd_ki,
        &d_upd_kj, &d_lvlb_idx, &d_block_in_lvl, &d_lvl_rows, &d_ix_in, &d_value_slot, &d_rfo;
    DevBuf &d_lq_cap, &d_lq_fr, &d_off_task, &d_upd4;

    culip_prec prec = CULIP_PREC_F64;
    bool uploaded = false;   // this instance's numeric buffers exist
    bool factored = false;
    double threshold = 0.0;
    // BlockiLDLt<T, bk, rook>: numeric pivoting inside pivot blocks
    culip_pivoting pivoting = CULIP_PIVOT_STATIC;
    // arithmetic of the 2x2 pivot steps and D solves of an FP32 factor
    culip_pivot_precision pivot_precision = CULIP_PIVOT2_F64;
    // this factor's pivot starts and row permutation (new -> old), and a
    // work vector for applying P in the solves
    DevBuf d_fis_piv, d_perm, d_perm_tmp;

    // numeric (per instance)
    DevBuf d_vals, d_val_in, d_bout, d_ix_out, d_val_out, d_tri, d_rn;
    DevBuf d_if_val, d_if_idx, d_if_cnt, d_rn_part, d_drop_cnt, d_overflow, d_piv_loc, d_carry;
    DevBuf d_keep, d_granted, d_chunk;
    DevBuf d_piv_umask, d_defer_list, d_defer_cnt, d_wave_list, d_wave_cnt, d_store_list,
        d_store_cnt;
    // solve layout
    DevBuf d_ptr[4], d_col[4], d_val[4], d_lr[4], d_task[2];
    DevBuf d_totals, d_scan_tmp;
    int32_t plan_tot[4] = {0, 0, 0, 0};   // entries of the uo, ui, to, ti rows
    bool plan_known = false;              // plan_tot read back since the factor
    // the numeric refactor captured as a HIP graph and replayed (explicit
    // streams only); captured again when the threshold changes
    bool use_graphs = false;
    // exact solves as level launches only: set where solves are captured
    // (SQMR segments of a server slot) and other threads capture too, so
    // that no capture depends on the single-launch admission, whose event
    // queries would touch streams other threads are capturing
    bool solve_levels = false;
    hipGraphExec_t factor_exec = nullptr;
    double graph_threshold = 0.0;
    // single-launch triangular solves: published values, and the half of
    // d_chain_pub the next such solve publishes into; h_trsv_err is set by
    // a solve whose dependency wait gave up (pinned: read without a copy)
    DevBuf d_chain_pub;
    mutable PinnedFlag h_trsv_err;
    culip_solve_algorithm solve_algo = CULIP_TRSV_SINGLE;
    int32_t jacobi_sweeps = 0;
    mutable int32_t chain_parity = 0;
    DevBuf d_jac[3];          // block-Jacobi right-hand side and iterates

    // host copies for inspection
    std::map<std::string, std::vector<int32_t>> host_int;
    std::map<std::string, std::vector<double>> host_real;

    Bildlt() : Bildlt(std::make_shared<BildltPattern>()) {}
    explicit Bildlt(std::shared_ptr<BildltPattern> p)
        : pat(p), A(p->A), piv_starts(p->piv_starts), block_starts(p->block_starts), sym(p->sym),
          have_sym(p->have_sym), d_bs(p->d_bs), d_is_piv(p->d_is_piv), d_bin(p->d_bin),
          d_csr_off(p->d_csr_off), d_csc_off(p->d_csc_off), d_csc_idx(p->d_csc_idx),
          d_upd_off(p->d_upd_off), d_upd_ki(p->d_upd_ki), d_upd_kj(p->d_upd_kj),
          d_lvlb_idx(p->d_lvlb_idx), d_block_in_lvl(p->d_block_in_lvl), d_lvl_rows(p->d_lvl_rows),
          d_ix_in(p->d_ix_in), d_value_slot(p->d_value_slot), d_rfo(p->d_rfo),
          d_lq_cap(p->d_lq_cap), d_lq_fr(p->d_lq_fr), d_off_task(p->d_off_task),
          d_upd4(p->d_upd4), prec(p->prec)
    {
    }
    Bildlt(const Bildlt&) = delete;
    Bildlt& operator=(const Bildlt&) = delete;
    ~Bildlt() override { release_graphs(); }

    int32_t n() const { return A.m; }
    bool f32() const { return prec == CULIP_PREC_F32; }
    size_t vt() const { return f32() ? sizeof(float) : sizeof(double); }

    void symbolic(int32_t fill_level, double fill_factor);
    // static structures of the pattern (once) and this instance's buffers
    void upload_pattern(hipStream_t st = nullptr);
    // device bytes of this instance's numeric buffers only
    int64_t numeric_bytes() const;
    // the same before allocation, from the symbolic analysis; the solve
    // layout (sized by the factor's entries) is bounded from above when
    // with_layout is set and left out otherwise
    int64_t numeric_bytes_estimate(bool with_layout) const;
    void factor(const double* vals, bool vals_on_device, double thr, hipStream_t st);
    // the end of a factor whose launches have completed (also after a
    // batched refactor): errors the device recorded, then the object is factored
    void after_factor();
    void solve(culip_side side, bool transpose, const double* b, double* x, hipStream_t st) const;
    // the same without allocating: exact solves as level launches,
    // block-Jacobi work vectors in the static scratch array (n must not
    // exceed NOALLOC_MAX_ROWS)
    void solve_noalloc(culip_side side, bool transpose, const double* b, double* x,
        hipStream_t st) const;
    // x <- P (U \ x) and x <- U' \ (P' x); the _unpermuted forms leave P out
    void trsv_u(double* x, hipStream_t st) const;
    void trsv_ut(double* x, hipStream_t st) const;
    void trsv_u_unpermuted(double* x, hipStream_t st) const;
    void trsv_ut_unpermuted(double* x, hipStream_t st) const;
    void jacobi(double* x, bool transpose, hipStream_t st) const;
    // set_solve_algorithm; allocates the block-Jacobi vectors when needed
    void set_solve(culip_solve_algorithm a, int32_t sweeps);
    // fails if a single-launch solve gave up waiting (synchronises st);
    // the flag is cleared, so the failure is reported once
    void check_solve_error(hipStream_t st) const;
    // the same without synchronising: reports what solves that already
    // completed recorded (entry of every solve, factor and SQMR call)
    void take_async_error() const;
    SolveTopo topo() const;
    int64_t device_bytes() const;
    void download_factor();
    // out_U.mtx, out_D.mtx, out_P.mtx under folder; returns the achieved
    // fill factor (2 nnz(U) + nnz(D)) / nnz(A)
    double export_factorized(const std::string& folder);

    template<typename VT> FactorArgs<VT> factor_args() const;
    template<typename VT> SolvePlan<VT> plan() const;
    template<typename VT> void factor_impl(hipStream_t st);
    // launches the refactor as a replay of its captured graph (capturing it
    // first); false when graphs are off or st cannot be captured
    bool factor_graph(hipStream_t st);
    void release_graphs();
    // entries of the four solve-layout rows (reads them back once)
    const int32_t* plan_sizes();

    // M = U' D U: left U^-T, right U^-1
    int32_t size() const override { return n(); }
    bool ready() const override { return factored; }
    void left(double* x, hipStream_t st) const override { trsv_ut(x, st); }
    void right(double* x, hipStream_t st) const override { trsv_u(x, st); }
    bool left_product(const DevCsr& K, const double* src, double* x, hipStream_t st) const override;
    const double* tri() const override { return d_tri.as<double>(); }
    const int32_t* is_piv() const override { return d_fis_piv.as<int32_t>(); }
    void after_solve(hipStream_t st) const override { check_solve_error(st); }
    // x <- x(perm) before a left solve, x(perm) <- x after a right solve
    // (k_permute_Pt / k_permute_P); nothing for static pivots
    void perm_pull(double* x, hipStream_t st) const;
    void perm_push(double* x, hipStream_t st) const;
};

// ImportLDUPreconditioner (import_ldu_preconditioner.cuh:36-125): an
// imported factor M = inv(S_r) P_r' L D R P_c' inv(S_c) with R = L' (the
// symmetric constructors) or U, each triangular operand resident with its
// level schedule; the transposed operands are stored explicitly so every
// substitution reads rows.
struct Ldu final : Precond {
    struct Operand {
        int32_t levels = 0;
        bool lower = true;
        std::vector<int32_t> lvl_off;   // host: levels + 1
        DevBuf row_ptr, col, val, diag_pos, order;
    };
    int32_t n = 0;
    culip_prec prec = CULIP_PREC_F64;
    bool symmetric = true, has_d = false, exact = true;
    Operand L, Lt, U, Ut;
    DevBuf d_inv_l, d_inv_u;            // approximate solves: 1 / diagonal
    DevBuf d_tri, d_is_piv;
    DevBuf d_pr, d_pc, d_sr, d_sc;      // permutation / scaling (optional)
    bool has_p = false, has_s = false;
    DevBuf d_t1, d_t2;                  // work vectors (solves are not reentrant)
    static constexpr int32_t APPROX_SWEEPS = 10;   // import_ldu_preconditioner.impl.cuh:223

    void create(const HostCsr& l, const HostCsr* u, const double* tri, const int32_t* is_piv,
        bool exact_solve, culip_prec p);
    void set_permutation(const int32_t* pr, const int32_t* pc);
    void set_scaling(const double* sr, const double* sc);
    void solve(culip_side side, bool transpose, const double* b, double* x, hipStream_t st) const;

    int32_t size() const override { return n; }
    bool ready() const override { return true; }
    void left(double* x, hipStream_t st) const override { solve(CULIP_SOLVE_LEFT, false, x, x, st); }
    void right(double* x, hipStream_t st) const override
    {
        solve(CULIP_SOLVE_RIGHT, false, x, x, st);
    }
    const double* tri() const override { return has_d ? d_tri.as<double>() : nullptr; }
    const int32_t* is_piv() const override { return has_d ? d_is_piv.as<int32_t>() : nullptr; }
    int64_t device_bytes() const;

private:
    void tri_apply(const Operand& o, const DevBuf& inv, const double* in, double* out,
        hipStream_t st) const;
    void ps(bool left_side, bool push, const double* b, double* x, hipStream_t st) const;
};

struct Spmv {
    bool kkt = false;
    bool normal = false;
    DevNormal N;
    DevBuf b_t;             // normal equations: work vector (not reentrant)
    DevCsr A, At;           // At only for a non-symmetric CSR transpose product
    bool have_at = false;
    DevKkt K;
    DevBuf b_row, b_col, b_val, bt_row, bt_col, bt_val, b_d, b_dix, b_arow, b_acol, b_aval,
        b_atrow, b_atcol, b_atval, b_hrow, b_hcol, b_hval;
    int32_t m = 0, n = 0;
    DevBuf b_eq;            // equilibration work vector
    void multiply(const double* x, double* y, double alpha, double beta, bool transpose,
        hipStream_t st) const;
    // CSR operators only: values <- D A D, d <- diag(D) (device)
    void equilibrate(int32_t sweeps, double* d, hipStream_t st);
};

// device copy of a host CSR matrix (row_ptr, col, val buffers)
void upload_csr(DevCsr& d, DevBuf& br, DevBuf& bc, DevBuf& bv, const HostCsr& h);

struct Sqmr {
    const Spmv* A = nullptr;
    const Precond* prec = nullptr;
    culip_sqmr_opts opts{};
    // true residuals measured on the unscaled system (device, borrowed;
    // NULL: as given)
    const double* res_scale = nullptr;
    // replay the iterations between true-residual checks from captured
    // graphs (set by owners that keep operator and preconditioner fixed)
    std::unique_ptr<SqmrGraphs> graphs;
    DevBuf work;
    int32_t* host_flags = nullptr;
    double* host_scal = nullptr;
    SqmrWork w;
    Sqmr() = default;
    Sqmr(const Sqmr&) = delete;
    Sqmr& operator=(const Sqmr&) = delete;
    ~Sqmr();
    // binds the operator and preconditioner (borrowed) and allocates the
    // work vectors
    void init(const Spmv* op, const Precond* p, const culip_sqmr_opts& o);
    // SQMR::solve on device vectors; synchronises st
    void run(const double* b, double* x, culip_sqmr_result& res, hipStream_t st);
};

}  // namespace culip
