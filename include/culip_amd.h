/*
 * culip_amd -- MI355X (gfx950) block incomplete LDL' (BILDLt) preconditioner
 * and SQMR solver for KKT systems [[H, A'], [A, -delta I]].
 *
 * Thin C-ABI: plain pointers and sizes, int32 indices (the reference's
 * mat_int_t, libs/utils/defines.h:56-62), explicit streams, status codes
 * instead of the reference's raise(SIGSEGV) (libs/utils/types.cuh:32).
 * Streams are hipStream_t passed as void* (NULL = default stream).
 * Device pointers are HIP device memory owned by the caller; every object
 * created here owns its own device storage.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository dthuerck/culip).
 */
#ifndef CULIP_AMD_H
#define CULIP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CULIP_OK = 0,
    CULIP_ERR_INVALID = 1,      /* bad argument / inconsistent structure */
    CULIP_ERR_HIP = 2,          /* HIP runtime error */
    CULIP_ERR_NOMEM = 3,        /* host or device allocation failed */
    CULIP_ERR_STATE = 4,        /* call out of order (e.g. solve before factor) */
    CULIP_ERR_NUMERIC = 5,      /* non-finite values / breakdown */
    CULIP_ERR_BUDGET = 6,       /* request exceeds the device memory budget */
    CULIP_ERR_UNSUPPORTED = 7,  /* option not supported by this build */
    CULIP_ERR_NO_DEVICE = 8     /* no HIP device available */
} culip_status;

typedef enum { CULIP_PREC_F64 = 0, CULIP_PREC_F32 = 1 } culip_prec;
typedef enum { CULIP_SOLVE_LEFT = 0, CULIP_SOLVE_MIDDLE = 1, CULIP_SOLVE_RIGHT = 2 } culip_side;
typedef enum { CULIP_ORDER_NATURAL = 0, CULIP_ORDER_BLOCK_COLOR = 1 } culip_ordering;

/* Message of the last failing call on this thread ("" if none). */
const char * culip_last_error(void);
const char * culip_version(void);

/* Host CSR view (borrowed pointers). */
typedef struct {
    int32_t m, n;
    int64_t nnz;
    const int32_t * row_ptr;
    const int32_t * col;
    const double * val;
} culip_csr;

/* ------------------------------------------------------------------------ */
/* Input preparation (host only). The reference reads the matrix, ordering,
 * blocking and pivot structure from MatrixMarket files
 * (apps/block_ildlt/block_ildlt_invariants.h:88-111); these entry points
 * produce them for the synthetic KKT workloads instead.                    */

typedef struct {
    int32_t n;                  /* primal size (H is n x n) */
    int32_t m;                  /* dual size (A is m x n) */
    int32_t h_offdiag_per_row;  /* average off-diagonals per row of H */
    int32_t a_nnz_per_col;      /* nonzeros per column of A */
    double h_diag_lo, h_diag_hi;/* H diagonal range */
    int32_t h_diag_log;         /* 1: log-uniform in [lo, hi]; 0: uniform */
    double h_offdiag_std;       /* std of H off-diagonal values */
    double delta;               /* K(2,2) = -delta I */
    uint64_t seed;
    /* Barrier term: barrier_hi > 0 adds w_i, log-uniform in
     * [barrier_lo, barrier_hi], to H_ii (H = H0 + X^-1 Z of an interior-
     * point method), drawn after the right-hand side. 0: none. */
    double barrier_lo, barrier_hi;
} culip_kkt_params;

typedef struct culip_kkt culip_kkt;
culip_status culip_kkt_generate(const culip_kkt_params * p, culip_kkt ** out);
/* Full symmetric CSR of K (sorted columns) and the N(0,1) right-hand side. */
culip_status culip_kkt_view(const culip_kkt * k, culip_csr * K, const double ** rhs);
/* Serving refactor input: fresh barrier weights, H_ii <- H0_ii + w_i with
 * w_i log-uniform in [w_lo, w_hi] drawn from `seed` (replacing any barrier
 * term of the generated system); writes the new values of K (same
 * pattern) into vals_out (nnz doubles). */
culip_status culip_kkt_barrier_values(const culip_kkt * k, double w_lo,
    double w_hi, uint64_t seed, double * vals_out);
void culip_kkt_destroy(culip_kkt * k);

/* Ordering, contiguous blocking (blocks <= max_block, split at n_primal) and
 * static 1x1 pivot starts; K_perm(i, j) = K(perm[i], perm[j]) as the
 * reference's host_permute with perm_is_old_to_new = false
 * (block_ildlt_invariants.h:29-69, sparse_la.impl.cuh:602-661). */
typedef struct culip_prep culip_prep;
culip_status culip_prep_create(const culip_csr * K, int32_t n_primal,
    int32_t max_block, culip_ordering ordering, culip_prep ** out);
culip_status culip_prep_view(const culip_prep * p, culip_csr * K_perm,
    const int32_t ** perm, int32_t * num_blocks, const int32_t ** block_starts,
    int32_t * num_pivs, const int32_t ** piv_starts);
/* Apply the prep permutation to a value array of K's pattern
 * (vals_K[nnz] -> vals_perm[nnz] in K_perm's entry order). */
culip_status culip_prep_permute_values(const culip_prep * p,
    const double * vals_K, double * vals_perm);
void culip_prep_destroy(culip_prep * p);

/* MatrixMarket (libs/test/test.impl.h:25-300 semantics: 1-based, symmetric
 * expansion, duplicates kept, per-row sort). */
typedef struct culip_host_matrix culip_host_matrix;
culip_status culip_mm_read_csr(const char * path, culip_host_matrix ** out);
culip_status culip_mm_read_vector(const char * path, int64_t * len, double ** vals);
culip_status culip_host_matrix_view(const culip_host_matrix * a, culip_csr * out);
void culip_host_matrix_destroy(culip_host_matrix * a);
void culip_host_free(void * p);

/* T_matrix_permute_scale (libs/la/sparse_la.impl.cuh:600-657), as the
 * reference driver's host_permute uses it (block_ildlt_invariants.h:29-69):
 * entry (r, c) of A moves to (P_r(r), P_c(c)) and is divided by
 * s_r[r] s_c[c] (scale_before_match) or s_r[P_r(r)] s_c[P_c(c)]; with
 * perm_is_old_to_new = 0 the vectors are new -> old, so B(i, j) =
 * A(p_r[i], p_c[j]) / scale. Scales may be NULL (1); rows of the result are
 * sorted by column. Host arrays in, a host matrix out. */
/* Maximum-product matching of a square A (WeightedBipartiteMatching::match
 * and get_scaling, libs/algorithms/matching.cuh:69-125): costs
 * c_ij = log2(max_k |a_kj|) - log2 |a_ij| (kept on a grid of 2^-20), a
 * perfect matching of least total cost with its dual variables u, v
 * (u_i + v_j <= c_ij, equality on the matching). row_scale r = 2^u and
 * col_scale c = 2^v / max_k |a_kj| bound |r_i a_ij c_j| by 1, with 1 on
 * the matching; sym_scale d = sqrt(r c) bounds |d_i a_ij d_j| by 1 for a
 * symmetric A. Every output may be NULL (n entries each). *feasible = 0
 * when A has no perfect matching (the outputs then hold a partial one).
 * Host only. */
culip_status culip_weighted_matching(const culip_csr * A, int32_t * row_match, double * u,
    double * v, double * row_scale, double * col_scale, double * sym_scale,
    double * objective, int32_t * feasible);
/* Structural rank of an m x n pattern (sprank, libs/algorithms/sprank.cuh:
 * 15-36) and a maximum matching: row_match[i] = column of row i or -1
 * (m entries, may be NULL). Host only. */
culip_status culip_structural_rank(const culip_csr * A, int32_t * rank, int32_t * row_match);
/* Lower bound on the rank of an m x n pattern by greedy peeling into a
 * triangular form (BTF::permute, libs/algorithms/permute_btf.h:28-42):
 * row_order (m) and col_order (n) receive 1-based orders with the peeled
 * upper-triangular part first, as the reference writes them (may be NULL).
 * Host only. */
culip_status culip_btf_rank_bound(const culip_csr * A, int32_t * bound, int32_t * row_order,
    int32_t * col_order);

culip_status culip_csr_permute_scale(const culip_csr * A, const int32_t * row_perm,
    const int32_t * col_perm, int32_t perm_is_old_to_new, const double * row_scale,
    const double * col_scale, int32_t scale_before_match, culip_host_matrix ** out);

/* ------------------------------------------------------------------------ */
/* BlockiLDLt (libs/la/preconditioner/block_ildlt.cuh:188-311).            */

typedef struct culip_bildlt culip_bildlt;

/* ctor (block_ildlt.impl.cuh:44-79): full symmetric host CSR with sorted
 * columns, num_pivs pivot starts, num_blocks block starts (no sentinels
 * needed). The matrix values are copied. prec selects BlockiLDLt<double>
 * or BlockiLDLt<float> semantics (storage and arithmetic of the factor). */
culip_status culip_bildlt_create(const culip_csr * hA, int32_t num_pivs,
    const int32_t * piv_starts, int32_t num_blocks,
    const int32_t * block_starts, culip_prec prec, culip_bildlt ** out);

/* compute(fill_level, fill_factor, threshold) (block_ildlt.impl.cuh:103-146):
 * symbolic analysis on the host (once per pattern), upload, numeric factor
 * on the device. Also performs solve_analysis() (:2050-2150). Returns after
 * the factor has finished on `stream`. */
culip_status culip_bildlt_compute(culip_bildlt * h, int32_t fill_level,
    double fill_factor, double threshold, void * stream);

/* Symbolic phase only (host): build_initial_coarse .. build_blocks and the
 * solve analysis. compute() calls it when it has not run yet. */
culip_status culip_bildlt_symbolic(culip_bildlt * h, int32_t fill_level,
    double fill_factor);

/* Numeric refactor for new values of the same pattern (interior-point
 * serving: barrier diagonals change). vals: nnz doubles in hA's entry order,
 * host (vals_on_device = 0) or device memory. No reference counterpart: the
 * reference re-runs compute() from scratch. */
culip_status culip_bildlt_refactor(culip_bildlt * h, const double * vals,
    int32_t vals_on_device, double threshold, void * stream);

/* solve_left / solve_middle / solve_right (block_ildlt.impl.cuh:183-246):
 * left x = U' \ (P' b), middle x = D \ b, right x = P (U \ b); transpose
 * swaps left and right. b, x: device doubles of length n; x may be b (the
 * solve then runs in place, as SQMR applies it). */
culip_status culip_bildlt_solve(const culip_bildlt * h, culip_side side,
    int32_t transpose, const double * b, double * x, void * stream);

/* Allocation-free variant of culip_bildlt_solve. It has no counterpart in
 * the reference; it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops). Same arguments and
 * results: exact solves (SINGLE or LEVELS) run as level launches, which
 * need no scratch and give the same bits; block-Jacobi solves keep b and
 * their two iterates in a static device array of 3 x 1,048,576 doubles
 * (24 MiB) that every GPU the library runs on holds until the process ends,
 * instead of the object's lazily allocated vectors. Systems above
 * 1,048,576 rows are refused with CULIP_ERR_INVALID. Every call on one GPU
 * shares that GPU's array and it keeps the last call's contents: calls
 * that overlap on one GPU (other streams, other threads) race. */
culip_status culip_bildlt_solve_noalloc(const culip_bildlt * h, culip_side side,
    int32_t transpose, const double * b, double * x, void * stream);

int32_t culip_bildlt_n(const culip_bildlt * h);

/* set_solve_algorithm(algorithm, num_jacobi_sweeps) (block_ildlt.cuh:
 * 226-228): how the block triangular solves run. SINGLE (default,
 * SOLVE_BLOCK): one launch per solve; resident waves take the block rows in
 * level order and each block row waits for the individual values it reads.
 * Such a solve holds a quarter of the wave slots the kernel can use until
 * it ends, so the library admits one only while fewer than 4 other streams
 * of the device have one in flight; otherwise (and on a stream under graph
 * capture) that solve runs as LEVELS, with the same bits. Solves that other
 * processes run on the same GPU are not counted: a single-launch solve that
 * then gives up waiting leaves NaN in x and the next solve, refactor or
 * SQMR call on the object returns CULIP_ERR_HIP (once). LEVELS: one launch
 * per level. SINGLE and LEVELS compute the same values in the same order.
 * JACOBI (SOLVE_JACOBI,
 * block_ildlt.impl.cuh:2454-2493): num_jacobi_sweeps >= 1 block-Jacobi
 * sweeps from x = 0, x <- U_ii^-1 (b - U_off x), one launch each — an
 * approximate solve that is a fixed linear operator (exact once the sweeps
 * reach the number of levels). Solves with one object must be ordered on
 * one stream. */
typedef enum {
    CULIP_TRSV_LEVELS = 0,
    CULIP_TRSV_SINGLE = 1,
    CULIP_TRSV_JACOBI = 2
} culip_solve_algorithm;
/* BlockiLDLt<T, bk, rook> (block_ildlt_instances.cu:54-60): static pivots
 * from piv_starts, or numeric Bunch-Kaufman / rook pivoting inside every
 * pivot block (block_ildlt_kernels.impl.cuh:743-948) with the symmetric
 * interchanges applied to the block row and to the columns above
 * (:1352-1418); the solves then apply P (k_permute_P / k_permute_Pt). Set
 * before compute / refactor. */
typedef enum {
    CULIP_PIVOT_STATIC = 0,
    CULIP_PIVOT_BUNCH_KAUFMAN = 1,
    CULIP_PIVOT_ROOK = 2
} culip_pivoting;
culip_status culip_bildlt_set_pivoting(culip_bildlt * h, culip_pivoting mode);
/* Arithmetic of the 2x2 pivot steps (d_ldlt_2x2_step,
 * block_ildlt_kernels.impl.cuh:589-675) and of the Givens-form D solves of
 * the off-diagonal blocks (:1533-1566) in an FP32 factor: F64 (the default)
 * forms 2x2 determinants, rotations and inverses in FP64 and stores the
 * results in FP32; FACTOR computes them in FP32 as BlockiLDLt<float> does
 * (bit for bit). No effect on an FP64 factor or on 1x1 pivots. Set before
 * compute / refactor. */
typedef enum {
    CULIP_PIVOT2_F64 = 0,
    CULIP_PIVOT2_FACTOR = 1
} culip_pivot_precision;
culip_status culip_bildlt_set_pivot_precision(culip_bildlt * h, culip_pivot_precision p);
culip_status culip_bildlt_set_solve_algorithm(culip_bildlt * h,
    culip_solve_algorithm algorithm, int32_t num_jacobi_sweeps);

/* export_factorized(folder) (block_ildlt.impl.cuh:1700-1916): writes
 * folder/out_U.mtx (the nonzero entries of every factored block as one fine
 * CSR matrix), folder/out_D.mtx (the tridiagonal D) and folder/out_P.mtx
 * (the identity of static pivoting) in the MatrixMarket form of
 * test.impl.h:480-545 ("%d %d %.12g" lines, 1-based). fill_factor (may be
 * NULL) receives (2 nnz(U) + nnz(D)) / nnz(A), the figure the reference
 * prints. Needs a computed factor. */
culip_status culip_bildlt_export_factorized(culip_bildlt * h, const char * folder,
    double * fill_factor);

/* Host views of the symbolic structures and (after compute) of the factor
 * downloaded from the device; valid until the next call on h.
 * int names: block_starts, is_piv, coarse_csr_off, coarse_csr_col,
 *   coarse_csc_off, coarse_csc_idx, a_block_nnz, lvl_off, lvl_rows,
 *   row_level, lvlb_off, lvlb_idx, block_in_lvl, row_first_block_offset,
 *   solve_off, solve_idx, blocks (9 int32 per block: id, format, nnz,
 *   max_nnz, got_extra, block_row, block_col, col_ptr, val_ptr), ix_store,
 *   upd_off, upd_ki, upd_kj, scalars, and after compute fblocks, fix_store,
 *   carry, plan_sizes, perm (new -> old rows), fis_piv (pivot starts as
 *   factored).
 * real names: val_store; after compute fval_store, tri (3 per row:
 *   D(i-1,i), D(i,i), D(i+1,i)), row_norms. */
culip_status culip_bildlt_get_int(culip_bildlt * h, const char * name,
    const int32_t ** ptr, int64_t * len);
culip_status culip_bildlt_get_real(culip_bildlt * h, const char * name,
    const double ** ptr, int64_t * len);
/* Device bytes held by this object. */
int64_t culip_bildlt_device_bytes(const culip_bildlt * h);
/* Numeric refactors as HIP graph replays (no counterpart in the reference;
 * the reference launches each level's kernels from a host loop): with
 * enable != 0, a refactor on an explicit (non-NULL) stream that is not
 * being captured replays a graph of the whole launch sequence, captured on
 * the first such refactor and again after a threshold change or a new
 * symbolic phase; same results as the launches themselves. Solves are not
 * affected. */
culip_status culip_bildlt_set_graphs(culip_bildlt * h, int32_t enable);
/* Testing hook, no counterpart in the reference: records, from the host, a
 * single-launch solve whose dependency wait gave up, as the kernel would
 * (after upload; CULIP_ERR_STATE before). */
culip_status culip_bildlt_inject_solve_error(culip_bildlt * h);
void culip_bildlt_destroy(culip_bildlt * h);

/* ------------------------------------------------------------------------ */
/* SPMV (libs/la/spmv.cuh:18-149).                                          */

typedef struct culip_spmv culip_spmv;
/* CSRMatrixSPMV (spmv.cuh:36-66): uploads a copy of the host CSR. */
culip_status culip_spmv_create_csr(const culip_csr * hA, culip_spmv ** out);
/* AugmentedMatrixSPMV (spmv.cuh:116-149): K = [diag(d) A'; A delta I(ix)]
 * matrix-free from A (m x n), d (n), delta and ix (n_ix rows of the dual
 * block; ix == NULL: none). Dimensions are n + m (the reference's m()/n()
 * return A->m, spmv.impl.cuh:319-337; fixed here). */
/* AugmentedMatrixSPMV's operator with a general H (spmv.cuh:116-149 takes
 * a diagonal d only): K = [H A'; A delta I(ix)] matrix-free from H (n x n
 * CSR, symmetric), A (m x n), delta and ix, so an interior-point KKT
 * system with a full Hessian needs no assembled K for SQMR. y_1 sums a
 * row of H, then the row of A' (eight lanes, fixed order). */
culip_status culip_spmv_create_kkt_csr(const culip_csr * hH, const culip_csr * hA,
    double delta, int32_t n_ix, const int32_t * ix, culip_spmv ** out);
culip_status culip_spmv_create_kkt(const culip_csr * hA, const double * d,
    double delta, int32_t n_ix, const int32_t * ix, culip_spmv ** out);
/* NormalMatrixSPMV (spmv.cuh:74-107, spmv.impl.cuh:122-265): the m x m
 * operator N = A diag(d)^-1 A' + delta I on the rows in ix, from host A
 * (m x n) and d (n). The reference adds delta * x[k] at row ix[k]
 * (k_add_app); here row ix[k] gets delta * x[ix[k]], which is the same when
 * ix lists rows in order from 0 or is empty. transpose is ignored (N is
 * symmetric). */
culip_status culip_spmv_create_normal(const culip_csr * A, const double * d, double delta,
    int32_t n_ix, const int32_t * ix, culip_spmv ** out);
/* Replace the values of a CSR operator (same pattern), host array. */
culip_status culip_spmv_update_values(culip_spmv * op, const double * vals,
    int32_t vals_on_device, void * stream);
/* multiply (spmv.cuh:24-31): y = alpha * op(A) x + beta * y. */
culip_status culip_spmv_multiply(const culip_spmv * op, const double * x,
    double * y, double alpha, double beta, int32_t transpose, void * stream);
/* Symmetric equilibration of a CSR operator's values on the device (no
 * counterpart in the reference; input preparation for barrier-conditioned
 * KKT systems): sweeps >= 1 sweeps from d = 1, each setting, for all rows
 * at once, d_i <- d_i / sqrt(max_j |(d_i a_ij) d_j|) (Ruiz's iteration);
 * then the operator's values become D A D. d: device, n doubles, receives
 * diag(D). Operators from culip_spmv_create_csr with a symmetric matrix. */
culip_status culip_spmv_equilibrate(culip_spmv * op, int32_t sweeps, double * d,
    void * stream);
int32_t culip_spmv_m(const culip_spmv * op);
int32_t culip_spmv_n(const culip_spmv * op);
void culip_spmv_destroy(culip_spmv * op);

/* x = U' \ (P' (K src)): the left solve of an operator product, as SQMR
 * applies it every iteration. Where the factor's solve forms its rows of
 * K src itself (single-launch and block-Jacobi schedules, static pivots) it
 * does so in the same launch; otherwise K src is formed first and solved
 * (the same bits either way). K: a CSR operator (culip_spmv_create_csr) of
 * the factor's size; src and x: device doubles, distinct. fused (may be
 * NULL) receives 1 when the product was folded into the solve. No
 * counterpart in the reference (its SQMR calls the SpMV and then the left
 * solve). */
culip_status culip_bildlt_solve_left_product(const culip_bildlt * h, const culip_spmv * K,
    const double * src, double * x, int32_t * fused, void * stream);

/* ------------------------------------------------------------------------ */
/* SQMR (libs/la/iterative_solvers/sqmr.cuh:21-72).                         */

typedef struct {
    double tol;             /* reference: 1e-8 in the driver */
    int32_t max_it;         /* reference: 1000 in the driver */
    /* 0: reference semantics only (quasi-residual test, 4-iteration
     * throttle). k > 0: every k iterations recompute x and the true
     * residual ||b - A x|| / ||b|| in FP64; stop when it is <= tol and
     * restart from x (iterative refinement) when the recurrence has
     * drifted. */
    int32_t true_residual_period;
    int32_t max_restarts;
    /* 0 (default): with a CSR operator, the left preconditioner forms
     * A x itself where it can (single-launch and block-Jacobi BILDLt solves),
     * with the same values; 1: apply A, then the left preconditioner, as two
     * steps */
    int32_t separate_product;
} culip_sqmr_opts;

typedef struct {
    double residual;            /* gamma / ||Pl \ b|| (sqmr.impl.cuh:343) */
    int32_t iterations;         /* total SQMR iterations */
    int32_t converged;          /* 1 if the stopping test was met */
    int32_t restarts;           /* refinement restarts */
    double true_rel_residual;   /* ||b - A x|| / ||b|| at exit, FP64 */
} culip_sqmr_result;

typedef struct culip_sqmr culip_sqmr;

/* SQMR ctor (sqmr.impl.cuh:19-56): operator, optional preconditioner
 * (NULL: none), options; allocates the device work vectors once. The
 * operator and preconditioner are borrowed and must outlive the solver. */
culip_status culip_sqmr_create(const culip_spmv * A, const culip_bildlt * prec,
    const culip_sqmr_opts * opts, culip_sqmr ** out);

/* SQMR::solve (sqmr.impl.cuh:109-355). b, x: device doubles. Synchronises
 * `stream` before returning. */
culip_status culip_sqmr_solve(culip_sqmr * s, const double * b, double * x,
    culip_sqmr_result * res, void * stream);
/* The true-residual gate and the reported true residual of later solves
 * measure ||(b - A x) ./ s|| / ||b ./ s|| (no counterpart in the
 * reference): with A = D K D, b = D b0 and s = diag(D) that is the residual
 * of K for x0 = D x. s: device, n doubles, borrowed until reset; NULL
 * resets to s = 1. */
culip_status culip_sqmr_set_residual_scale(culip_sqmr * s, const double * scale);
void culip_sqmr_destroy(culip_sqmr * s);

/* ------------------------------------------------------------------------ */
/* ImportLDUPreconditioner (libs/la/preconditioner/import_ldu_preconditioner.cuh:
 * 36-125): an imported factor M = inv(S_r) P_r' L D R P_c' inv(S_c). */

typedef struct culip_ldu culip_ldu;

/* The three ctors (import_ldu_preconditioner.impl.cuh:19-130) in one:
 * L (host CSR, n x n; its lower triangle with a non-unit diagonal is used)
 * and U (NULL: R = L', the symmetric ctors; else its upper triangle);
 * tri / is_piv (3n doubles / n ints, host; both NULL: no D) in the layout
 * of the middle solve (helper_kernels.impl.cuh:720-772). exact_solve = 0
 * applies 10 Jacobi sweeps instead of the exact substitution
 * (sparse_la.impl.cuh:79-126). prec = F32 stores L and U in FP32 (sums in
 * FP64). Everything is copied to the device; a row without a diagonal
 * entry divides by 1. */
culip_status culip_ldu_create(const culip_csr * L, const culip_csr * U, const double * tri,
    const int32_t * is_piv, int32_t exact_solve, culip_prec prec, culip_ldu ** out);

/* set_permutation (:142-155): n-entry host arrays, push convention (entry
 * i moves to p[i]); both NULL removes the permutation. */
culip_status culip_ldu_set_permutation(culip_ldu * h, const int32_t * p_r, const int32_t * p_c);

/* set_scaling (:157-170): componentwise factors, host; both NULL removes. */
culip_status culip_ldu_set_scaling(culip_ldu * h, const double * s_r, const double * s_c);

/* solve_left / solve_middle / solve_right (:197-330). b, x: n device
 * doubles (b == x allowed). Asynchronous on `stream`; one handle's solves
 * share its work vectors, so they must not overlap. */
culip_status culip_ldu_solve(const culip_ldu * h, culip_side side, int32_t transpose,
    const double * b, double * x, void * stream);
int32_t culip_ldu_n(const culip_ldu * h);
/* is_middle (:180-186): 1 when D was given */
int32_t culip_ldu_is_middle(const culip_ldu * h);
void culip_ldu_destroy(culip_ldu * h);

/* SQMR ctor with an imported factor as the preconditioner
 * (sqmr.impl.cuh:19-56, Preconditioner<T> = ImportLDUPreconditioner<T>). */
culip_status culip_sqmr_create_ldu(const culip_spmv * A, const culip_ldu * prec,
    const culip_sqmr_opts * opts, culip_sqmr ** out);

/* ------------------------------------------------------------------------ */
/* Request loop for KKT solves (no counterpart in the reference, which
 * solves one system per process): sparsity patterns are registered once
 * (symbolic analysis and static device structures stay resident), requests
 * carry new values of K and a right-hand side and are served by `slots`
 * worker threads with one HIP stream each, so solves that arrive while
 * others run start as soon as a slot and device memory are free and
 * overlap on the device. Per request: BlockiLDLt refactor + SQMR.       */

typedef struct {
    int32_t slots;              /* concurrent solves, 1..64 */
    int64_t hbm_budget;         /* device bytes for patterns and per-slot
                                   instances; 0: 90% of free memory */
    culip_prec prec;            /* factor precision (SQMR is FP64) */
    int32_t fill_level;
    double fill_factor;
    double threshold;
    culip_sqmr_opts sqmr;
    /* triangular-solve schedule of every pattern unless set per pattern
     * (culip_server_set_pattern_solve); as culip_bildlt_set_solve_algorithm */
    culip_solve_algorithm solve_algorithm;
    int32_t jacobi_sweeps;
    /* > 0: every request's K is equilibrated on the device with this many
     * sweeps (culip_spmv_equilibrate) and solved as D K D y = D b with the
     * true residual measured on K, x = D y; 0: K as given */
    int32_t equilibrate_sweeps;
    /* != 0: every instance replays its refactor and the SQMR iterations
     * between true-residual checks from captured HIP graphs (same results,
     * far fewer launches for small systems) */
    int32_t graphs;
} culip_server_opts;

typedef struct culip_server culip_server;
culip_status culip_server_create(const culip_server_opts * o, culip_server ** out);
/* Symbolic analysis and upload of a pattern: K full symmetric CSR (its
 * values are not used), ordering and blocking as for culip_bildlt_create.
 * Thread-safe; CULIP_ERR_BUDGET if the static data does not fit. */
culip_status culip_server_add_pattern(culip_server * s, const culip_csr * K,
    int32_t num_pivs, const int32_t * piv_starts, int32_t num_blocks,
    const int32_t * block_starts, int32_t * pattern);
/* The triangular-solve schedule of one pattern (default: the server's);
 * only before the pattern's first submit (CULIP_ERR_STATE after). With
 * CULIP_TRSV_SINGLE and more than 4 slots, solves beyond 4 at a time run as
 * level launches (same bits). */
culip_status culip_server_set_pattern_solve(culip_server * s, int32_t pattern,
    culip_solve_algorithm algorithm, int32_t jacobi_sweeps);
/* Queues a solve: vals = nnz values of K in the pattern's entry order,
 * rhs = N doubles (host; copied before return). */
culip_status culip_server_submit(culip_server * s, int32_t pattern,
    const double * vals, const double * rhs, int64_t * request);
/* Waits for a request; x: N doubles (host, may be NULL). timing (may be
 * NULL): {queued, service, latency, submit time} in seconds. slot: the
 * worker that served it (may be NULL). The request is forgotten after. */
culip_status culip_server_wait(culip_server * s, int64_t request, double * x,
    culip_sqmr_result * res, double * timing, int32_t * slot);
/* out[5]: device bytes in use, peak, budget, instances, completed. */
culip_status culip_server_stats(const culip_server * s, int64_t * out);
void culip_server_destroy(culip_server * s);

/* ------------------------------------------------------------------------ */
/* Batches of independent systems (no counterpart in the reference): every
 * system of the batch is refactored and solved by SQMR with the FP64
 * true-residual gate and block-Jacobi triangular solves; each step of the
 * refactor (a level of all systems) and of the recurrence is one launch over
 * all systems that take part in it. Each system's x, residuals and
 * iteration counts are bit-identical to its solve alone (culip_bildlt with
 * CULIP_TRSV_JACOBI and the same sweeps, culip_sqmr with the same options).
 * Not thread-safe: one solve at a time per batch.                        */

typedef struct {
    culip_prec prec;            /* factor precision (SQMR is FP64) */
    int32_t fill_level;
    double fill_factor;
    double threshold;           /* drop tolerance of every refactor */
    culip_sqmr_opts sqmr;       /* true_residual_period must be > 0 */
    int32_t jacobi_sweeps;      /* block-Jacobi sweeps per triangular solve, >= 1 */
} culip_batch_opts;

typedef struct culip_batch culip_batch;
culip_status culip_batch_create(const culip_batch_opts * o, culip_batch ** out);
/* Symbolic analysis and upload of one system (K full symmetric CSR, its
 * values not used; ordering and blocking as for culip_bildlt_create).
 * Thread-safe; only before the first solve (CULIP_ERR_STATE after).
 * index: the system's position in the arrays of culip_batch_solve. */
culip_status culip_batch_add(culip_batch * b, const culip_csr * K, int32_t num_pivs,
    const int32_t * piv_starts, int32_t num_blocks, const int32_t * block_starts,
    int32_t * index);
/* Refactors system i with vals[i] (nnz values in its entry order) and
 * solves it for rhs[i] into x[i] for every system; the arrays are host
 * memory, or device memory when device_arrays != 0. res: one result per
 * system (may be NULL). Synchronises stream. */
culip_status culip_batch_solve(culip_batch * b, const double * const * vals,
    const double * const * rhs, double * const * x, culip_sqmr_result * res,
    int32_t device_arrays, void * stream);
/* systems registered, device bytes held, and seconds[2] of the last solve:
 * the refactors, then the SQMR (any may be NULL) */
culip_status culip_batch_info(const culip_batch * b, int32_t * count, int64_t * device_bytes,
    double * seconds);
void culip_batch_destroy(culip_batch * b);

#ifdef __cplusplus
}
#endif

#endif
