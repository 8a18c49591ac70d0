/*
 * CPU oracle for the BILDLt + SQMR path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference algorithm (dthuerck/culip,
 * libs/la/preconditioner/block_ildlt*.cuh, libs/la/iterative_solvers/sqmr*,
 * libs/la/helper_kernels.impl.cuh). It is the checker for the product path
 * and the CPU baseline of bench.py; no product code may link or call it.
 *
 * Every function names the reference lines it follows. Where the reference's
 * GPU execution order is not deterministic (OpenMP frontier order, atomics),
 * the oracle fixes the sequential order documented at that function.
 *
 * Real-typed entry points exist twice: suffix _d (T = double) and _f
 * (T = float, the reference's BlockiLDLt<float> semantics). Vectors crossing
 * this interface are always double; the _f variant rounds them to float.
 */
#ifndef CULIP_ORACLE_H
#define CULIP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_bildlt_d or_bildlt_d;
typedef struct or_bildlt_f or_bildlt_f;

/* ---- BILDLt (block_ildlt.impl.cuh) ---- */
#define OR_DECLARE(SUF)                                                        \
    or_bildlt_##SUF * or_bildlt_create_##SUF(int n, const int * row_ptr,        \
        const int * col, const double * val, int num_pivs,                     \
        const int * piv_starts, int num_blocks, const int * block_starts);     \
    int or_bildlt_symbolic_##SUF(or_bildlt_##SUF * h, int fill_level,          \
        double fill_factor);                                                   \
    int or_bildlt_factor_##SUF(or_bildlt_##SUF * h, double threshold);         \
    void or_bildlt_set_pivoting_##SUF(or_bildlt_##SUF * h, int mode);          \
    void or_bildlt_set_pivot_precision_##SUF(or_bildlt_##SUF * h, int f64);    \
    void or_bildlt_solve_left_##SUF(const or_bildlt_##SUF * h,                 \
        const double * b, double * x);                                         \
    void or_bildlt_solve_middle_##SUF(const or_bildlt_##SUF * h,               \
        const double * b, double * x);                                         \
    void or_bildlt_solve_right_##SUF(const or_bildlt_##SUF * h,                \
        const double * b, double * x);                                         \
    long or_bildlt_get_##SUF(const or_bildlt_##SUF * h, const char * name,     \
        void ** ptr);                                                          \
    void or_bildlt_destroy_##SUF(or_bildlt_##SUF * h);

OR_DECLARE(d)
OR_DECLARE(f)

/* ---- SpMV (cuSPARSE csrmv semantics, y = A x) ---- */
void or_csr_spmv(int m, const int * row_ptr, const int * col,
    const double * val, const double * x, double * y);

/* KKT adapter: K = [diag(d) A'; A delta*I(ix)], x = [x1 (n); x2 (m)],
 * y = alpha K x + beta y (spmv.impl.cuh:362-427 with the two bugs fixed) */
void or_kkt_spmv(int m, int n, const int * a_row, const int * a_col,
    const double * a_val, const double * d, double delta, int n_ix,
    const int * ix, const double * x, double * y, double alpha, double beta);

/* ---- SQMR (sqmr.impl.cuh:109-355) ---- */
typedef struct {
    double residual;    /* gamma / ||Pl \ b||, sqmr.impl.cuh:343 */
    int iterations;     /* loop counter at exit (throttled test) */
    int flag_delta;     /* test[0]: |delta_{k+1}| <= tol */
    int flag_res;       /* test[1]: quasi-residual test */
} or_sqmr_result;

/* SQMR on CSR A with a BILDLt oracle preconditioner (NULL: none) */
int or_sqmr_bildlt_d(int n, const int * row_ptr, const int * col,
    const double * val, const or_bildlt_d * prec, const double * b, double * x,
    double tol, int max_it, or_sqmr_result * res);
int or_sqmr_bildlt_f(int n, const int * row_ptr, const int * col,
    const double * val, const or_bildlt_f * prec, const double * b, double * x,
    double tol, int max_it, or_sqmr_result * res);

/* SQMR with the ImportLDUPreconditioner semantics the reference's own SQMR
 * fixtures use (import_ldu_preconditioner.impl.cuh; test_sqmr.cc:58-185):
 * M = S^-1 P L D L' P' S^-1; L lower CSR (non-unit), D tridiagonal + is_piv,
 * P 0-based push permutation, S scaling. L == NULL: no preconditioner. */
int or_sqmr_ldu_d(int n, const int * row_ptr, const int * col,
    const double * val, const int * l_row, const int * l_col,
    const double * l_val, const double * tri, const int * is_piv,
    const int * perm, const double * scale, const double * b, double * x,
    double tol, int max_it, or_sqmr_result * res);

/* k_solve_D semantics (helper_kernels.impl.cuh:720-772), in place */
void or_solve_D(int n, const double * tri, const int * is_piv, double * b);

#ifdef __cplusplus
}
#endif

#endif
