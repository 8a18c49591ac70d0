/*
 * Oracle SQMR, SpMV, KKT adapter and ImportLDU preconditioner (double).
 * TEST INFRASTRUCTURE ONLY: the checker for the product path and the CPU
 * baseline of bench.py.
 *
 * SQMR follows libs/la/iterative_solvers/sqmr.impl.cuh:109-355 step by step,
 * including the 4-iteration throttled convergence test (:309-331) and the
 * final x <- Pr \ (D \ x) (:343-351). BLAS-1 reductions are plain sequential
 * sums in double (the reference uses cuBLAS, whose reduction order is not
 * documented; results therefore agree to rounding, not bitwise).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void or_csr_spmv(int m, const int * row_ptr, const int * col,
    const double * val, const double * x, double * y)
{
    for(int i = 0; i < m; ++i) {
        double acc = 0;
        for(int p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
            acc = fma(val[p], x[col[p]], acc);
        y[i] = acc;
    }
}

/* AugmentedMatrixSPMV::multiply (spmv.impl.cuh:362-427) with the operator
 * its own fixture pins (tests/la/data/spmv_K.mtx == [diag(d) A'; A 0]):
 * b1 = d .* x1 + A' x2, b2 = delta*x2(ix) + A x1, then b = alpha b + beta v.
 * The reference's A' product reads x_1 (:416-421) and m()/n() return A->m
 * (:319-337); both are bugs the fixture exposes, restated here fixed. */
void or_kkt_spmv(int m, int n, const int * a_row, const int * a_col,
    const double * a_val, const double * d, double delta, int n_ix,
    const int * ix, const double * x, double * y, double alpha, double beta)
{
    double * b = (double *) calloc((size_t) (n + m), sizeof(double));
    const double * x1 = x;
    const double * x2 = x + n;
    for(int i = 0; i < n; ++i) b[i] = x1[i] * d[i];
    for(int q = 0; q < n_ix; ++q) b[n + ix[q]] += delta * x2[ix[q]];
    for(int i = 0; i < m; ++i) {
        double acc = b[n + i];
        for(int p = a_row[i]; p < a_row[i + 1]; ++p)
            acc = fma(a_val[p], x1[a_col[p]], acc);
        b[n + i] = acc;
    }
    for(int i = 0; i < m; ++i)
        for(int p = a_row[i]; p < a_row[i + 1]; ++p)
            b[a_col[p]] = fma(a_val[p], x2[i], b[a_col[p]]);
    for(int i = 0; i < n + m; ++i) y[i] = alpha * b[i] + beta * y[i];
    free(b);
}

void or_solve_D(int n, const double * tri, const int * is_piv, double * b)
{
    double * in = (double *) malloc(sizeof(double) * (size_t) (n > 0 ? n : 1));
    memcpy(in, b, sizeof(double) * (size_t) n);
    for(int g = 0; g < n; ++g) {
        const int this_piv = is_piv[g];
        const int next_piv = (g < n - 1) ? is_piv[g + 1] : 1;
        const double p11 = tri[3 * g + 1];
        const double p12 = tri[3 * g + 2];
        const double p22 = !next_piv ? tri[3 * (g + 1) + 1] : 1.0;
        const double r1 = in[g];
        const double r2 = (g < n - 1) ? in[g + 1] : 0.0;
        const double gr = sqrt(fma(p11, p11, p12 * p12));
        const double c = p11 / gr, s = p12 / gr;
        double g1 = fma(c, r1, s * r2);
        double g2 = fma(-s, r1, c * r2);
        const double t12 = fma(c, p12, s * p22);
        const double t22 = fma(-s, p12, c * p22);
        g2 = g2 / t22;
        g1 = fma(-t12, g2, g1) / gr;
        if(this_piv) b[g] = g1;
        if(!next_piv) b[g + 1] = g2;
    }
    free(in);
}

/* ------------------------------------------------------------------------ */

typedef struct {
    int n;
    const int * row_ptr;
    const int * col;
    const double * val;
    void * ctx;
    void (*left)(void * ctx, const double * b, double * x);
    void (*middle)(void * ctx, const double * b, double * x);
    void (*right)(void * ctx, const double * b, double * x);
} sq_ops;

static void sq_copy(int n, const double * b, double * x) { memcpy(x, b, sizeof(double) * (size_t) n); }
static void sq_left(const sq_ops * o, const double * b, double * x) { if(o->left) o->left(o->ctx, b, x); else sq_copy(o->n, b, x); }
static void sq_middle(const sq_ops * o, const double * b, double * x) { if(o->middle) o->middle(o->ctx, b, x); else sq_copy(o->n, b, x); }
static void sq_right(const sq_ops * o, const double * b, double * x) { if(o->right) o->right(o->ctx, b, x); else sq_copy(o->n, b, x); }

static double sq_dot(int n, const double * a, const double * b)
{
    double s = 0;
    for(int i = 0; i < n; ++i) s = fma(a[i], b[i], s);
    return s;
}
static double sq_nrm2(int n, const double * a) { return sqrt(sq_dot(n, a, a)); }
static void sq_axpy(int n, double a, const double * x, double * y) { for(int i = 0; i < n; ++i) y[i] = fma(a, x[i], y[i]); }
static void sq_scal(int n, double a, double * x) { for(int i = 0; i < n; ++i) x[i] *= a; }
static double sq_sign(double v) { return v < 0 ? -1.0 : 1.0; }

/* k_sym_ortho (helper_kernels.impl.cuh:662-699) */
static void sq_sym_ortho(double a, double b, double * c, double * s, double * r)
{
    if(b == 0) {
        *s = 0; *r = fabs(a); *c = (a == 0) ? 1 : sq_sign(a);
    } else if(a == 0) {
        *c = 0; *s = sq_sign(b); *r = fabs(b);
    } else if(fabs(b) >= fabs(a)) {
        const double tao = a / b;
        *s = sq_sign(b) / sqrt(1 + tao * tao);
        *c = *s * tao;
        *r = b / *s;
    } else if(fabs(a) > fabs(b)) {
        const double tao = b / a;
        *c = sq_sign(a) / sqrt(1 + tao * tao);
        *s = *c * tao;
        *r = a / *c;
    }
}

/* prec_spmv (sqmr.impl.cuh:359-369): b = Pl \ (A (Pr \ x)) */
static void sq_prec_spmv(const sq_ops * o, const double * x, double * b,
    double * ev, double * ev2)
{
    sq_right(o, x, ev);
    or_csr_spmv(o->n, o->row_ptr, o->col, o->val, ev, ev2);
    sq_left(o, ev2, b);
}

#define ACC(r) (((r) % 3) + (((r) < 0) ? 3 : 0))

static int sq_solve(const sq_ops * o, const double * bvec, double * x,
    double tol, int max_it, or_sqmr_result * res)
{
    const int n = o->n;
    double * v[3], * w[3], * p[3];
    for(int i = 0; i < 3; ++i) {
        v[i] = (double *) calloc((size_t) n, sizeof(double));
        w[i] = (double *) calloc((size_t) n, sizeof(double));
        p[i] = (double *) calloc((size_t) n, sizeof(double));
    }
    double * Av = (double *) calloc((size_t) n, sizeof(double));
    double * tv = (double *) calloc((size_t) n, sizeof(double));
    double * ev = (double *) calloc((size_t) n, sizeof(double));
    double * ev2 = (double *) calloc((size_t) n, sizeof(double));

    double alpha = 0, beta[3] = {0, 0, 0}, delta[3] = {0, 0, 0}, normA = 0;
    double epsilon = 0, beta_rot = 0, beta_rot_rot = 0, alpha_rot = 0;
    double c[3] = {0, 0, 0}, s[3] = {0, 0, 0};
    double tau = 0, rho = 0, gamma = 0, norm_pb = 0, norm_x = 0, eta = 0, tmp;
    int test[2] = {0, 0}, host_test[2] = {0, 0};
    memset(x, 0, sizeof(double) * (size_t) n);
    c[ACC(0)] = -1.0;

    sq_left(o, bvec, v[0]);
    norm_pb = sq_nrm2(n, v[0]);
    tmp = 1.0 / norm_pb;
    sq_scal(n, tmp, v[0]);
    sq_middle(o, v[0], w[0]);
    eta = sq_nrm2(n, w[0]);
    eta = 1.0 / eta;
    sq_scal(n, eta, w[0]);
    gamma = norm_pb;
    sq_middle(o, v[0], tv);
    sq_prec_spmv(o, tv, Av, ev, ev2);

    int it = 0;
    while(it < max_it) {
        const int k0 = ACC(it), k1 = ACC(it + 1), km1 = ACC(it - 1), km2 = ACC(it - 2);
        sq_middle(o, Av, tv);
        alpha = sq_dot(n, v[k0], tv);
        sq_middle(o, v[k0], tv);
        tmp = sq_dot(n, v[k0], tv);
        alpha = alpha / tmp;

        memcpy(v[k1], Av, sizeof(double) * (size_t) n);
        sq_axpy(n, -alpha, v[k0], v[k1]);
        sq_axpy(n, -beta[k0], v[km1], v[k1]);
        delta[k1] = sq_nrm2(n, v[k1]);
        tmp = 1.0 / delta[k1];
        sq_scal(n, tmp, v[k1]);

        sq_middle(o, v[k1], w[k1]);
        tmp = eta * delta[k1];
        sq_scal(n, tmp, w[k1]);

        sq_middle(o, v[k1], tv);
        sq_prec_spmv(o, tv, Av, ev, ev2);

        beta[k1] = sq_dot(n, w[k0], Av);
        tmp = sq_dot(n, w[k0], v[k0]);
        beta[k1] = beta[k1] / tmp;
        tmp = 1.0 / beta[k1];
        sq_scal(n, tmp, w[k1]);
        tmp = delta[k1] / beta[k1];
        eta = tmp * eta;

        tmp = sqrt(beta[k0] * beta[k0] + alpha * alpha + delta[k1] * delta[k1]);
        normA = fmax(normA, tmp);

        epsilon = s[km1] * beta[k0];
        tmp = -c[km1];
        beta_rot = tmp * beta[k0];
        beta_rot_rot = c[k0] * beta_rot + s[k0] * alpha;
        tmp = -c[k0];
        alpha_rot = s[k0] * beta_rot + tmp * alpha;

        sq_sym_ortho(alpha_rot, delta[k1], &c[k1], &s[k1], &rho);
        tau = c[k1] * gamma;

        memcpy(p[k0], v[k0], sizeof(double) * (size_t) n);
        sq_axpy(n, -beta_rot_rot, p[km1], p[k0]);
        sq_axpy(n, -epsilon, p[km2], p[k0]);
        tmp = 1.0 / rho;
        sq_scal(n, tmp, p[k0]);
        sq_axpy(n, tau, p[k0], x);
        norm_x = sq_nrm2(n, x);

        gamma = s[k1] * gamma;

        /* k_sqmr_test (:89-107) */
        test[0] |= (fabs(delta[k1]) <= tol);
        test[1] |= (gamma / (normA * norm_x + norm_pb) <= tol);

        /* throttled host check (:309-331) */
        if(it > 0 && (it % 4) == 0) {
            if(host_test[0] || host_test[1]) break;
        } else if(((it + 1) % 4) == 0) {
            host_test[0] = test[0];
            host_test[1] = test[1];
        }
        ++it;
    }

    res->residual = gamma / norm_pb;
    res->iterations = it;
    res->flag_delta = test[0];
    res->flag_res = test[1];
    memcpy(Av, x, sizeof(double) * (size_t) n);
    sq_middle(o, Av, tv);
    sq_right(o, tv, x);

    for(int i = 0; i < 3; ++i) { free(v[i]); free(w[i]); free(p[i]); }
    free(Av); free(tv); free(ev); free(ev2);
    return 0;
}

/* ---- BILDLt preconditioner ---- */

static void bl_left_d(void * c, const double * b, double * x) { or_bildlt_solve_left_d((const or_bildlt_d *) c, b, x); }
static void bl_middle_d(void * c, const double * b, double * x) { or_bildlt_solve_middle_d((const or_bildlt_d *) c, b, x); }
static void bl_right_d(void * c, const double * b, double * x) { or_bildlt_solve_right_d((const or_bildlt_d *) c, b, x); }
static void bl_left_f(void * c, const double * b, double * x) { or_bildlt_solve_left_f((const or_bildlt_f *) c, b, x); }
static void bl_middle_f(void * c, const double * b, double * x) { or_bildlt_solve_middle_f((const or_bildlt_f *) c, b, x); }
static void bl_right_f(void * c, const double * b, double * x) { or_bildlt_solve_right_f((const or_bildlt_f *) c, b, x); }

int or_sqmr_bildlt_d(int n, const int * row_ptr, const int * col,
    const double * val, const or_bildlt_d * prec, const double * b, double * x,
    double tol, int max_it, or_sqmr_result * res)
{
    sq_ops o = { n, row_ptr, col, val, (void *) prec,
        prec ? bl_left_d : NULL, prec ? bl_middle_d : NULL, prec ? bl_right_d : NULL };
    return sq_solve(&o, b, x, tol, max_it, res);
}

int or_sqmr_bildlt_f(int n, const int * row_ptr, const int * col,
    const double * val, const or_bildlt_f * prec, const double * b, double * x,
    double tol, int max_it, or_sqmr_result * res)
{
    sq_ops o = { n, row_ptr, col, val, (void *) prec,
        prec ? bl_left_f : NULL, prec ? bl_middle_f : NULL, prec ? bl_right_f : NULL };
    return sq_solve(&o, b, x, tol, max_it, res);
}

/* ---- ImportLDUPreconditioner (import_ldu_preconditioner.impl.cuh) ---- */

typedef struct {
    int n;
    const int * l_row; const int * l_col; const double * l_val;
    const double * tri; const int * is_piv; const int * perm;
    const double * scale;
    double * tmp;
} ldu_ctx;

/* lower-triangular part of L (diagonal included, non-unit), as the legacy
 * csrsv with a GENERAL descriptor (types.impl.cuh:899-909) */
static void ldu_lower_solve(const ldu_ctx * c, double * x)
{
    for(int i = 0; i < c->n; ++i) {
        double acc = x[i], diag = 1.0;
        for(int p = c->l_row[i]; p < c->l_row[i + 1]; ++p) {
            const int j = c->l_col[p];
            if(j < i) acc = fma(-c->l_val[p], x[j], acc);
            else if(j == i) diag = c->l_val[p];
        }
        x[i] = acc / diag;
    }
}

static void ldu_lower_t_solve(const ldu_ctx * c, double * x)
{
    for(int i = c->n - 1; i >= 0; --i) {
        double diag = 1.0;
        for(int p = c->l_row[i]; p < c->l_row[i + 1]; ++p)
            if(c->l_col[p] == i) diag = c->l_val[p];
        x[i] = x[i] / diag;
        for(int p = c->l_row[i]; p < c->l_row[i + 1]; ++p) {
            const int j = c->l_col[p];
            if(j < i) x[j] = fma(-c->l_val[p], x[i], x[j]);
        }
    }
}

/* solve_left: x = L \ (P' (S .* b)) ; solve_PS(true, false) then csrsv */
static void ldu_left(void * cv, const double * b, double * x)
{
    const ldu_ctx * c = (const ldu_ctx *) cv;
    for(int i = 0; i < c->n; ++i) c->tmp[i] = b[i] * c->scale[i];
    for(int i = 0; i < c->n; ++i) x[i] = c->tmp[c->perm[i]];
    ldu_lower_solve(c, x);
}

static void ldu_middle(void * cv, const double * b, double * x)
{
    const ldu_ctx * c = (const ldu_ctx *) cv;
    memcpy(x, b, sizeof(double) * (size_t) c->n);
    or_solve_D(c->n, c->tri, c->is_piv, x);
}

/* solve_right: x = S .* (P (L' \ b)) */
static void ldu_right(void * cv, const double * b, double * x)
{
    const ldu_ctx * c = (const ldu_ctx *) cv;
    memcpy(c->tmp, b, sizeof(double) * (size_t) c->n);
    ldu_lower_t_solve(c, c->tmp);
    for(int i = 0; i < c->n; ++i) x[c->perm[i]] = c->tmp[i];
    for(int i = 0; i < c->n; ++i) x[i] = x[i] * c->scale[i];
}

int or_sqmr_ldu_d(int n, const int * row_ptr, const int * col,
    const double * val, const int * l_row, const int * l_col,
    const double * l_val, const double * tri, const int * is_piv,
    const int * perm, const double * scale, const double * b, double * x,
    double tol, int max_it, or_sqmr_result * res)
{
    ldu_ctx c = { n, l_row, l_col, l_val, tri, is_piv, perm, scale, NULL };
    c.tmp = (double *) calloc((size_t) (n > 0 ? n : 1), sizeof(double));
    sq_ops o = { n, row_ptr, col, val, &c,
        l_row ? ldu_left : NULL, (l_row && tri) ? ldu_middle : NULL,
        l_row ? ldu_right : NULL };
    const int r = sq_solve(&o, b, x, tol, max_it, res);
    free(c.tmp);
    return r;
}
