// The reference driver's call sequence (apps/block_ildlt/block_ildlt_invariants.h:
// 126-194: BlockiLDLt ctor -> compute -> solve_analysis -> b = A * 1 ->
// CSRMatrixSPMV -> SQMR(1e-8, 1000) -> solve) written against culip_amd.hpp.
// Usage: kkt_driver N_PRIMAL N_DUAL [f32] [bk|rook]. Prints one JSON line; exit 0 when
// the true relative residual is <= 1e-8. With --files the driver reads its
// inputs as the reference driver does (run_files below).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "culip_amd.hpp"

using namespace culip_amd;

#define HIPOK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if(e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
            return 2;                                                                  \
        }                                                                              \
    } while(0)

template<typename T, bool BK = false, bool ROOK = false>
static int run(int32_t n, int32_t m)
{
    culip_kkt_params p{};
    p.n = n;
    p.m = m;
    p.h_offdiag_per_row = 6;
    p.a_nnz_per_col = 4;
    p.h_diag_lo = 1.0;
    p.h_diag_hi = 10.0;
    p.h_diag_log = 0;
    p.h_offdiag_std = 0.1;
    p.delta = 1e-6;
    p.seed = 0;
    culip_kkt* kkt = nullptr;
    check(culip_kkt_generate(&p, &kkt));
    culip_csr K;
    const double* rhs = nullptr;
    check(culip_kkt_view(kkt, &K, &rhs));
    culip_prep* prep = nullptr;
    check(culip_prep_create(&K, n, 32, CULIP_ORDER_BLOCK_COLOR, &prep));
    culip_csr Kp;
    const int32_t *perm, *bs, *ps;
    int32_t nb, np;
    check(culip_prep_view(prep, &Kp, &perm, &nb, &bs, &np, &ps));

    int rc = 1;
    {
        BlockiLDLt<T, BK, ROOK> blocky(nullptr, &Kp, np, ps, nb, bs);
        blocky.compute(1, T(1.0), T(1e-4));
        blocky.set_solve_algorithm(SOLVE_BLOCK, 0);
        blocky.solve_analysis();

        const int32_t N = Kp.m;
        double *d_x = nullptr, *d_b = nullptr, *d_sol = nullptr;
        HIPOK(hipMalloc(&d_x, sizeof(double) * N));
        HIPOK(hipMalloc(&d_b, sizeof(double) * N));
        HIPOK(hipMalloc(&d_sol, sizeof(double) * N));
        std::vector<double> ones(N, 1.0);
        HIPOK(hipMemcpy(d_x, ones.data(), sizeof(double) * N, hipMemcpyHostToDevice));

        CSRMatrixSPMV spmv(nullptr, &Kp);
        spmv.multiply(d_x, d_b, 1.0, 0.0);

        SQMR sqmr(nullptr, &spmv, 1e-8, 1000, &blocky, 20, 10);
        double residual = 0;
        int32_t iterations = 0;
        sqmr.solve(d_b, d_sol, residual, iterations);

        std::vector<double> x(N);
        HIPOK(hipMemcpy(x.data(), d_sol, sizeof(double) * N, hipMemcpyDeviceToHost));
        double err = 0;
        for(double v : x) err = std::fmax(err, std::fabs(v - 1.0));
        const culip_sqmr_result& r = sqmr.last();
        std::printf("{\"n\": %d, \"prec\": \"%s\", \"residual\": %.3e, \"iterations\": %d, "
                    "\"restarts\": %d, \"true_rel_residual\": %.3e, \"max_err\": %.3e}\n",
            N, sizeof(T) == 4 ? "f32" : "f64", residual, iterations, r.restarts,
            r.true_rel_residual, err);
        rc = (r.true_rel_residual <= 1e-8) ? 0 : 1;
        HIPOK(hipFree(d_x));
        HIPOK(hipFree(d_b));
        HIPOK(hipFree(d_sol));
    }
    culip_prep_destroy(prep);
    culip_kkt_destroy(kkt);
    return rc;
}

// bildlt_fun (block_ildlt_invariants.h:73-194) from files: matrix, block
// starts (with the sentinel N), permutation (new -> old) and is_piv as
// MatrixMarket; pivot starts from is_piv plus the sentinel; host_permute;
// BlockiLDLt; b = K 1; SQMR(1e-8, 1000) with the reference's stopping test.
static std::vector<int32_t> read_ints(const char* path)
{
    int64_t len = 0;
    double* v = nullptr;
    check(culip_mm_read_vector(path, &len, &v));
    std::vector<int32_t> out(v, v + len);
    culip_host_free(v);
    return out;
}

template<typename T>
static int run_files(char** f, int32_t fill_level, double fill_factor, double threshold)
{
    culip_host_matrix* hA = nullptr;
    check(culip_mm_read_csr(f[0], &hA));
    culip_csr A;
    check(culip_host_matrix_view(hA, &A));
    const std::vector<int32_t> bs = read_ints(f[1]), perm = read_ints(f[2]), is_piv = read_ints(f[3]);
    std::vector<int32_t> piv;
    for(size_t i = 0; i < is_piv.size(); ++i)
        if(is_piv[i] == 1) piv.push_back((int32_t) i);
    piv.push_back(A.m);
    culip_host_matrix* hP = nullptr;
    check(culip_csr_permute_scale(&A, perm.data(), perm.data(), 0, nullptr, nullptr, 0, &hP));
    culip_csr Kp;
    check(culip_host_matrix_view(hP, &Kp));
    int rc = 1;
    {
        BlockiLDLt<T> blocky(nullptr, &Kp, (int32_t) piv.size() - 1, piv.data(),
            (int32_t) bs.size() - 1, bs.data());
        blocky.compute(fill_level, T(fill_factor), T(threshold));
        blocky.set_solve_algorithm(SOLVE_BLOCK, 0);
        blocky.solve_analysis();
        const int32_t N = Kp.m;
        double *d_x = nullptr, *d_b = nullptr, *d_sol = nullptr;
        HIPOK(hipMalloc(&d_x, sizeof(double) * N));
        HIPOK(hipMalloc(&d_b, sizeof(double) * N));
        HIPOK(hipMalloc(&d_sol, sizeof(double) * N));
        std::vector<double> ones(N, 1.0);
        HIPOK(hipMemcpy(d_x, ones.data(), sizeof(double) * N, hipMemcpyHostToDevice));
        CSRMatrixSPMV spmv(nullptr, &Kp);
        spmv.multiply(d_x, d_b, 1.0, 0.0);
        SQMR sqmr(nullptr, &spmv, 1e-8, 1000, &blocky);
        double residual = 0;
        int32_t iterations = 0;
        sqmr.solve(d_b, d_sol, residual, iterations);
        std::vector<double> x(N);
        HIPOK(hipMemcpy(x.data(), d_sol, sizeof(double) * N, hipMemcpyDeviceToHost));
        double err = 0;
        for(double v : x) err = std::fmax(err, std::fabs(v - 1.0));
        const culip_sqmr_result& r = sqmr.last();
        std::printf("{\"n\": %d, \"prec\": \"%s\", \"residual\": %.17g, \"iterations\": %d, "
                    "\"true_rel_residual\": %.3e, \"max_err\": %.3e}\n",
            N, sizeof(T) == 4 ? "f32" : "f64", residual, iterations, r.true_rel_residual, err);
        rc = 0;
        HIPOK(hipFree(d_x));
        HIPOK(hipFree(d_b));
        HIPOK(hipFree(d_sol));
    }
    culip_host_matrix_destroy(hP);
    culip_host_matrix_destroy(hA);
    return rc;
}

int main(int argc, char** argv)
{
    if(argc >= 2 && std::strcmp(argv[1], "--files") == 0) {
        if(argc < 9) {
            std::fprintf(stderr, "usage: %s --files A BLOCK_STARTS PERM IS_PIV FILL FF THRESH [f32]\n",
                argv[0]);
            return 2;
        }
        const bool f32 = argc > 9 && std::strcmp(argv[9], "f32") == 0;
        try {
            return f32 ? run_files<float>(argv + 2, std::atoi(argv[6]), std::atof(argv[7]),
                             std::atof(argv[8]))
                       : run_files<double>(argv + 2, std::atoi(argv[6]), std::atof(argv[7]),
                             std::atof(argv[8]));
        } catch(const Error& e) {
            std::fprintf(stderr, "culip_amd error %d: %s\n", (int) e.status, e.what());
            return 3;
        }
    }
    if(argc < 3) {
        std::fprintf(stderr, "usage: %s N_PRIMAL N_DUAL [f32] [bk|rook]\n"
                             "       %s --files A BLOCK_STARTS PERM IS_PIV FILL FF THRESH [f32]\n",
            argv[0], argv[0]);
        return 2;
    }
    const int32_t n = std::atoi(argv[1]), m = std::atoi(argv[2]);
    bool f32 = false, bk = false, rook = false;
    for(int a = 3; a < argc; ++a) {
        f32 = f32 || std::strcmp(argv[a], "f32") == 0;
        bk = bk || std::strcmp(argv[a], "bk") == 0;
        rook = rook || std::strcmp(argv[a], "rook") == 0;
    }
    try {
        if(bk) return f32 ? run<float, true, false>(n, m) : run<double, true, false>(n, m);
        if(rook) return f32 ? run<float, false, true>(n, m) : run<double, false, true>(n, m);
        return f32 ? run<float>(n, m) : run<double>(n, m);
    } catch(const Error& e) {
        std::fprintf(stderr, "culip_amd error %d: %s\n", (int) e.status, e.what());
        return 3;
    }
}
