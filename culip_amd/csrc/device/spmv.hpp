// Device sparse operators.
#pragma once

#include <cstdint>

namespace culip {

struct DevCsr {
    int32_t m = 0, n = 0;
    int64_t nnz = 0;
    int32_t* row_ptr = nullptr;
    int32_t* col = nullptr;
    double* val = nullptr;
};

// K = [H A'; A diag(dix)] with H = diag(d) or a CSR matrix (H.row_ptr set)
// and dix = delta on the rows in ix, else 0.
struct DevKkt {
    int32_t n = 0, m = 0;
    DevCsr A, At, H;
    double* d = nullptr;    // n (diagonal H)
    double* dix = nullptr;  // m
};

// N = A diag(d)^-1 A' + diag(dix), m x m (NormalMatrixSPMV,
// spmv.impl.cuh:122-265); t: n doubles of work.
struct DevNormal {
    int32_t m = 0, n = 0;
    DevCsr A, At;
    double* d = nullptr;    // n
    double* dix = nullptr;  // m
    double* t = nullptr;    // n
};

void csr_spmv(const DevCsr& A, const double* x, double* y, double alpha, double beta,
    void* stream);
void normal_spmv(const DevNormal& N, const double* x, double* y, double alpha, double beta,
    void* stream);
void kkt_spmv(const DevKkt& K, const double* x, double* y, double alpha, double beta,
    void* stream);

}  // namespace culip
