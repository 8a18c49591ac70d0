// Sparse triangular substitution for an imported factor (CSR rows grouped
// by dependency level) and the permutation / scaling steps around it.
#pragma once

#include <cstdint>

namespace culip {

// One triangular operand: CSR rows (values in VT), the diagonal position of
// each row (-1: none, taken as 1), and the rows in level order. `lower`
// selects which triangle the substitution reads (entries of the other
// triangle are ignored, as a lower/upper fill mode does).
template<typename VT>
struct TriOperand {
    int32_t n = 0;
    bool lower = true;
    const int32_t* row_ptr = nullptr;
    const int32_t* col = nullptr;
    const VT* val = nullptr;
    const int32_t* diag_pos = nullptr;
    const int32_t* order = nullptr;     // n rows, level by level
    const double* inv_diag = nullptr;   // approximate solves: 1 / diagonal
};

// x = M \ x by levels; lvl_off is a host array of levels + 1 offsets into
// `order`.
template<typename VT>
void tri_solve_levels(const TriOperand<VT>& M, const int32_t* lvl_off, int32_t levels, double* x,
    void* stream);

// x = approximate M \ b by `sweeps` Jacobi steps from x = D^-1 b
// (sparse_la.impl.cuh:79-126); tmp: n doubles.
template<typename VT>
void tri_solve_jacobi(const TriOperand<VT>& M, const double* b, double* x, double* tmp,
    int32_t sweeps, void* stream);

// x = P' (s .* b) (pull) or x = s .* (P b) (push); perm / scale may be NULL
void permute_scale(int32_t n, const int32_t* perm, const double* scale, bool push,
    const double* b, double* x, void* stream);

}  // namespace culip
