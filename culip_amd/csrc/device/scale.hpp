// Symmetric equilibration of a CSR matrix on the device.
#pragma once

#include <cstdint>

#include "spmv.hpp"

namespace culip {

// `sweeps` >= 1 Ruiz sweeps from d = 1 on A (square, symmetric pattern and
// values), then A <- D A D in place. d receives the scaling (A.m doubles);
// tmp is A.m doubles of work.
void equilibrate_csr(const DevCsr& A, int32_t sweeps, double* d, double* tmp, void* stream);

// out = d .* in (out may be in)
void scale_vector(int32_t n, const double* d, const double* in, double* out, void* stream);

}  // namespace culip
