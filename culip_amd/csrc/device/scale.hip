// Symmetric equilibration of a CSR matrix on the device (Ruiz's iteration):
// D = diag(d) with every row of D A D scaled towards a largest |entry| of 1.
// Interior-point KKT systems whose barrier diagonals span many decades
// become far better conditioned for the incomplete factor and for SQMR's
// stopping rules; the solve then runs on D A D y = D b and returns x = D y.
//
// Each sweep, for every row i at once: m_i = max_j |(d_i a_ij) d_j| with
// the current d, then d_i <- d_i / sqrt(m_i) (rows without a nonzero keep
// d_i). Maxima are exact and sqrt and division are correctly rounded, so
// the result does not depend on the order of the work.
#include "scale.hpp"

#include <hip/hip_runtime.h>

namespace culip {

namespace {

constexpr int LPR = 8;   // lanes per row

__global__ __launch_bounds__(256) void k_ruiz_sweep(int32_t m, const int32_t* __restrict__ rp,
    const int32_t* __restrict__ ci, const double* __restrict__ cv,
    const double* __restrict__ d_in, double* __restrict__ d_out)
{
    const int64_t tid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t stride = (int64_t) gridDim.x * blockDim.x / LPR;
    for(int64_t row = tid / LPR; row < m; row += stride) {
        const double di = d_in ? d_in[row] : 1.0;
        double mx = 0.0;
        const int32_t s = rp[row], e = rp[row + 1];
        for(int32_t q = s + sub; q < e; q += LPR) {
            const double dj = d_in ? d_in[ci[q]] : 1.0;
            mx = fmax(mx, fabs((di * cv[q]) * dj));
        }
#pragma unroll
        for(int o = LPR / 2; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, LPR));
        if(sub == 0) d_out[row] = (mx > 0.0) ? di / sqrt(mx) : di;
    }
}

__global__ __launch_bounds__(256) void k_scale_values(int32_t m, const int32_t* __restrict__ rp,
    const int32_t* __restrict__ ci, double* __restrict__ cv, const double* __restrict__ d)
{
    const int64_t tid = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t stride = (int64_t) gridDim.x * blockDim.x / LPR;
    for(int64_t row = tid / LPR; row < m; row += stride) {
        const double di = d[row];
        const int32_t s = rp[row], e = rp[row + 1];
        for(int32_t q = s + sub; q < e; q += LPR) cv[q] = (di * cv[q]) * d[ci[q]];
    }
}

__global__ void k_scale_vector(int32_t n, const double* d, const double* in, double* out)
{
    const int32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if(g < n) out[g] = d[g] * in[g];
}

unsigned row_blocks(int32_t m)
{
    const int64_t want = ((int64_t) m * LPR + 255) / 256;
    return (unsigned) (want < 1 ? 1 : (want > 65536 ? 65536 : want));
}

}  // namespace

void equilibrate_csr(const DevCsr& A, int32_t sweeps, double* d, double* tmp, void* stream)
{
    hipStream_t st = (hipStream_t) stream;
    const int32_t m = A.m;
    if(m <= 0 || sweeps <= 0) return;
    const unsigned g = row_blocks(m);
    // sweep s writes into d when the sweeps left after it are even, so the
    // last one lands in d
    const double* in = nullptr;
    for(int32_t s = 0; s < sweeps; ++s) {
        double* out = ((sweeps - 1 - s) % 2 == 0) ? d : tmp;
        k_ruiz_sweep<<<g, 256, 0, st>>>(m, A.row_ptr, A.col, A.val, in, out);
        in = out;
    }
    k_scale_values<<<g, 256, 0, st>>>(m, A.row_ptr, A.col, A.val, d);
}

void scale_vector(int32_t n, const double* d, const double* in, double* out, void* stream)
{
    if(n <= 0) return;
    k_scale_vector<<<(n + 255) / 256, 256, 0, (hipStream_t) stream>>>(n, d, in, out);
}

}  // namespace culip
