// Batched SQMR over many independent systems: one launch per step of the
// recurrence covers every listed system.
//
// A launch is a list of work items {system, local index}. Each item runs the
// body of the single-system kernel at the grid position (local index, the
// system's own grid size) it would have there, on that system's buffers, so
// that every system computes exactly what it computes when it is solved
// alone: same partial sums in the same order, same bits.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace culip {

// One system as the batched kernels see it (device pointers).
struct BatchSys {
    int32_t n = 0;        // rows
    int32_t G = 0;        // grid of the SQMR vector kernels (sqmr_grid(n))
    int32_t nb = 0;       // block rows of the factor
    int32_t brun_rhs = 0; // the current SQMR run solves for rhs (else for b)
    // SQMR work vectors (as SqmrWork)
    double* v[3] = {};
    double* w[3] = {};
    double* p[3] = {};
    double* z = nullptr;
    double* tv = nullptr;
    double* ev = nullptr;
    double* av[2] = {};
    double* xt = nullptr;
    double* rt = nullptr;
    double* xacc = nullptr;
    double* rhs = nullptr;
    double* part = nullptr;
    double* scal = nullptr;     // 64 doubles: the recurrence's scalars, [40..41] norms
    int32_t* flags = nullptr;   // [0] breakdown, [1] quasi-residual test
    int32_t* iter = nullptr;    // iteration of the current run
    // D of the factor (tridiagonal form and pivot starts)
    const double* tri = nullptr;
    const int32_t* ip = nullptr;
    // K (CSR) for the residual product and the fused first sweep
    const int32_t* k_rp = nullptr;
    const int32_t* k_col = nullptr;
    const double* k_val = nullptr;
    const double* b = nullptr;  // right-hand side
    double* x = nullptr;        // solution
    double* jb[3] = {};         // block-Jacobi vectors: b, two sweep outputs
    // block-Jacobi solve layout, [0] the U solve, [1] the U' solve; values
    // in the factor's type
    const int4* task[2] = {};
    const int32_t* o_col[2] = {};
    const void* o_val[2] = {};
    const uint8_t* o_lr[2] = {};
    const int32_t* i_col[2] = {};
    const void* i_val[2] = {};
    const uint8_t* i_lr[2] = {};
};

// vectors the generic batched steps address
enum BatchVec : int32_t {
    BV_B = 0,     // right-hand side
    BV_X,         // solution
    BV_RT,        // residual
    BV_XACC,      // accumulated solution of a refinement
    BV_RHS,       // right-hand side of a refinement run
    BV_TV,
    BV_EV,
    BV_XT,
    BV_BRUN,      // right-hand side of the current run (b or rhs)
    BV_AVCUR,     // av[iter & 1]
    BV_AVNXT,     // av[(iter & 1) ^ 1]
    BV_JB0,
    BV_JB1,
    BV_JB2,
    BV_NONE = -1
};

// rows of K per item of the batched CSR product, block rows per item of the
// batched block-Jacobi sweep
constexpr int BATCH_SPMV_ROWS = 128;
constexpr int JACOBI_BATCH_ROWS = 4;

// XCD-aware lists: a 1D grid's workgroups go to the XCDs round-robin
// (workgroup w to XCD w % BATCH_XCDS). The lists of the product and of the
// sweeps give each system one XCD: item t of a group sits at position
// BATCH_XCDS t + group, so every XCD runs the items of its own systems and
// keeps their gathered x in its own L2. The groups are balanced by item
// count; positions past a group's end hold {-1, 0}, which the kernels skip.
// Below BATCH_XCD_MIN_SYSTEMS running systems the lists stay plain (one
// system on one XCD would leave seven idle).
constexpr int BATCH_XCDS = 8;
constexpr int BATCH_XCD_MIN_SYSTEMS = 64;

// Work items of one kind for a set of systems: items[i] = {system, local}.
struct BatchList {
    const int2* items = nullptr;
    int32_t count = 0;
};

// the lists of one set of systems, one per item kind
struct BatchLists {
    BatchList vec;    // sqmr_grid(n) items of 256 threads (vector kernels)
    BatchList row;    // ceil(n / 256) items of 256 threads (elementwise)
    BatchList blk;    // ceil(nb / JACOBI_BATCH_ROWS) items of 64 threads (block-Jacobi)
    BatchList spmv;   // ceil(n / BATCH_SPMV_ROWS) items of 256 threads (CSR product)
    BatchList sys;    // one item per system (scalars)
    // every listed system's vector grid covers its rows in one round
    // (n <= G * 256): the vector kernels then hold one element per thread
    // instead of a round of four (same arithmetic, fewer registers)
    bool one_round = false;
};

// the vector `id` of system s
__device__ inline double* batch_vec(const BatchSys& s, int32_t id)
{
    switch(id) {
    case BV_B: return const_cast<double*>(s.b);
    case BV_X: return s.x;
    case BV_RT: return s.rt;
    case BV_XACC: return s.xacc;
    case BV_RHS: return s.rhs;
    case BV_TV: return s.tv;
    case BV_EV: return s.ev;
    case BV_XT: return s.xt;
    case BV_BRUN: return s.brun_rhs ? s.rhs : const_cast<double*>(s.b);
    case BV_AVCUR: return s.av[*s.iter & 1];
    case BV_AVNXT: return s.av[(*s.iter & 1) ^ 1];
    case BV_JB0: return s.jb[0];
    case BV_JB1: return s.jb[1];
    case BV_JB2: return s.jb[2];
    default: return nullptr;
    }
}

// ent[i] = {system, count, offset, group}: items[offset + j] = {system, j}
// (plain), or with interleaved_total > 0 items[BATCH_XCDS (offset + j) +
// group] = {system, j} after every one of the interleaved_total items is
// set to {-1, 0}
void batch_expand(const int4* ent, int32_t count, int2* items, int32_t interleaved_total,
    hipStream_t st);

// the steps of the recurrence (sqmr.hip)
enum BatchStep : int32_t {
    BS_ALPHA, BS_VNEXT, BS_VSCALE, BS_BETA, BS_PUPDATE, BS_SCALARS,
    BS_INIT_FILL,     // the run's vectors to zero
    BS_SUMSQ,         // part slot 6 = ||a||^2
    BS_INIT_V, BS_INIT_W,
    BS_NORM2_PAIR,    // part slots 0, 1 = ||a||^2, ||b||^2
    BS_NORM_FINAL,    // scal[40], scal[41] = their square roots
    BS_COPY,          // b = a
    BS_DSOLVE,        // b = D \ a
    BS_RESIDUAL,      // b = a - b
    BS_AXPY1,         // b += a
    BS_ZERO           // a = 0
};
void batch_step(BatchStep s, const BatchLists& L, const BatchSys* sys, int32_t a, int32_t b,
    double tol, double brel, hipStream_t st);

// one block-Jacobi apply on vector `target` of every listed system
// (solve.hip): copy target to jb[0], then `sweeps` sweeps, the last one
// into target. With product, jb[0] = K src first (the fused product).
template<typename VT>
void batch_jacobi(const BatchLists& L, const BatchSys* sys, bool transpose, int32_t sweeps,
    int32_t target, int32_t product_src, hipStream_t st);

// y = K x for every listed system (spmv.hip)
void batch_csr_spmv(const BatchLists& L, const BatchSys* sys, int32_t x, int32_t y,
    hipStream_t st);

}  // namespace culip
