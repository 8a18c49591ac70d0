// SQMR on the device, FP64 recurrences.
//
// Reference: libs/la/iterative_solvers/sqmr.impl.cuh:109-355. The reference
// issues per iteration 4 dots, 2 nrm2, 5 axpy, 4 scal, 3 D-solves, 6 copies
// and ~22 single-thread scalar kernels. Here one iteration is six vector
// kernels plus the preconditioned operator:
//   k_alpha   alpha's two dots; D \ Av is formed on the fly, never stored
//   k_vnext   v_{k+1} = Av - alpha v_k - beta_k v_{k-1}, ||v_{k+1}||^2
//   k_vscale  v_{k+1} /= delta, z = D \ v_{k+1}, w_{k+1} = eta delta z, ev = z
//   [right(ev); ev2 = A ev; left(ev2)]                   (Av for k + 1)
//   k_beta    beta's two dots
//   k_pupdate w_{k+1} /= beta, p_k, x += tau p_k, ||x||^2
//   k_scalars the scalar recurrence, quasi-residual and the stopping test
// D \ v_k is reused from the previous iteration (the reference recomputes the
// same vector). Every workgroup reduces the per-workgroup partial sums in the
// same fixed order, so all scalars are identical everywhere and run to run.
// The host checks the stopping flags with the reference's 4-iteration
// throttle (:309-331).
#include "sqmr.hpp"
#include "batch.hpp"
#include "wave.hpp"
#include "../host/common.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

namespace culip {

namespace {

constexpr int TPB = 256;

// scalar slots (sqmr.impl.cuh:138-156)
enum {
    S_ALPHA = 0, S_BETA = 1, S_DELTA = 4, S_NORMA = 7, S_EPS = 8, S_BROT = 9, S_BRR = 10,
    S_AROT = 11, S_C = 12, S_S = 15, S_TAU = 18, S_RHO = 19, S_GAMMA = 20, S_NORMPB = 21,
    S_NORMX = 22, S_ETA = 26, S_ITER = 27, NSCAL = 32
};
static_assert(S_GAMMA == SQMR_S_GAMMA && S_NORMPB == SQMR_S_NORMPB, "scalar slots");

#define HIPCHK(x)                                                                       \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if(e_ != hipSuccess) fail(CULIP_ERR_HIP, std::string("HIP: ") + hipGetErrorString(e_)); \
    } while(0)

__device__ double block_sum(double v, double* sh)
{
    const int tid = threadIdx.x;
    v = wave_sum(v);
    __syncthreads();
    if((tid & 63) == 0) sh[tid >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if(tid == 0) {
        for(int q = 0; q < (int) (blockDim.x >> 6); ++q) t += sh[q];
        sh[8] = t;
    }
    __syncthreads();
    t = sh[8];
    __syncthreads();
    return t;
}

// Partial sums: slot s of workgroup q at part[s * G + q] (G workgroups), so
// that a slot's partials are contiguous. Every workgroup sums the partials
// of a slot the same way.
__device__ double reduce_part(const double* part, int G, int slot, double* sh)
{
    double v = 0.0;
    for(int q = threadIdx.x; q < G; q += blockDim.x) v += part[slot * G + q];
    return block_sum(v, sh);
}

// reduce_part for K slots at once: the loads of all slots are issued
// together, and every slot is summed exactly as reduce_part sums it (per
// thread in q order, wave sums, then the wave totals in order). sh holds at
// least K * (waves + 1) doubles.
template<int K>
__device__ void reduce_parts(const double* part, int G, const int (&slot)[K], double (&out)[K],
    double* sh)
{
    const int tid = threadIdx.x, nw = (int) (blockDim.x >> 6);
    double v[K];
#pragma unroll
    for(int k = 0; k < K; ++k) v[k] = 0.0;
    for(int q = tid; q < G; q += blockDim.x) {
        double x[K];
#pragma unroll
        for(int k = 0; k < K; ++k) x[k] = part[slot[k] * G + q];
#pragma unroll
        for(int k = 0; k < K; ++k) v[k] += x[k];
    }
#pragma unroll
    for(int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if((tid & 63) == 0) {
#pragma unroll
        for(int k = 0; k < K; ++k) sh[k * nw + (tid >> 6)] = v[k];
    }
    __syncthreads();
    if(tid < K) {
        double t = 0.0;
        for(int q = 0; q < nw; ++q) t += sh[tid * nw + q];
        sh[K * nw + tid] = t;
    }
    __syncthreads();
#pragma unroll
    for(int k = 0; k < K; ++k) out[k] = sh[K * nw + k];
    __syncthreads();
}

// block_sum of two values at once (each summed exactly as block_sum sums it)
__device__ void block_sum2(double& a, double& b, double* sh)
{
    const int tid = threadIdx.x, nw = (int) (blockDim.x >> 6);
    a = wave_sum(a);
    b = wave_sum(b);
    __syncthreads();
    if((tid & 63) == 0) { sh[tid >> 6] = a; sh[nw + (tid >> 6)] = b; }
    __syncthreads();
    if(tid < 2) {
        double t = 0.0;
        for(int q = 0; q < nw; ++q) t += sh[tid * nw + q];
        sh[2 * nw + tid] = t;
    }
    __syncthreads();
    a = sh[2 * nw];
    b = sh[2 * nw + 1];
    __syncthreads();
}

__device__ __forceinline__ double& part_at(double* part, int slot, int q, int G)
{
    return part[slot * G + q];
}

struct DP { double g1, g2; bool w1, w2; };

// k_solve_D at row g (helper_kernels.impl.cuh:720-772)
__device__ __forceinline__ DP dsolve(int32_t g, int32_t n, const double* tri, const int32_t* ip,
    double r1, double r2)
{
    DP o;
    if(!tri) { o.g1 = r1; o.g2 = 0; o.w1 = true; o.w2 = false; return o; }
    const int this_piv = ip[g];
    const int next_piv = (g < n - 1) ? ip[g + 1] : 1;
    const double p11 = tri[3 * (int64_t) g + 1];
    const double p12 = tri[3 * (int64_t) g + 2];
    const double p22 = !next_piv ? tri[3 * (int64_t) (g + 1) + 1] : 1.0;
    const double gr = sqrt(fma(p11, p11, p12 * p12));
    const double c = p11 / gr, s = p12 / gr;
    double g1 = fma(c, r1, s * r2);
    double g2 = fma(-s, r1, c * r2);
    const double t12 = fma(c, p12, s * p22);
    const double t22 = fma(-s, p12, c * p22);
    g2 = g2 / t22;
    g1 = fma(-t12, g2, g1) / gr;
    o.g1 = g1; o.g2 = g2; o.w1 = this_piv != 0; o.w2 = !next_piv;
    return o;
}

__device__ __forceinline__ double sgn(double v) { return v < 0 ? -1.0 : 1.0; }

// k_sym_ortho (helper_kernels.impl.cuh:662-699)
__device__ void sym_ortho(double a, double b, double& c, double& s, double& r)
{
    if(b == 0) { s = 0; r = fabs(a); c = (a == 0) ? 1 : sgn(a); }
    else if(a == 0) { c = 0; s = sgn(b); r = fabs(b); }
    else if(fabs(b) >= fabs(a)) {
        const double tao = a / b;
        s = sgn(b) / sqrt(1 + tao * tao);
        c = s * tao;
        r = b / s;
    } else {
        const double tao = b / a;
        c = sgn(a) / sqrt(1 + tao * tao);
        s = c * tao;
        r = a / c;
    }
}

// scalar part of one iteration (sqmr.impl.cuh:192-283), on a private copy;
// K0 = k mod 3 is a template argument so that every index into S is a
// constant and S stays in registers
template<int K0>
__device__ __forceinline__ void scalar_step(double* S, double alpha, double delta1, double beta1)
{
    constexpr int k0 = K0, k1 = (K0 + 1) % 3, km1 = (K0 + 2) % 3;
    S[S_ALPHA] = alpha;
    S[S_DELTA + k1] = delta1;
    S[S_BETA + k1] = beta1;
    S[S_ETA] = (delta1 / beta1) * S[S_ETA];
    const double bk = S[S_BETA + k0];
    const double nrm3 = sqrt(bk * bk + alpha * alpha + delta1 * delta1);
    S[S_NORMA] = fmax(S[S_NORMA], nrm3);
    S[S_EPS] = S[S_S + km1] * bk;
    S[S_BROT] = (-S[S_C + km1]) * bk;
    S[S_BRR] = S[S_C + k0] * S[S_BROT] + S[S_S + k0] * alpha;
    S[S_AROT] = S[S_S + k0] * S[S_BROT] + (-S[S_C + k0]) * alpha;
    double c, s, r;
    sym_ortho(S[S_AROT], delta1, c, s, r);
    S[S_C + k1] = c;
    S[S_S + k1] = s;
    S[S_RHO] = r;
    S[S_TAU] = c * S[S_GAMMA];
}

// The vector kernels give each thread VR elements per round, g0 + r S
// (S = G * TPB), and issue every load of a round before the first use, so a
// round costs one memory round trip; per thread the elements are still
// visited in increasing g, so every partial sum keeps its order.
constexpr int VR = 4;

// what the D solve of row g reads (k_solve_D's inputs), loaded up front
struct DRow {
    double p11, p12, p22n;
    int this_piv, next_piv;
};

__device__ __forceinline__ DRow dload(int64_t g, int32_t n, const double* tri, const int32_t* ip)
{
    DRow d;
    const int64_t g1 = min(g + 1, (int64_t) n - 1);
    d.this_piv = ip[g];
    const int nxt = ip[g1];
    d.next_piv = (g < n - 1) ? nxt : 1;
    d.p11 = tri[3 * g + 1];
    d.p12 = tri[3 * g + 2];
    d.p22n = tri[3 * g1 + 1];
    return d;
}

// dsolve on loaded inputs (the same operations)
__device__ __forceinline__ DP dsolve_loaded(const DRow& d, double r1, double r2)
{
    DP o;
    const double p22 = !d.next_piv ? d.p22n : 1.0;
    const double gr = sqrt(fma(d.p11, d.p11, d.p12 * d.p12));
    const double c = d.p11 / gr, s = d.p12 / gr;
    double g1 = fma(c, r1, s * r2);
    double g2 = fma(-s, r1, c * r2);
    const double t12 = fma(c, d.p12, s * p22);
    const double t22 = fma(-s, d.p12, c * p22);
    g2 = g2 / t22;
    g1 = fma(-t12, g2, g1) / gr;
    o.g1 = g1; o.g2 = g2; o.w1 = d.this_piv != 0; o.w2 = !d.next_piv;
    return o;
}

template<int R = VR>
__device__ __forceinline__ void alpha_body(int q_, int G_, int32_t n, const double* Av, const double* v,
    const double* z, const double* tri, const int32_t* ip, double* part)
{
    __shared__ double sh[16];
    double num = 0.0, den = 0.0;
    const int64_t S = (int64_t) G_ * TPB;
    for(int64_t g0 = (int64_t) q_ * TPB + threadIdx.x; g0 < n; g0 += S * R) {
        double av[R], av1[R], vv[R], vv1[R], zz[R];
        DRow d[R];
#pragma unroll
        for(int r = 0; r < R; ++r) {
            const int64_t g = min(g0 + r * S, (int64_t) n - 1);
            const int64_t g1 = min(g + 1, (int64_t) n - 1);
            av[r] = Av[g];
            av1[r] = Av[g1];
            vv[r] = v[g];
            vv1[r] = v[g1];
            zz[r] = z[g];
            if(tri) d[r] = dload(g, n, tri, ip);
        }
#pragma unroll
        for(int r = 0; r < R; ++r) {
            const int64_t g = g0 + r * S;
            if(g >= n) break;
            DP o;
            if(tri) o = dsolve_loaded(d[r], av[r], (g < n - 1) ? av1[r] : 0.0);
            else { o.g1 = av[r]; o.g2 = 0; o.w1 = true; o.w2 = false; }
            if(o.w1) num = fma(vv[r], o.g1, num);
            if(o.w2) num = fma(vv1[r], o.g2, num);
            den = fma(vv[r], zz[r], den);
        }
    }
    block_sum2(num, den, sh);
    if(threadIdx.x == 0) { part_at(part, 0, q_, G_) = num; part_at(part, 1, q_, G_) = den; }
}

// The kernels that start with a reduction of the previous kernel's partial
// sums issue the loads of their first round before it, so its memory round
// trip overlaps the reduction's.
template<int R = VR>
__device__ __forceinline__ void vnext_body(int q_, int G_, int32_t n, int k, const double* Av,
    const double* vk, const double* vkm1, double* tv, const double* scal, double* part)
{
    __shared__ double sh[16];
    const int G = G_;
    const int64_t S = (int64_t) G * TPB;
    double a[R], b[R], c[R];
    auto load = [&](int64_t g0) {
#pragma unroll
        for(int r = 0; r < R; ++r) {
            const int64_t g = min(g0 + r * S, (int64_t) n - 1);
            a[r] = Av[g];
            b[r] = vk[g];
            c[r] = vkm1[g];
        }
    };
    int64_t g0 = (int64_t) q_ * TPB + threadIdx.x;
    if(g0 < n) load(g0);
    double rp[2];
    reduce_parts<2>(part, G, {0, 1}, rp, sh);
    const double alpha = rp[0] / rp[1];
    const double beta = scal[S_BETA + k % 3];
    double ss = 0.0;
    for(; g0 < n; g0 += S * R) {
#pragma unroll
        for(int r = 0; r < R; ++r) {
            const int64_t g = g0 + r * S;
            if(g >= n) break;
            const double t = fma(-beta, c[r], fma(-alpha, b[r], a[r]));
            tv[g] = t;
            ss = fma(t, t, ss);
        }
        if(g0 + S * R < n) load(g0 + S * R);
    }
    ss = block_sum(ss, sh);
    if(threadIdx.x == 0) part_at(part, 2, q_, G_) = ss;
}

template<int R = VR>
__device__ __forceinline__ void vscale_body(int q_, int G_, int32_t n, const double* tv, double* v1,
    double* z, double* w1, double* ev, const double* tri, const int32_t* ip, const double* scal,
    const double* part)
{
    __shared__ double sh[16];
    const int G = G_;
    const int64_t S = (int64_t) G * TPB;
    double a[R], a1[R];
    DRow d[R];
    auto load = [&](int64_t g0) {
#pragma unroll
        for(int r = 0; r < R; ++r) {
            const int64_t g = min(g0 + r * S, (int64_t) n - 1);
            a[r] = tv[g];
            a1[r] = tv[min(g + 1, (int64_t) n - 1)];
            if(tri) d[r] = dload(g, n, tri, ip);
        }
    };
    int64_t g0 = (int64_t) q_ * TPB + threadIdx.x;
    if(g0 < n) load(g0);
    const double delta = sqrt(reduce_part(part, G, 2, sh));
    const double t = 1.0 / delta;
    const double we = scal[S_ETA] * delta;
    for(; g0 < n; g0 += S * R) {
#pragma unroll
        for(int r = 0; r < R; ++r) {
            const int64_t g = g0 + r * S;
            if(g >= n) break;
            const double r1 = a[r] * t;
            const double r2 = (g < n - 1) ? a1[r] * t : 0.0;
            v1[g] = r1;
            DP o;
            if(tri) o = dsolve_loaded(d[r], r1, r2);
            else { o.g1 = r1; o.g2 = 0; o.w1 = true; o.w2 = false; }
            if(o.w1) { z[g] = o.g1; w1[g] = o.g1 * we; ev[g] = o.g1; }
            if(o.w2) { z[g + 1] = o.g2; w1[g + 1] = o.g2 * we; ev[g + 1] = o.g2; }
        }
        if(g0 + S * R < n) load(g0 + S * R);
    }
}

template<int R = VR>
__device__ __forceinline__ void beta_body(int q_, int G_, int32_t n, const double* wk, const double* Av,
    const double* vk, double* part)
{
    __shared__ double sh[16];
    double num = 0.0, den = 0.0;
    const int64_t S = (int64_t) G_ * TPB;
    for(int64_t g0 = (int64_t) q_ * TPB + threadIdx.x; g0 < n; g0 += S * R) {
        double a[R], b[R], c[R];
#pragma unroll
        for(int r = 0; r < R; ++r) {
            const int64_t g = min(g0 + r * S, (int64_t) n - 1);
            a[r] = wk[g];
            b[r] = Av[g];
            c[r] = vk[g];
        }
#pragma unroll
        for(int r = 0; r < R; ++r) {
            if(g0 + r * S >= n) break;
            num = fma(a[r], b[r], num);
            den = fma(a[r], c[r], den);
        }
    }
    block_sum2(num, den, sh);
    if(threadIdx.x == 0) { part_at(part, 3, q_, G_) = num; part_at(part, 4, q_, G_) = den; }
}

template<int K0>
__device__ __forceinline__ void iteration_scalars(double* S, const double* scal,
    const double (&r)[5])
{
#pragma unroll
    for(int q = 0; q < NSCAL; ++q) S[q] = scal[q];
    const double alpha = r[0] / r[1];
    const double delta = sqrt(r[2]);
    const double beta = r[3] / r[4];
    scalar_step<K0>(S, alpha, delta, beta);
}

template<int K0, int R>
__device__ __forceinline__ void pupdate(int q_, int G_, int32_t n, double* w1, const double* vk, double* pk,
    const double* pkm1, const double* pkm2, double* x, const double* scal, double* part,
    const int32_t* flags)
{
    __shared__ double sh[64];
    const int G = G_;
    const int64_t SG = (int64_t) G * TPB;
    double ww[R], a[R], b[R], c[R], xx[R];
    auto load = [&](int64_t g0) {
#pragma unroll
        for(int r = 0; r < R; ++r) {
            const int64_t g = min(g0 + r * SG, (int64_t) n - 1);
            ww[r] = w1[g];
            a[r] = vk[g];
            b[r] = pkm1[g];
            c[r] = pkm2[g];
            xx[r] = x[g];
        }
    };
    int64_t g0 = (int64_t) q_ * TPB + threadIdx.x;
    if(g0 < n) load(g0);
    double rp[5];
    reduce_parts<5>(part, G, {0, 1, 2, 3, 4}, rp, sh);
    // one thread runs the scalar step and shares the five factors the
    // update needs (the step's 32 scalars held by every thread would cost
    // the registers of two more waves per SIMD)
    if(threadIdx.x == 0) {
        double S[NSCAL];
        iteration_scalars<K0>(S, scal, rp);
        sh[32] = 1.0 / S[S_BETA + (K0 + 1) % 3];
        sh[33] = -S[S_BRR];
        sh[34] = -S[S_EPS];
        sh[35] = 1.0 / S[S_RHO];
        sh[36] = S[S_TAU];
    }
    __syncthreads();
    const double ib = sh[32], nbrr = sh[33], neps = sh[34], ir = sh[35], tau = sh[36];
    // after a Lanczos breakdown (flagged by the previous iteration) the
    // recurrence divides by a vanishing delta: x keeps the value it had when
    // the breakdown was detected
    const bool frozen = flags[0] != 0;
    double ss = 0.0;
    for(; g0 < n; g0 += SG * R) {
#pragma unroll
        for(int r = 0; r < R; ++r) {
            const int64_t g = g0 + r * SG;
            if(g >= n) break;
            w1[g] = ww[r] * ib;
            double p = a[r];
            p = fma(nbrr, b[r], p);
            p = fma(neps, c[r], p);
            p = p * ir;
            pk[g] = p;
            const double xv = frozen ? xx[r] : fma(tau, p, xx[r]);
            if(!frozen) x[g] = xv;
            ss = fma(xv, xv, ss);
        }
        if(g0 + SG * R < n) load(g0 + SG * R);
    }
    ss = block_sum(ss, sh);
    if(threadIdx.x == 0) part_at(part, 5, q_, G_) = ss;
}

template<int R = VR>
__device__ __forceinline__ void pupdate_body(int q_, int G_, int32_t n, int k, double* w1, const double* vk,
    double* pk, const double* pkm1, const double* pkm2, double* x, const double* scal,
    double* part, const int32_t* flags)
{
    switch(k % 3) {
    case 0: pupdate<0, R>(q_, G_, n, w1, vk, pk, pkm1, pkm2, x, scal, part, flags); break;
    case 1: pupdate<1, R>(q_, G_, n, w1, vk, pk, pkm1, pkm2, x, scal, part, flags); break;
    default: pupdate<2, R>(q_, G_, n, w1, vk, pk, pkm1, pkm2, x, scal, part, flags); break;
    }
}

// k_sqmr_test (:89-107) and the state update; one workgroup
// brel < 0: the reference's breakdown test |delta| <= tol; brel >= 0:
// |delta| <= brel * ||A||est, scale-free (the delta of a badly scaled but
// healthy recurrence can fall below an absolute 1e-8)
template<int K0>
__device__ __forceinline__ void scalars(int k, int G, double* scal, const double* part,
    int32_t* flags, double tol, double brel)
{
    __shared__ double sh[64];
    double S[NSCAL];
    double r6[6];
    reduce_parts<6>(part, G, {0, 1, 2, 3, 4, 5}, r6, sh);
    const double r[5] = {r6[0], r6[1], r6[2], r6[3], r6[4]};
    iteration_scalars<K0>(S, scal, r);
    const double nx = sqrt(r6[5]);
    if(threadIdx.x == 0) {
        constexpr int k1 = (K0 + 1) % 3;
        S[S_NORMX] = nx;
        S[S_GAMMA] = S[S_S + k1] * S[S_GAMMA];
        const double dl = fabs(S[S_DELTA + k1]);
        flags[0] |= (brel < 0) ? (dl <= tol) : (dl <= brel * S[S_NORMA]);
        flags[1] |= (S[S_GAMMA] / (S[S_NORMA] * S[S_NORMX] + S[S_NORMPB]) <= tol);
        S[S_ITER] = (double) k;
#pragma unroll
        for(int q = 0; q < NSCAL; ++q) scal[q] = S[q];
    }
}

__global__ __launch_bounds__(TPB) void k_scalars(int k, int G, double* scal, const double* part,
    int32_t* flags, double tol, double brel)
{
    switch(k % 3) {
    case 0: scalars<0>(k, G, scal, part, flags, tol, brel); break;
    case 1: scalars<1>(k, G, scal, part, flags, tol, brel); break;
    default: scalars<2>(k, G, scal, part, flags, tol, brel); break;
    }
}

// start: tv = Pl \ b (done by the caller); v0 = tv / ||tv||, z = D \ v0
__device__ __forceinline__ void sumsq_body(int q_, int G_, int32_t n, const double* a, double* part, int slot)
{
    __shared__ double sh[16];
    double ss = 0.0;
    for(int32_t g = q_ * TPB + threadIdx.x; g < n; g += G_ * TPB) ss = fma(a[g], a[g], ss);
    ss = block_sum(ss, sh);
    if(threadIdx.x == 0) part_at(part, slot, q_, G_) = ss;
}

__device__ __forceinline__ void init_v_body(int q_, int G_, int32_t n, const double* tv, double* v0, double* z,
    const double* tri, const int32_t* ip, double* part)
{
    __shared__ double sh[16];
    const int G = G_;
    const double t = 1.0 / sqrt(reduce_part(part, G, 6, sh));
    double ss = 0.0;
    for(int32_t g = q_ * TPB + threadIdx.x; g < n; g += G * TPB) {
        const double r1 = tv[g] * t;
        const double r2 = (g < n - 1) ? tv[g + 1] * t : 0.0;
        v0[g] = r1;
        const DP d = dsolve(g, n, tri, ip, r1, r2);
        if(d.w1) { z[g] = d.g1; ss = fma(d.g1, d.g1, ss); }
        if(d.w2) { z[g + 1] = d.g2; ss = fma(d.g2, d.g2, ss); }
    }
    ss = block_sum(ss, sh);
    if(threadIdx.x == 0) part_at(part, 7, q_, G_) = ss;
}

__device__ __forceinline__ void init_w_body(int q_, int G_, int32_t n, const double* z, double* w0, double* ev,
    double* scal, const double* part, int32_t* flags)
{
    __shared__ double sh[16];
    const int G = G_;
    const double norm_pb = sqrt(reduce_part(part, G, 6, sh));
    const double eta = 1.0 / sqrt(reduce_part(part, G, 7, sh));
    for(int32_t g = q_ * TPB + threadIdx.x; g < n; g += G * TPB) {
        w0[g] = z[g] * eta;
        ev[g] = z[g];
    }
    if(q_ == 0 && threadIdx.x == 0) {
        for(int q = 0; q < NSCAL; ++q) scal[q] = 0.0;
        scal[S_C + 0] = -1.0;
        scal[23] = 1.0;
        scal[S_NORMPB] = norm_pb;
        scal[S_GAMMA] = norm_pb;
        scal[S_ETA] = eta;
        flags[0] = flags[1] = 0;
    }
}

__device__ __forceinline__ void fill_body(int q_, int G_, int32_t n, double* a, double v)
{
    const int32_t g = q_ * blockDim.x + threadIdx.x;
    if(g < n) a[g] = v;
}

__device__ __forceinline__ void copy_body(int q_, int G_, int32_t n, const double* a, double* b)
{
    const int32_t g = q_ * blockDim.x + threadIdx.x;
    if(g < n) b[g] = a[g];
}

// b <- D \ a (out of place)
__device__ __forceinline__ void dsolve_body(int q_, int G_, int32_t n, const double* a, double* b, const double* tri, const int32_t* ip)
{
    const int32_t g = q_ * blockDim.x + threadIdx.x;
    if(g >= n) return;
    const DP d = dsolve(g, n, tri, ip, a[g], (g < n - 1) ? a[g + 1] : 0.0);
    if(d.w1) b[g] = d.g1;
    if(d.w2) b[g + 1] = d.g2;
}

// r = b - r (r holds A x); x_acc += x
__device__ __forceinline__ void residual_body(int q_, int G_, int32_t n, const double* b, double* r)
{
    const int32_t g = q_ * blockDim.x + threadIdx.x;
    if(g < n) r[g] = b[g] - r[g];
}

__device__ __forceinline__ void axpy1_body(int q_, int G_, int32_t n, const double* a, double* b)
{
    const int32_t g = q_ * blockDim.x + threadIdx.x;
    if(g < n) b[g] += a[g];
}

// sums of squares of a and b, each element divided by s[g] when s is set
// (the norms of the unscaled system when a and b belong to D A D)
__device__ __forceinline__ void norm2_pair_body(int q_, int G_, int32_t n, const double* a, const double* b,
    const double* s, double* part)
{
    __shared__ double sh[16];
    double sa = 0.0, sb = 0.0;
    for(int32_t g = q_ * TPB + threadIdx.x; g < n; g += G_ * TPB) {
        double av = a[g], bv = b[g];
        if(s) {
            av /= s[g];
            bv /= s[g];
        }
        sa = fma(av, av, sa);
        sb = fma(bv, bv, sb);
    }
    sa = block_sum(sa, sh);
    sb = block_sum(sb, sh);
    if(threadIdx.x == 0) { part_at(part, 0, q_, G_) = sa; part_at(part, 1, q_, G_) = sb; }
}

// the single-system kernels: the bodies at their own grid position; R = 1
// when the grid covers the rows in one round (fewer registers, same
// arithmetic), else VR
template<int R>
__global__ __launch_bounds__(TPB) void k_alpha(int32_t n, const double* Av, const double* v,
    const double* z, const double* tri, const int32_t* ip, double* part)
{
    alpha_body<R>(blockIdx.x, gridDim.x, n, Av, v, z, tri, ip, part);
}

template<int R>
__global__ __launch_bounds__(TPB) void k_vnext(int32_t n, int k, const double* Av,
    const double* vk, const double* vkm1, double* tv, const double* scal, double* part)
{
    vnext_body<R>(blockIdx.x, gridDim.x, n, k, Av, vk, vkm1, tv, scal, part);
}

template<int R>
__global__ __launch_bounds__(TPB) void k_vscale(int32_t n, const double* tv, double* v1,
    double* z, double* w1, double* ev, const double* tri, const int32_t* ip, const double* scal,
    const double* part)
{
    vscale_body<R>(blockIdx.x, gridDim.x, n, tv, v1, z, w1, ev, tri, ip, scal, part);
}

template<int R>
__global__ __launch_bounds__(TPB) void k_beta(int32_t n, const double* wk, const double* Av,
    const double* vk, double* part)
{
    beta_body<R>(blockIdx.x, gridDim.x, n, wk, Av, vk, part);
}

template<int R>
__global__ __launch_bounds__(TPB) void k_pupdate(int32_t n, int k, double* w1, const double* vk,
    double* pk, const double* pkm1, const double* pkm2, double* x, const double* scal,
    double* part, const int32_t* flags)
{
    pupdate_body<R>(blockIdx.x, gridDim.x, n, k, w1, vk, pk, pkm1, pkm2, x, scal, part, flags);
}

__global__ __launch_bounds__(TPB) void k_sumsq(int32_t n, const double* a, double* part, int slot)
{
    sumsq_body(blockIdx.x, gridDim.x, n, a, part, slot);
}

__global__ __launch_bounds__(TPB) void k_init_v(int32_t n, const double* tv, double* v0, double* z,
    const double* tri, const int32_t* ip, double* part)
{
    init_v_body(blockIdx.x, gridDim.x, n, tv, v0, z, tri, ip, part);
}

__global__ __launch_bounds__(TPB) void k_init_w(int32_t n, const double* z, double* w0, double* ev,
    double* scal, const double* part, int32_t* flags)
{
    init_w_body(blockIdx.x, gridDim.x, n, z, w0, ev, scal, part, flags);
}

__global__ void k_fill(int32_t n, double* a, double v)
{
    fill_body(blockIdx.x, gridDim.x, n, a, v);
}

__global__ void k_copy(int32_t n, const double* a, double* b)
{
    copy_body(blockIdx.x, gridDim.x, n, a, b);
}

__global__ void k_dsolve(int32_t n, const double* a, double* b, const double* tri, const int32_t* ip)
{
    dsolve_body(blockIdx.x, gridDim.x, n, a, b, tri, ip);
}

__global__ void k_residual(int32_t n, const double* b, double* r)
{
    residual_body(blockIdx.x, gridDim.x, n, b, r);
}

__global__ void k_axpy1(int32_t n, const double* a, double* b)
{
    axpy1_body(blockIdx.x, gridDim.x, n, a, b);
}

__global__ __launch_bounds__(TPB) void k_norm2_pair(int32_t n, const double* a, const double* b,
    const double* s, double* part)
{
    norm2_pair_body(blockIdx.x, gridDim.x, n, a, b, s, part);
}

__device__ __forceinline__ void norm_final_body(int G, const double* part, double* out)
{
    __shared__ double sh[16];
    const double a = reduce_part(part, G, 0, sh);
    const double b = reduce_part(part, G, 1, sh);
    if(threadIdx.x == 0) { out[0] = sqrt(a); out[1] = sqrt(b); }
}

__global__ __launch_bounds__(TPB) void k_norm_final(int G, const double* part, double* out)
{
    norm_final_body(G, part, out);
}

// ---------------------------------------------------------------------------
// Batched steps: one launch over the items of a list; each item is the
// single-system kernel's workgroup `it.y` of system `it.x` (the iteration's
// cyclic vectors chosen from the system's own iteration counter).

__global__ void k_batch_expand(const int4* ent, int2* items, int32_t interleaved)
{
    const int4 e = ent[blockIdx.x];
    for(int j = threadIdx.x; j < e.y; j += blockDim.x) {
        const int64_t pos = interleaved ? (int64_t) BATCH_XCDS * (e.z + j) + e.w : e.z + j;
        items[pos] = make_int2(e.x, j);
    }
}

__global__ void k_batch_holes(int2* items, int32_t total)
{
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < total) items[i] = make_int2(-1, 0);
}

template<int R>
__global__ __launch_bounds__(TPB) void bk_alpha(const int2* items, const BatchSys* sys)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    const int k = *S.iter;
    alpha_body<R>(it.y, S.G, S.n, S.av[k & 1], S.v[k % 3], S.z, S.tri, S.ip, S.part);
}

template<int R>
__global__ __launch_bounds__(TPB) void bk_vnext(const int2* items, const BatchSys* sys)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    const int k = *S.iter;
    vnext_body<R>(it.y, S.G, S.n, k, S.av[k & 1], S.v[k % 3], S.v[(k + 2) % 3], S.tv, S.scal,
        S.part);
}

template<int R>
__global__ __launch_bounds__(TPB) void bk_vscale(const int2* items, const BatchSys* sys)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    const int k1 = (*S.iter + 1) % 3;
    vscale_body<R>(it.y, S.G, S.n, S.tv, S.v[k1], S.z, S.w[k1], S.ev, S.tri, S.ip, S.scal, S.part);
}

template<int R>
__global__ __launch_bounds__(TPB) void bk_beta(const int2* items, const BatchSys* sys)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    const int k = *S.iter;
    beta_body<R>(it.y, S.G, S.n, S.w[k % 3], S.av[(k & 1) ^ 1], S.v[k % 3], S.part);
}

template<int R>
__global__ __launch_bounds__(TPB) void bk_pupdate(const int2* items, const BatchSys* sys)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    const int k = *S.iter;
    const int k0 = k % 3, k1 = (k + 1) % 3, km1 = (k + 2) % 3;
    pupdate_body<R>(it.y, S.G, S.n, k, S.w[k1], S.v[k0], S.p[k0], S.p[km1], S.p[k1], S.xt, S.scal,
        S.part, S.flags);
}

// the scalar recurrence of every listed system, then its iteration counter
// moves on (every thread has read it before the first barrier in scalars)
__global__ __launch_bounds__(TPB) void bk_scalars(const int2* items, const BatchSys* sys,
    double tol, double brel)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    const int k = *S.iter;
    switch(k % 3) {
    case 0: scalars<0>(k, S.G, S.scal, S.part, S.flags, tol, brel); break;
    case 1: scalars<1>(k, S.G, S.scal, S.part, S.flags, tol, brel); break;
    default: scalars<2>(k, S.G, S.scal, S.part, S.flags, tol, brel); break;
    }
    if(threadIdx.x == 0) *S.iter = k + 1;
}

// the vectors a run starts from zero (sqmr.impl.cuh:126-135)
__global__ void bk_init_fill(const int2* items, const BatchSys* sys)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    for(double* a : {S.v[0], S.v[1], S.v[2], S.w[0], S.w[1], S.w[2], S.p[0], S.p[1], S.p[2], S.xt})
        fill_body(it.y, 0, S.n, a, 0.0);
}

__global__ __launch_bounds__(TPB) void bk_sumsq(const int2* items, const BatchSys* sys, int32_t a)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    sumsq_body(it.y, S.G, S.n, batch_vec(S, a), S.part, 6);
}

__global__ __launch_bounds__(TPB) void bk_init_v(const int2* items, const BatchSys* sys)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    init_v_body(it.y, S.G, S.n, S.tv, S.v[0], S.z, S.tri, S.ip, S.part);
}

__global__ __launch_bounds__(TPB) void bk_init_w(const int2* items, const BatchSys* sys)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    init_w_body(it.y, S.G, S.n, S.z, S.w[0], S.ev, S.scal, S.part, S.flags);
    if(it.y == 0 && threadIdx.x == 0) *S.iter = 0;
}

__global__ __launch_bounds__(TPB) void bk_norm2_pair(const int2* items, const BatchSys* sys,
    int32_t a, int32_t b)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    norm2_pair_body(it.y, S.G, S.n, batch_vec(S, a), batch_vec(S, b), nullptr, S.part);
}

__global__ __launch_bounds__(TPB) void bk_norm_final(const int2* items, const BatchSys* sys)
{
    const BatchSys& S = sys[items[blockIdx.x].x];
    norm_final_body(S.G, S.part, S.scal + 40);
}

// elementwise steps over 256-row chunks; op 0 copy b = a, 1 b = D \ a,
// 2 b = a - b, 3 b += a, 4 a = 0
__global__ void bk_elementwise(const int2* items, const BatchSys* sys, int32_t op, int32_t a,
    int32_t b)
{
    const int2 it = items[blockIdx.x];
    const BatchSys& S = sys[it.x];
    double* va = batch_vec(S, a);
    double* vb = batch_vec(S, b);
    switch(op) {
    case 0: copy_body(it.y, 0, S.n, va, vb); break;
    case 1: dsolve_body(it.y, 0, S.n, va, vb, S.tri, S.ip); break;
    case 2: residual_body(it.y, 0, S.n, va, vb); break;
    case 3: axpy1_body(it.y, 0, S.n, va, vb); break;
    default: fill_body(it.y, 0, S.n, va, 0.0); break;
    }
}

inline unsigned blocks_for(int32_t n) { return (unsigned) ((n + 255) / 256); }

}  // namespace

int32_t sqmr_grid(int32_t n)
{
    int32_t g = (n + TPB - 1) / TPB;
    if(g > 512) g = 512;
    if(g < 1) g = 1;
    return g;
}

static size_t align_up(size_t v) { return (v + 255) & ~(size_t) 255; }

size_t sqmr_work_bytes(int32_t n)
{
    const size_t vec = align_up(sizeof(double) * (size_t) (n > 0 ? n : 1));
    return 19 * vec + align_up(sizeof(double) * 8 * sqmr_grid(n)) + align_up(sizeof(double) * 64) +
        align_up(sizeof(int32_t) * 8);
}

void sqmr_work_bind(SqmrWork& w, int32_t n, char* base, int32_t* host_flags, double* host_scal)
{
    const size_t vec = align_up(sizeof(double) * (size_t) (n > 0 ? n : 1));
    w.n = n;
    w.grid = sqmr_grid(n);
    double** vecs[18] = {&w.v[0], &w.v[1], &w.v[2], &w.w[0], &w.w[1], &w.w[2], &w.p[0], &w.p[1],
        &w.p[2], &w.z, &w.tv, &w.ev, &w.av[0], &w.av[1], &w.xt, &w.rt, &w.xacc, &w.rhs};
    size_t off = 0;
    for(int q = 0; q < 18; ++q) { *vecs[q] = (double*) (base + off); off += vec; }
    off += vec;  // spare
    w.part = (double*) (base + off);
    off += align_up(sizeof(double) * 8 * w.grid);
    w.scal = (double*) (base + off);
    off += align_up(sizeof(double) * 64);
    w.flags = (int32_t*) (base + off);
    w.host_flags = host_flags;
    w.host_scal = host_scal;
}

// ||b - A x|| and ||b|| in FP64; r = b - A x is left in w.rt
static void residual_norms(const SqmrOps& ops, SqmrWork& w, const double* b, const double* x,
    hipStream_t st, double& nr, double& nb)
{
    const int32_t n = ops.n;
    ops.spmv(x, w.rt, st);
    k_residual<<<blocks_for(n), 256, 0, st>>>(n, b, w.rt);
    k_norm2_pair<<<w.grid, TPB, 0, st>>>(n, w.rt, b, ops.res_scale, w.part);
    k_norm_final<<<1, TPB, 0, st>>>(w.grid, w.part, w.scal + 40);
    HIPCHK(hipMemcpyAsync(w.host_scal + 40, w.scal + 40, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    nr = w.host_scal[40];
    nb = w.host_scal[41];
}

double true_rel_residual(const SqmrOps& ops, SqmrWork& w, const double* b, const double* x,
    void* stream)
{
    double nr, nb;
    residual_norms(ops, w, b, x, (hipStream_t) stream, nr, nb);
    return nb > 0 ? nr / nb : nr;
}

// One SQMR run from x = 0 (sqmr.impl.cuh:109-355). Returns the solution
// x = Pr \ (D \ y) in xout. check_every > 0 adds the true-residual gate.
static void sqmr_run(const SqmrOps& ops, SqmrWork& w, const double* b, double* xout,
    const culip_sqmr_opts& opt, culip_sqmr_result& res, hipStream_t st, int32_t& total_it,
    bool& gate_met, double gate_abs, SqmrGraphs* graphs)
{
    const int32_t n = ops.n;
    const int G = w.grid;
    const unsigned nb = blocks_for(n);
    gate_met = false;
    // x (y-iterate) lives in w.xt; vectors start at zero (:126-135)
    for(double* a : {w.v[0], w.v[1], w.v[2], w.w[0], w.w[1], w.w[2], w.p[0], w.p[1], w.p[2], w.xt})
        k_fill<<<nb, 256, 0, st>>>(n, a, 0.0);
    // v0 = Pl \ b normalised; w0 = eta D \ v0; Av = Pl \ A Pr \ D \ v0
    k_copy<<<nb, 256, 0, st>>>(n, b, w.tv);
    if(ops.left) ops.left(w.tv, st);
    k_sumsq<<<G, TPB, 0, st>>>(n, w.tv, w.part, 6);
    k_init_v<<<G, TPB, 0, st>>>(n, w.tv, w.v[0], w.z, ops.tri, ops.is_piv, w.part);
    k_init_w<<<G, TPB, 0, st>>>(n, w.z, w.w[0], w.ev, w.scal, w.part, w.flags);
    // Av = Pl \ (A x), folded into one pass where the preconditioner can
    auto apply_left_op = [&](const double* xin, double* yout) {
        if(ops.spmv_left && ops.spmv_left(xin, yout, st)) return;
        ops.spmv(xin, yout, st);
        if(ops.left) ops.left(yout, st);
    };
    int cur = 0;
    if(ops.right) ops.right(w.ev, st);
    apply_left_op(w.ev, w.av[cur]);

    hipEvent_t ev_copy;
    HIPCHK(hipEventCreateWithFlags(&ev_copy, hipEventDisableTiming));
    int32_t it = 0;
    const int period = opt.true_residual_period;
    // long plateaus of the true residual are normal on barrier-conditioned
    // systems (a few hundred iterations); restarting inside one throws the
    // Krylov space away
    constexpr int STALL_PERIODS = SQMR_STALL_PERIODS;
    double best_nr = -1.0, gam_at_best = 0.0;
    int stall = 0;
    // the kernels of iteration `i`: they depend on i only through i mod 3
    // (the 3-cyclic vectors) and the parity of `cur` (the two A ev vectors),
    // so iterations i and i + 6 launch the same work
    const bool one_round = (int64_t) n <= (int64_t) G * TPB;
    auto launch_iteration = [&](int i) {
        const int k0 = i % 3, k1 = (i + 1) % 3, km1 = (i + 2) % 3, km2 = (i + 1) % 3;
        const int nxt = cur ^ 1;
        if(one_round) {
            k_alpha<1><<<G, TPB, 0, st>>>(n, w.av[cur], w.v[k0], w.z, ops.tri, ops.is_piv, w.part);
            k_vnext<1><<<G, TPB, 0, st>>>(n, i, w.av[cur], w.v[k0], w.v[km1], w.tv, w.scal, w.part);
            k_vscale<1><<<G, TPB, 0, st>>>(n, w.tv, w.v[k1], w.z, w.w[k1], w.ev, ops.tri,
                ops.is_piv, w.scal, w.part);
        } else {
            k_alpha<VR><<<G, TPB, 0, st>>>(n, w.av[cur], w.v[k0], w.z, ops.tri, ops.is_piv, w.part);
            k_vnext<VR><<<G, TPB, 0, st>>>(n, i, w.av[cur], w.v[k0], w.v[km1], w.tv, w.scal,
                w.part);
            k_vscale<VR><<<G, TPB, 0, st>>>(n, w.tv, w.v[k1], w.z, w.w[k1], w.ev, ops.tri,
                ops.is_piv, w.scal, w.part);
        }
        if(ops.right) ops.right(w.ev, st);
        apply_left_op(w.ev, w.av[nxt]);
        if(one_round) {
            k_beta<1><<<G, TPB, 0, st>>>(n, w.w[k0], w.av[nxt], w.v[k0], w.part);
            k_pupdate<1><<<G, TPB, 0, st>>>(n, i, w.w[k1], w.v[k0], w.p[k0], w.p[km1], w.p[km2],
                w.xt, w.scal, w.part, w.flags);
        } else {
            k_beta<VR><<<G, TPB, 0, st>>>(n, w.w[k0], w.av[nxt], w.v[k0], w.part);
            k_pupdate<VR><<<G, TPB, 0, st>>>(n, i, w.w[k1], w.v[k0], w.p[k0], w.p[km1], w.p[km2],
                w.xt, w.scal, w.part, w.flags);
        }
        k_scalars<<<1, TPB, 0, st>>>(i, G, w.scal, w.part, w.flags, opt.tol,
            period > 0 ? SQMR_BREAKDOWN_REL : -1.0);
        cur = nxt;
        HIPCHK(hipGetLastError());
    };
    // with the true-residual gate, the `period` iterations between two
    // checks can replay a captured graph (one per phase it mod 6; the phase
    // fixes every buffer the kernels use)
    hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
    const bool segments = graphs && period > 0 && st != nullptr &&
        hipStreamIsCapturing(st, &cst) == hipSuccess && cst == hipStreamCaptureStatusNone;
    if(segments && graphs->period != period) {
        graphs->release();
        graphs->period = period;
    }
    while(it < opt.max_it) {
        if(segments && it % period == 0 && it + period <= opt.max_it) {
            const int phase = it % 6;
            hipGraphExec_t& ex = graphs->seg[phase];
            const int cur0 = cur;
            if(!ex) {
                HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                try {
                    for(int j = 0; j < period; ++j) launch_iteration(it + j);
                } catch(...) {
                    // leave the stream usable: end the broken capture
                    hipGraph_t bad = nullptr;
                    (void) hipStreamEndCapture(st, &bad);
                    if(bad) (void) hipGraphDestroy(bad);
                    throw;
                }
                hipGraph_t g = nullptr;
                HIPCHK(hipStreamEndCapture(st, &g));
                const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void) hipGraphDestroy(g);
                HIPCHK(e);
            }
            HIPCHK(hipGraphLaunch(ex, st));
            cur = cur0 ^ (period & 1);
            it += period - 1;
        } else {
            launch_iteration(it);
        }

        if(period > 0) {
            // true-residual gate: the recurrence runs on past the reference's
            // quasi-residual test; it stops when the FP64 true residual meets
            // the tolerance, and hands over to a refinement restart on a
            // Lanczos breakdown, when the best true residual has not halved
            // for STALL_PERIODS checks, or when the quasi-residual has fallen
            // a thousandfold since the true residual last halved (in both
            // cases the recurrence has drifted from the true residual). The
            // quasi-residual test is not used here: with a large ||A||
            // estimate it fires long before the true residual is small.
            if((it + 1) % period == 0) {
                k_dsolve<<<nb, 256, 0, st>>>(n, w.xt, xout, ops.tri, ops.is_piv);
                if(ops.right) ops.right(xout, st);
                double nr, nbn;
                residual_norms(ops, w, b, xout, st, nr, nbn);
                if(nr <= gate_abs) { gate_met = true; ++it; break; }
                // on the solve's own stream (copies on the null stream could
                // disturb graph captures of other threads)
                HIPCHK(hipMemcpyAsync(w.host_flags, w.flags, 2 * sizeof(int32_t),
                    hipMemcpyDeviceToHost, st));
                HIPCHK(hipMemcpyAsync(w.host_scal + S_GAMMA, w.scal + S_GAMMA, sizeof(double),
                    hipMemcpyDeviceToHost, st));
                HIPCHK(hipStreamSynchronize(st));
                double gam = w.host_scal[S_GAMMA];
                gam = fabs(gam);
                if(best_nr < 0 || nr < 0.5 * best_nr) {
                    best_nr = nr;
                    gam_at_best = gam;
                    stall = 0;
                } else {
                    ++stall;
                }
                // drift: the recurrence's own residual estimate has fallen a
                // thousandfold since the true residual last halved
                const bool drifted = stall > 0 && gam < 1e-3 * gam_at_best;
                if(w.host_flags[0] || stall >= STALL_PERIODS || drifted) {
                    ++it;
                    break;
                }
            }
            ++it;
            continue;
        }
        // throttled convergence check (:309-331)
        if(it > 0 && (it % 4) == 0) {
            HIPCHK(hipEventSynchronize(ev_copy));
            if(w.host_flags[0] || w.host_flags[1]) break;
        } else if(((it + 1) % 4) == 0) {
            HIPCHK(hipMemcpyAsync(w.host_flags, w.flags, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            HIPCHK(hipEventRecord(ev_copy, st));
        }
        ++it;
    }
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipEventDestroy(ev_copy));
    total_it += it;
    HIPCHK(hipMemcpyAsync(w.host_scal, w.scal, NSCAL * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(w.host_flags, w.flags, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    res.residual = w.host_scal[S_GAMMA] / w.host_scal[S_NORMPB];
    res.converged = gate_met || w.host_flags[0] || w.host_flags[1];
    if(!gate_met) {
        // x <- Pr \ (D \ x) (:346-351)
        if(ops.tri) k_dsolve<<<nb, 256, 0, st>>>(n, w.xt, xout, ops.tri, ops.is_piv);
        else k_copy<<<nb, 256, 0, st>>>(n, w.xt, xout);
        if(ops.right) ops.right(xout, st);
    }
}

void SqmrGraphs::release()
{
    for(hipGraphExec_t& e : seg) {
        if(e) (void) hipGraphExecDestroy(e);
        e = nullptr;
    }
}

void sqmr_solve(const SqmrOps& ops, SqmrWork& w, const double* b, double* x,
    const culip_sqmr_opts& opt, culip_sqmr_result& res, void* stream, SqmrGraphs* graphs)
{
    hipStream_t st = (hipStream_t) stream;
    const int32_t n = ops.n;
    const unsigned nb = blocks_for(n);
    std::memset(&res, 0, sizeof(res));
    int32_t total_it = 0;
    bool gate = false;
    double gate_abs = -1.0;
    if(opt.true_residual_period > 0) {
        double nbn;
        k_norm2_pair<<<w.grid, TPB, 0, st>>>(n, b, b, ops.res_scale, w.part);
        k_norm_final<<<1, TPB, 0, st>>>(w.grid, w.part, w.scal + 40);
        HIPCHK(hipMemcpyAsync(w.host_scal + 40, w.scal + 40, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        nbn = w.host_scal[40];
        gate_abs = opt.tol * nbn;
    }
    sqmr_run(ops, w, b, x, opt, res, st, total_it, gate, gate_abs, graphs);
    double nr, nbn;
    residual_norms(ops, w, b, x, st, nr, nbn);
    double rr = nbn > 0 ? nr / nbn : nr;
    if((opt.true_residual_period > 0 || opt.max_restarts > 0) && nbn > 0 && !(rr <= 1.0)) {
        // worse than the zero start: keep x = 0 (residual b)
        k_fill<<<nb, 256, 0, st>>>(n, x, 0.0);
        k_copy<<<nb, 256, 0, st>>>(n, b, w.rt);
        rr = 1.0;
    }
    // iterative refinement: solve A d = r with a fresh recurrence, x += d,
    // while the FP64 true residual misses the tolerance. A correction that
    // does not lower the true residual is undone and ends the refinement,
    // so x is never worse than the best iterate seen.
    int restarts = 0;
    while(rr > opt.tol && restarts < opt.max_restarts && total_it < opt.max_it) {
        k_copy<<<nb, 256, 0, st>>>(n, x, w.xacc);
        k_copy<<<nb, 256, 0, st>>>(n, w.rt, w.rhs);
        culip_sqmr_opts o2 = opt;
        o2.max_it = opt.max_it - total_it;
        culip_sqmr_result r2{};
        bool g2 = false;
        sqmr_run(ops, w, w.rhs, x, o2, r2, st, total_it, g2, gate_abs, graphs);
        k_axpy1<<<nb, 256, 0, st>>>(n, w.xacc, x);
        residual_norms(ops, w, b, x, st, nr, nbn);
        const double rr_new = nbn > 0 ? nr / nbn : nr;
        ++restarts;
        if(!(rr_new < rr)) {
            k_copy<<<nb, 256, 0, st>>>(n, w.xacc, x);
            HIPCHK(hipStreamSynchronize(st));
            break;
        }
        rr = rr_new;
        res.residual = r2.residual;
    }
    res.iterations = total_it;
    res.restarts = restarts;
    res.true_rel_residual = rr;
    if(opt.true_residual_period > 0 || opt.max_restarts > 0) res.converged = rr <= opt.tol;
}


void batch_expand(const int4* ent, int32_t count, int2* items, int32_t interleaved_total,
    hipStream_t st)
{
    if(interleaved_total > 0)
        k_batch_holes<<<(unsigned) ((interleaved_total + 255) / 256), 256, 0, st>>>(items,
            interleaved_total);
    if(count > 0) k_batch_expand<<<count, 256, 0, st>>>(ent, items, interleaved_total > 0);
}

void batch_step(BatchStep s, const BatchLists& L, const BatchSys* sys, int32_t a, int32_t b,
    double tol, double brel, hipStream_t st)
{
    auto grid = [](const BatchList& l) { return (unsigned) l.count; };
    switch(s) {
    case BS_ALPHA:
        if(L.vec.count) {
            if(L.one_round) bk_alpha<1><<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
            else bk_alpha<VR><<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
        }
        break;
    case BS_VNEXT:
        if(L.vec.count) {
            if(L.one_round) bk_vnext<1><<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
            else bk_vnext<VR><<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
        }
        break;
    case BS_VSCALE:
        if(L.vec.count) {
            if(L.one_round) bk_vscale<1><<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
            else bk_vscale<VR><<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
        }
        break;
    case BS_BETA:
        if(L.vec.count) {
            if(L.one_round) bk_beta<1><<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
            else bk_beta<VR><<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
        }
        break;
    case BS_PUPDATE:
        if(L.vec.count) {
            if(L.one_round) bk_pupdate<1><<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
            else bk_pupdate<VR><<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
        }
        break;
    case BS_SCALARS:
        if(L.sys.count) bk_scalars<<<grid(L.sys), TPB, 0, st>>>(L.sys.items, sys, tol, brel);
        break;
    case BS_INIT_FILL:
        if(L.row.count) bk_init_fill<<<grid(L.row), 256, 0, st>>>(L.row.items, sys);
        break;
    case BS_SUMSQ:
        if(L.vec.count) bk_sumsq<<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys, a);
        break;
    case BS_INIT_V:
        if(L.vec.count) bk_init_v<<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
        break;
    case BS_INIT_W:
        if(L.vec.count) bk_init_w<<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys);
        break;
    case BS_NORM2_PAIR:
        if(L.vec.count) bk_norm2_pair<<<grid(L.vec), TPB, 0, st>>>(L.vec.items, sys, a, b);
        break;
    case BS_NORM_FINAL:
        if(L.sys.count) bk_norm_final<<<grid(L.sys), TPB, 0, st>>>(L.sys.items, sys);
        break;
    case BS_COPY:
    case BS_DSOLVE:
    case BS_RESIDUAL:
    case BS_AXPY1:
    case BS_ZERO: {
        const int op = s == BS_COPY ? 0 : s == BS_DSOLVE ? 1 : s == BS_RESIDUAL ? 2 : s == BS_AXPY1 ? 3 : 4;
        if(L.row.count) bk_elementwise<<<grid(L.row), 256, 0, st>>>(L.row.items, sys, op, a, b);
        break;
    }
    }
    HIPCHK(hipGetLastError());
}

}  // namespace culip
