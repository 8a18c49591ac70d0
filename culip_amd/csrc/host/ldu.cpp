// ImportLDUPreconditioner (libs/la/preconditioner/import_ldu_preconditioner.*):
// host analysis (triangle selection, transposes, level schedules, diagonal
// positions) and the solve sequences of solve_left / solve_middle /
// solve_right (import_ldu_preconditioner.impl.cuh:197-330) over the device
// substitution kernels.
#include <algorithm>
#include <cstring>

#include "objects.hpp"
#include "../device/ldu.hpp"

namespace culip {

namespace {

// the lower (or upper) triangle of a, diagonal included, columns sorted
HostCsr triangle(const HostCsr& a, bool lower)
{
    HostCsr t;
    t.m = a.m;
    t.n = a.n;
    t.row_ptr.assign(a.m + 1, 0);
    std::vector<std::pair<int32_t, double>> row;
    for(int32_t i = 0; i < a.m; ++i) {
        row.clear();
        for(int32_t p = a.row_ptr[i]; p < a.row_ptr[i + 1]; ++p) {
            const int32_t j = a.col[p];
            if(lower ? j <= i : j >= i) row.emplace_back(j, a.val[p]);
        }
        std::stable_sort(row.begin(), row.end(),
            [](const std::pair<int32_t, double>& x, const std::pair<int32_t, double>& y) {
                return x.first < y.first;
            });
        for(const auto& e : row) {
            t.col.push_back(e.first);
            t.val.push_back(e.second);
        }
        t.row_ptr[i + 1] = (int32_t) t.col.size();
    }
    return t;
}

HostCsr transpose(const HostCsr& a)
{
    HostCsr t;
    t.m = a.n;
    t.n = a.m;
    t.row_ptr.assign(a.n + 1, 0);
    for(int32_t j : a.col) ++t.row_ptr[j + 1];
    for(int32_t i = 0; i < a.n; ++i) t.row_ptr[i + 1] += t.row_ptr[i];
    t.col.resize(a.col.size());
    t.val.resize(a.val.size());
    std::vector<int32_t> next(t.row_ptr.begin(), t.row_ptr.end() - 1);
    for(int32_t i = 0; i < a.m; ++i)
        for(int32_t p = a.row_ptr[i]; p < a.row_ptr[i + 1]; ++p) {
            const int32_t q = next[a.col[p]]++;
            t.col[q] = i;
            t.val[q] = a.val[p];
        }
    return t;
}

// 1 / diagonal when a row's first or last entry is its diagonal, else 0
// (k_extract_diag, sparse_la.impl.cuh:31-54), rounded through the factor's
// precision
std::vector<double> inverse_diagonal(const HostCsr& a, bool f32)
{
    std::vector<double> inv(a.m, 0.0);
    for(int32_t i = 0; i < a.m; ++i) {
        const int32_t s = a.row_ptr[i], e = a.row_ptr[i + 1];
        if(s == e) continue;
        auto one = [&](double v) {
            return f32 ? (double) (float) (1.0 / (double) (float) v) : 1.0 / v;
        };
        if(a.col[s] == i) inv[i] = one(a.val[s]);
        if(a.col[e - 1] == i) inv[i] = one(a.val[e - 1]);
    }
    return inv;
}

void build_operand(Ldu::Operand& o, const HostCsr& t, bool lower, bool f32)
{
    const int32_t n = t.m;
    o.lower = lower;
    std::vector<int32_t> lev(n, 0), dpos(n, -1);
    int32_t levels = n ? 1 : 0;
    for(int32_t k = 0; k < n; ++k) {
        const int32_t i = lower ? k : n - 1 - k;
        int32_t l = 0;
        for(int32_t p = t.row_ptr[i]; p < t.row_ptr[i + 1]; ++p) {
            const int32_t j = t.col[p];
            if(j == i) dpos[i] = p;
            else l = std::max(l, lev[j] + 1);
        }
        lev[i] = l;
        levels = std::max(levels, l + 1);
    }
    o.levels = levels;
    o.lvl_off.assign(levels + 1, 0);
    for(int32_t i = 0; i < n; ++i) ++o.lvl_off[lev[i] + 1];
    for(int32_t l = 0; l < levels; ++l) o.lvl_off[l + 1] += o.lvl_off[l];
    std::vector<int32_t> order(n), fill(o.lvl_off.begin(), o.lvl_off.end() - 1);
    for(int32_t i = 0; i < n; ++i) order[fill[lev[i]]++] = i;
    upload(o.row_ptr, t.row_ptr);
    upload(o.col, t.col);
    upload(o.diag_pos, dpos);
    upload(o.order, order);
    if(f32) upload(o.val, std::vector<float>(t.val.begin(), t.val.end()));
    else upload(o.val, t.val);
}

template<typename VT>
TriOperand<VT> view(const Ldu::Operand& o, int32_t n, const DevBuf& inv)
{
    TriOperand<VT> t;
    t.n = n;
    t.lower = o.lower;
    t.row_ptr = o.row_ptr.as<int32_t>();
    t.col = o.col.as<int32_t>();
    t.val = o.val.as<VT>();
    t.diag_pos = o.diag_pos.as<int32_t>();
    t.order = o.order.as<int32_t>();
    t.inv_diag = inv.as<double>();
    return t;
}

}  // namespace

void Ldu::create(const HostCsr& l, const HostCsr* u, const double* tri_in,
    const int32_t* is_piv_in, bool exact_solve, culip_prec p)
{
    if(l.m != l.n) fail(CULIP_ERR_INVALID, "ldu: L must be square");
    if(u && (u->m != l.m || u->n != l.n)) fail(CULIP_ERR_INVALID, "ldu: U must match L");
    if((tri_in == nullptr) != (is_piv_in == nullptr))
        fail(CULIP_ERR_INVALID, "ldu: D needs both the tridiagonal and the pivot flags");
    n = l.m;
    prec = p;
    symmetric = (u == nullptr);
    has_d = (tri_in != nullptr);
    exact = exact_solve;
    const bool f32 = (p == CULIP_PREC_F32);
    const HostCsr lt = triangle(l, true);
    build_operand(L, lt, true, f32);
    build_operand(Lt, transpose(lt), false, f32);
    upload(d_inv_l, inverse_diagonal(l, f32));
    if(u) {
        const HostCsr ut = triangle(*u, false);
        build_operand(U, ut, false, f32);
        build_operand(Ut, transpose(ut), true, f32);
        upload(d_inv_u, inverse_diagonal(*u, f32));
    }
    if(has_d) {
        upload(d_tri, std::vector<double>(tri_in, tri_in + 3 * (size_t) n));
        upload(d_is_piv, std::vector<int32_t>(is_piv_in, is_piv_in + n));
    }
    d_t1.ensure(sizeof(double) * (size_t) std::max(n, 1));
    d_t2.ensure(sizeof(double) * (size_t) std::max(n, 1));
}

void Ldu::set_permutation(const int32_t* pr, const int32_t* pc)
{
    if((pr == nullptr) != (pc == nullptr))
        fail(CULIP_ERR_INVALID, "ldu: give both permutations or neither");
    has_p = (pr != nullptr);
    if(!has_p) return;
    for(const int32_t* q : {pr, pc}) {
        std::vector<char> seen(n, 0);
        for(int32_t i = 0; i < n; ++i) {
            if(q[i] < 0 || q[i] >= n || seen[q[i]]) fail(CULIP_ERR_INVALID, "ldu: not a permutation");
            seen[q[i]] = 1;
        }
    }
    upload(d_pr, std::vector<int32_t>(pr, pr + n));
    upload(d_pc, std::vector<int32_t>(pc, pc + n));
}

void Ldu::set_scaling(const double* sr, const double* sc)
{
    if((sr == nullptr) != (sc == nullptr))
        fail(CULIP_ERR_INVALID, "ldu: give both scalings or neither");
    has_s = (sr != nullptr);
    if(!has_s) return;
    upload(d_sr, std::vector<double>(sr, sr + n));
    upload(d_sc, std::vector<double>(sc, sc + n));
}

// out = op \ in (in may equal out; out must not alias the work vector t2)
void Ldu::tri_apply(const Operand& o, const DevBuf& inv, const double* in, double* out,
    hipStream_t st) const
{
    const bool f32 = (prec == CULIP_PREC_F32);
    if(exact) {
        if(in != out)
            HIP_CHECK(hipMemcpyAsync(out, in, sizeof(double) * (size_t) n, hipMemcpyDeviceToDevice,
                st));
        if(f32) tri_solve_levels(view<float>(o, n, inv), o.lvl_off.data(), o.levels, out, st);
        else tri_solve_levels(view<double>(o, n, inv), o.lvl_off.data(), o.levels, out, st);
        return;
    }
    const double* b = in;
    if(in == out) {
        HIP_CHECK(hipMemcpyAsync(d_t1.as<double>(), in, sizeof(double) * (size_t) n,
            hipMemcpyDeviceToDevice, st));
        b = d_t1.as<double>();
    }
    if(f32) tri_solve_jacobi(view<float>(o, n, inv), b, out, d_t2.as<double>(), APPROX_SWEEPS, st);
    else tri_solve_jacobi(view<double>(o, n, inv), b, out, d_t2.as<double>(), APPROX_SWEEPS, st);
}

// solve_PS (import_ldu_preconditioner.impl.cuh:414-484): pull x = P' (S b),
// push x = S (P b); row data (P_r, S_r) for the left side
void Ldu::ps(bool left_side, bool push, const double* b, double* x, hipStream_t st) const
{
    const int32_t* p = has_p ? (left_side ? d_pr : d_pc).as<int32_t>() : nullptr;
    const double* s = has_s ? (left_side ? d_sr : d_sc).as<double>() : nullptr;
    if(!p && !s) {
        if(b != x)
            HIP_CHECK(hipMemcpyAsync(x, b, sizeof(double) * (size_t) n, hipMemcpyDeviceToDevice, st));
        return;
    }
    permute_scale(n, p, s, push, b, x, st);
}

void Ldu::solve(culip_side side, bool transpose, const double* b, double* x, hipStream_t st) const
{
    double* t1 = d_t1.as<double>();
    if(side == CULIP_SOLVE_MIDDLE) {
        // solve_middle (:260-279): copy, then the 1x1 / 2x2 D solve
        if(has_d) solve_D(n, d_tri.as<double>(), d_is_piv.as<int32_t>(), b, x, st);
        else if(b != x)
            HIP_CHECK(hipMemcpyAsync(x, b, sizeof(double) * (size_t) n, hipMemcpyDeviceToDevice, st));
        return;
    }
    // the permutation step writes t1 and never in place; an approximate
    // solve reading t1 gets its own copy in tri_apply only when in == out
    if(side == CULIP_SOLVE_LEFT) {
        if(!transpose) {
            // inv(P_l) = inv(L) P_r' S_r
            ps(true, false, b, t1, st);
            if(exact) tri_apply(L, d_inv_l, t1, x, st);
            else {
                HIP_CHECK(hipMemcpyAsync(x, t1, sizeof(double) * (size_t) n,
                    hipMemcpyDeviceToDevice, st));
                tri_apply(L, d_inv_l, x, x, st);
            }
        } else {
            // inv(P_l') = S_r P_r inv(L')
            tri_apply(Lt, d_inv_l, b, x, st);
            ps(true, true, x, t1, st);
            HIP_CHECK(hipMemcpyAsync(x, t1, sizeof(double) * (size_t) n, hipMemcpyDeviceToDevice,
                st));
        }
        return;
    }
    const Operand& r = symmetric ? Lt : U;     // R
    const Operand& rt = symmetric ? L : Ut;    // R'
    const DevBuf& inv = symmetric ? d_inv_l : d_inv_u;
    if(!transpose) {
        // inv(P_r) = S_c P_c inv(R)
        tri_apply(r, inv, b, x, st);
        ps(false, true, x, t1, st);
        HIP_CHECK(hipMemcpyAsync(x, t1, sizeof(double) * (size_t) n, hipMemcpyDeviceToDevice, st));
    } else {
        // inv(P_r') = inv(R') P_c' S_c
        ps(false, false, b, t1, st);
        HIP_CHECK(hipMemcpyAsync(x, t1, sizeof(double) * (size_t) n, hipMemcpyDeviceToDevice, st));
        tri_apply(rt, inv, x, x, st);
    }
}

int64_t Ldu::device_bytes() const
{
    int64_t b = 0;
    for(const Operand* o : {&L, &Lt, &U, &Ut})
        b += (int64_t) (o->row_ptr.bytes + o->col.bytes + o->val.bytes + o->diag_pos.bytes +
            o->order.bytes);
    for(const DevBuf* d : {&d_inv_l, &d_inv_u, &d_tri, &d_is_piv, &d_pr, &d_pc, &d_sr, &d_sc,
            &d_t1, &d_t2})
        b += (int64_t) d->bytes;
    return b;
}

}  // namespace culip
