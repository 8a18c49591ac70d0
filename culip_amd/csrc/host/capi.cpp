// C-ABI of culip_amd (include/culip_amd.h). Every entry point converts
// library errors into status codes; the message is kept per thread.
#include "objects.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "kkt.hpp"
#include "matching.hpp"
#include "serve.hpp"
#include "batch_solver.hpp"

struct culip_server { std::unique_ptr<culip::Server> s; };
struct culip_batch { std::unique_ptr<culip::BatchSolver> b; };
struct culip_kkt { culip::Kkt* k; };
struct culip_prep { culip::Prep* p; };
struct culip_host_matrix { culip::HostCsr a; };
struct culip_bildlt { culip::Bildlt b; };
struct culip_spmv { culip::Spmv s; };
struct culip_sqmr { culip::Sqmr q; };
struct culip_ldu { culip::Ldu l; };

using namespace culip;

static thread_local std::string g_last_error;

template<typename F>
static culip_status guarded(F&& f)
{
    try {
        f();
        g_last_error.clear();
        return CULIP_OK;
    } catch(const Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch(const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return CULIP_ERR_NOMEM;
    } catch(const std::exception& e) {
        g_last_error = e.what();
        return CULIP_ERR_INVALID;
    }
}

static HostCsr to_host(const culip_csr* c)
{
    CULIP_REQUIRE(c && c->row_ptr && c->m >= 0 && c->n >= 0, "csr: null or malformed matrix");
    HostCsr h;
    h.m = c->m;
    h.n = c->n;
    h.row_ptr.assign(c->row_ptr, c->row_ptr + c->m + 1);
    CULIP_REQUIRE(h.row_ptr[0] == 0 && (int64_t) h.row_ptr[c->m] == c->nnz, "csr: row_ptr inconsistent with nnz");
    for(int32_t i = 0; i < c->m; ++i) CULIP_REQUIRE(h.row_ptr[i] <= h.row_ptr[i + 1], "csr: row_ptr not monotone");
    if(c->nnz) {
        CULIP_REQUIRE(c->col && c->val, "csr: null col/val");
        h.col.assign(c->col, c->col + c->nnz);
        h.val.assign(c->val, c->val + c->nnz);
    }
    for(int32_t q : h.col) CULIP_REQUIRE(q >= 0 && q < h.n, "csr: column index out of range");
    return h;
}

static void view(const HostCsr& h, culip_csr* out)
{
    out->m = h.m;
    out->n = h.n;
    out->nnz = h.nnz();
    out->row_ptr = h.row_ptr.data();
    out->col = h.col.data();
    out->val = h.val.data();
}

static void device_present()
{
    int n = 0;
    if(hipGetDeviceCount(&n) != hipSuccess || n == 0)
        fail(CULIP_ERR_NO_DEVICE, "no HIP device available");
}

extern "C" {

const char* culip_last_error(void) { return g_last_error.c_str(); }
const char* culip_version(void) { return "culip_amd 0.1 (gfx950)"; }

// ---- KKT generator / prep -------------------------------------------------

culip_status culip_kkt_generate(const culip_kkt_params* p, culip_kkt** out)
{
    return guarded([&] {
        CULIP_REQUIRE(p && out, "kkt: null argument");
        auto* h = new culip_kkt;
        h->k = kkt_generate(*p);
        *out = h;
    });
}

culip_status culip_kkt_view(const culip_kkt* k, culip_csr* K, const double** rhs)
{
    return guarded([&] {
        CULIP_REQUIRE(k && K, "kkt: null argument");
        view(k->k->K, K);
        if(rhs) *rhs = k->k->rhs.data();
    });
}

culip_status culip_kkt_barrier_values(const culip_kkt* k, double w_lo, double w_hi,
    uint64_t seed, double* vals_out)
{
    return guarded([&] {
        CULIP_REQUIRE(k && vals_out, "kkt: null argument");
        kkt_barrier_values(*k->k, w_lo, w_hi, seed, vals_out);
    });
}

void culip_kkt_destroy(culip_kkt* k)
{
    if(k) { delete k->k; delete k; }
}

culip_status culip_prep_create(const culip_csr* K, int32_t n_primal, int32_t max_block,
    culip_ordering ordering, culip_prep** out)
{
    return guarded([&] {
        CULIP_REQUIRE(out, "prep: null argument");
        HostCsr h = to_host(K);
        auto* p = new culip_prep;
        p->p = prep_create(h, n_primal, max_block, ordering);
        *out = p;
    });
}

culip_status culip_prep_view(const culip_prep* p, culip_csr* K_perm, const int32_t** perm,
    int32_t* num_blocks, const int32_t** block_starts, int32_t* num_pivs,
    const int32_t** piv_starts)
{
    return guarded([&] {
        CULIP_REQUIRE(p, "prep: null argument");
        if(K_perm) view(p->p->K, K_perm);
        if(perm) *perm = p->p->perm.data();
        if(num_blocks) *num_blocks = (int32_t) p->p->block_starts.size();
        if(block_starts) *block_starts = p->p->block_starts.data();
        if(num_pivs) *num_pivs = (int32_t) p->p->piv_starts.size();
        if(piv_starts) *piv_starts = p->p->piv_starts.data();
    });
}

culip_status culip_prep_permute_values(const culip_prep* p, const double* vals_K,
    double* vals_perm)
{
    return guarded([&] {
        CULIP_REQUIRE(p && vals_K && vals_perm, "prep: null argument");
        const auto& src = p->p->entry_src;
        for(size_t q = 0; q < src.size(); ++q) vals_perm[q] = vals_K[src[q]];
    });
}

void culip_prep_destroy(culip_prep* p)
{
    if(p) { delete p->p; delete p; }
}

// ---- MatrixMarket (libs/test/test.impl.h:25-300) ---------------------------

static std::vector<std::string> mm_lines(const char* path, std::string& banner)
{
    std::ifstream f(path);
    if(!f.good()) fail(CULIP_ERR_INVALID, std::string("cannot open ") + path);
    std::vector<std::string> lines;
    std::string line;
    bool first = true;
    while(std::getline(f, line)) {
        if(first) { banner = line; first = false; continue; }
        if(!line.empty() && line[0] == '%') continue;
        if(line.find_first_not_of(" \t\r") == std::string::npos) continue;
        lines.push_back(line);
    }
    std::transform(banner.begin(), banner.end(), banner.begin(), ::tolower);
    return lines;
}

culip_status culip_mm_read_csr(const char* path, culip_host_matrix** out)
{
    return guarded([&] {
        CULIP_REQUIRE(path && out, "mm: null argument");
        std::string banner;
        auto lines = mm_lines(path, banner);
        CULIP_REQUIRE(banner.find("coordinate") != std::string::npos, "mm: only coordinate matrices");
        const bool sym = banner.find("symmetric") != std::string::npos;
        CULIP_REQUIRE(!lines.empty(), "mm: missing size line");
        int64_t m, n, nnz;
        { std::istringstream ss(lines[0]); ss >> m >> n >> nnz; }
        std::vector<int32_t> ii, jj;
        std::vector<double> cc;
        for(int64_t q = 0; q < nnz; ++q) {
            CULIP_REQUIRE((size_t) q + 1 < lines.size(), "mm: truncated file");
            std::istringstream ss(lines[q + 1]);
            int64_t i, j;
            double v = 1.0;
            ss >> i >> j >> v;
            ii.push_back((int32_t) (i - 1));
            jj.push_back((int32_t) (j - 1));
            cc.push_back(v);
        }
        // symmetric expansion: mirror off-diagonal entries (:76-91)
        if(sym) {
            const size_t base = ii.size();
            for(size_t q = 0; q < base; ++q)
                if(ii[q] != jj[q]) { ii.push_back(jj[q]); jj.push_back(ii[q]); cc.push_back(cc[q]); }
        }
        // COO -> CSR, per-row sort by column keeping duplicates (:111-172)
        auto* h = new culip_host_matrix;
        HostCsr& a = h->a;
        a.m = (int32_t) m;
        a.n = (int32_t) n;
        a.row_ptr.assign(m + 1, 0);
        for(int32_t r : ii) a.row_ptr[r + 1]++;
        for(int64_t r = 0; r < m; ++r) a.row_ptr[r + 1] += a.row_ptr[r];
        a.col.resize(ii.size());
        a.val.resize(ii.size());
        std::vector<int32_t> fillp(a.row_ptr.begin(), a.row_ptr.end() - 1);
        for(size_t q = 0; q < ii.size(); ++q) {
            const int32_t pos = fillp[ii[q]]++;
            a.col[pos] = jj[q];
            a.val[pos] = cc[q];
        }
        for(int64_t r = 0; r < m; ++r) {
            std::vector<std::pair<int32_t, double>> row;
            for(int32_t q = a.row_ptr[r]; q < a.row_ptr[r + 1]; ++q) row.emplace_back(a.col[q], a.val[q]);
            std::stable_sort(row.begin(), row.end(),
                [](const auto& x, const auto& y) { return x.first < y.first; });
            for(size_t q = 0; q < row.size(); ++q) {
                a.col[a.row_ptr[r] + q] = row[q].first;
                a.val[a.row_ptr[r] + q] = row[q].second;
            }
        }
        *out = h;
    });
}

culip_status culip_mm_read_vector(const char* path, int64_t* len, double** vals)
{
    return guarded([&] {
        CULIP_REQUIRE(path && len && vals, "mm: null argument");
        std::string banner;
        auto lines = mm_lines(path, banner);
        CULIP_REQUIRE(banner.find("array") != std::string::npos, "mm: only array vectors");
        int64_t m, n;
        { std::istringstream ss(lines.at(0)); ss >> m >> n; }
        const int64_t k = std::max(m, n);
        CULIP_REQUIRE(k >= 0 && (size_t) k + 1 <= lines.size(), "mm: truncated vector");
        double* v = (double*) std::malloc(sizeof(double) * (size_t) std::max<int64_t>(1, k));
        if(!v) throw std::bad_alloc();
        for(int64_t q = 0; q < k; ++q) v[q] = std::strtod(lines[q + 1].c_str(), nullptr);
        *len = k;
        *vals = v;
    });
}

culip_status culip_host_matrix_view(const culip_host_matrix* a, culip_csr* out)
{
    return guarded([&] {
        CULIP_REQUIRE(a && out, "mm: null argument");
        view(a->a, out);
    });
}

void culip_host_matrix_destroy(culip_host_matrix* a) { delete a; }

culip_status culip_csr_permute_scale(const culip_csr* A, const int32_t* row_perm,
    const int32_t* col_perm, int32_t perm_is_old_to_new, const double* row_scale,
    const double* col_scale, int32_t scale_before_match, culip_host_matrix** out)
{
    return guarded([&] {
        CULIP_REQUIRE(out && row_perm && col_perm, "permute: null argument");
        HostCsr a = to_host(A);
        CULIP_REQUIRE(a.m == a.n, "permute: matrix must be square");
        const int32_t n = a.m;
        // old -> new maps (k_inv_perm inverts a new -> old vector)
        auto to_new = [&](const int32_t* p) {
            std::vector<int32_t> q(n), seen(n, 0);
            for(int32_t i = 0; i < n; ++i) {
                CULIP_REQUIRE(p[i] >= 0 && p[i] < n && !seen[p[i]], "permute: not a permutation");
                seen[p[i]] = 1;
            }
            for(int32_t i = 0; i < n; ++i) {
                if(perm_is_old_to_new) q[i] = p[i];
                else q[p[i]] = i;
            }
            return q;
        };
        const std::vector<int32_t> pr = to_new(row_perm), pc = to_new(col_perm);
        struct E { int32_t r, c; double v; };
        std::vector<E> es;
        es.reserve((size_t) a.nnz());
        for(int32_t r = 0; r < n; ++r)
            for(int32_t q = a.row_ptr[r]; q < a.row_ptr[r + 1]; ++q) {
                const int32_t c = a.col[q], nr = pr[r], nc = pc[c];
                const double sr = row_scale ? row_scale[scale_before_match ? r : nr] : 1.0;
                const double sc = col_scale ? col_scale[scale_before_match ? c : nc] : 1.0;
                es.push_back({nr, nc, a.val[q] / (sr * sc)});
            }
        std::stable_sort(es.begin(), es.end(),
            [](const E& x, const E& y) { return x.r < y.r || (x.r == y.r && x.c < y.c); });
        auto* h = new culip_host_matrix;
        HostCsr& b = h->a;
        b.m = b.n = n;
        b.row_ptr.assign(n + 1, 0);
        b.col.resize(es.size());
        b.val.resize(es.size());
        for(size_t q = 0; q < es.size(); ++q) {
            ++b.row_ptr[es[q].r + 1];
            b.col[q] = es[q].c;
            b.val[q] = es[q].v;
        }
        for(int32_t i = 0; i < n; ++i) b.row_ptr[i + 1] += b.row_ptr[i];
        *out = h;
    });
}
void culip_host_free(void* p) { std::free(p); }

culip_status culip_weighted_matching(const culip_csr* A, int32_t* row_match, double* u, double* v,
    double* row_scale, double* col_scale, double* sym_scale, double* objective, int32_t* feasible)
{
    return guarded([&] {
        const HostCsr a = to_host(A);
        const MatchingResult r = weighted_matching(a);
        const size_t n = (size_t) a.m;
        if(row_match) std::copy(r.row_match.begin(), r.row_match.end(), row_match);
        if(u) std::copy(r.u.begin(), r.u.end(), u);
        if(v) std::copy(r.v.begin(), r.v.end(), v);
        if(row_scale) std::copy(r.row_scale.begin(), r.row_scale.end(), row_scale);
        if(col_scale) std::copy(r.col_scale.begin(), r.col_scale.end(), col_scale);
        if(sym_scale) {
            const std::vector<double> d = symmetric_matching_scaling(r);
            std::copy(d.begin(), d.begin() + n, sym_scale);
        }
        if(objective) *objective = r.objective;
        if(feasible) *feasible = r.feasible ? 1 : 0;
    });
}

culip_status culip_structural_rank(const culip_csr* A, int32_t* rank, int32_t* row_match)
{
    return guarded([&] {
        CULIP_REQUIRE(A && A->row_ptr && rank, "structural_rank: null argument");
        CULIP_REQUIRE(A->nnz == 0 || A->col, "structural_rank: null col");
        std::vector<int32_t> rm;
        *rank = structural_rank(A->m, A->n, A->row_ptr, A->col, row_match ? &rm : nullptr);
        if(row_match) std::copy(rm.begin(), rm.end(), row_match);
    });
}

culip_status culip_btf_rank_bound(const culip_csr* A, int32_t* bound, int32_t* row_order,
    int32_t* col_order)
{
    return guarded([&] {
        CULIP_REQUIRE(A && A->row_ptr && bound, "btf_rank_bound: null argument");
        CULIP_REQUIRE(A->nnz == 0 || A->col, "btf_rank_bound: null col");
        const int32_t m = A->m, n = A->n;
        // the pattern by columns (A' as CSR)
        std::vector<int32_t> cp(n + 1, 0), cr((size_t) A->nnz);
        for(int64_t q = 0; q < A->nnz; ++q) {
            CULIP_REQUIRE(A->col[q] >= 0 && A->col[q] < n, "btf_rank_bound: column out of range");
            ++cp[A->col[q] + 1];
        }
        for(int32_t j = 0; j < n; ++j) cp[j + 1] += cp[j];
        std::vector<int32_t> fill(cp.begin(), cp.end() - 1);
        for(int32_t i = 0; i < m; ++i)
            for(int32_t q = A->row_ptr[i]; q < A->row_ptr[i + 1]; ++q) cr[fill[A->col[q]]++] = i;
        std::vector<int32_t> ro(m), co(n);
        *bound = btf_rank_bound(m, n, cp.data(), cr.data(), A->row_ptr, A->col, ro.data(), co.data());
        if(row_order) std::copy(ro.begin(), ro.end(), row_order);
        if(col_order) std::copy(co.begin(), co.end(), col_order);
    });
}

// ---- BlockiLDLt -------------------------------------------------------------

culip_status culip_bildlt_create(const culip_csr* hA, int32_t num_pivs, const int32_t* piv_starts,
    int32_t num_blocks, const int32_t* block_starts, culip_prec prec, culip_bildlt** out)
{
    return guarded([&] {
        CULIP_REQUIRE(out, "bildlt: null argument");
        CULIP_REQUIRE(prec == CULIP_PREC_F64 || prec == CULIP_PREC_F32, "bildlt: unknown precision");
        CULIP_REQUIRE(num_pivs > 0 && piv_starts, "bildlt: pivot starts required");
        CULIP_REQUIRE(num_blocks > 0 && block_starts, "bildlt: block starts required");
        auto* h = new culip_bildlt;
        try {
            h->b.A = to_host(hA);
            CULIP_REQUIRE(h->b.A.m == h->b.A.n, "bildlt: matrix must be square");
            h->b.prec = prec;
            h->b.pat->prec = prec;
            h->b.piv_starts.assign(piv_starts, piv_starts + num_pivs);
            h->b.block_starts.assign(block_starts, block_starts + num_blocks);
        } catch(...) { delete h; throw; }
        *out = h;
    });
}

culip_status culip_bildlt_symbolic(culip_bildlt* h, int32_t fill_level, double fill_factor)
{
    return guarded([&] {
        CULIP_REQUIRE(h, "bildlt: null handle");
        h->b.symbolic(fill_level, fill_factor);
    });
}

culip_status culip_bildlt_compute(culip_bildlt* h, int32_t fill_level, double fill_factor,
    double threshold, void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(h, "bildlt: null handle");
        device_present();
        if(!h->b.have_sym || h->b.sym.fill_level != fill_level || h->b.sym.fill_factor != fill_factor)
            h->b.symbolic(fill_level, fill_factor);
        h->b.factor(nullptr, false, threshold, (hipStream_t) stream);
    });
}

culip_status culip_bildlt_refactor(culip_bildlt* h, const double* vals, int32_t vals_on_device,
    double threshold, void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(h && vals, "bildlt: null argument");
        device_present();
        h->b.factor(vals, vals_on_device != 0, threshold, (hipStream_t) stream);
    });
}

culip_status culip_bildlt_solve(const culip_bildlt* h, culip_side side, int32_t transpose,
    const double* b, double* x, void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(h && b && x, "bildlt: null argument");
        CULIP_REQUIRE(side == CULIP_SOLVE_LEFT || side == CULIP_SOLVE_MIDDLE || side == CULIP_SOLVE_RIGHT,
            "bildlt: unknown side");
        h->b.solve(side, transpose != 0, b, x, (hipStream_t) stream);
        HIP_CHECK(hipGetLastError());
    });
}

culip_status culip_bildlt_solve_left_product(const culip_bildlt* h, const culip_spmv* K,
    const double* src, double* x, int32_t* fused, void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(h && K && src && x, "bildlt: null argument");
        CULIP_REQUIRE(src != x, "bildlt: the product's source and result must differ");
        const Spmv& op = K->s;
        CULIP_REQUIRE(!op.kkt && !op.normal && op.m == h->b.n() && op.n == h->b.n(),
            "bildlt: the product needs a CSR operator of the factor's size");
        if(!h->b.factored) fail(CULIP_ERR_STATE, "solve: preconditioner has not been computed");
        h->b.take_async_error();
        hipStream_t st = (hipStream_t) stream;
        const bool f = h->b.left_product(op.A, src, x, st);
        if(!f) {
            op.multiply(src, x, 1.0, 0.0, false, st);
            h->b.left(x, st);
        }
        HIP_CHECK(hipGetLastError());
        if(fused) *fused = f ? 1 : 0;
    });
}

culip_status culip_bildlt_solve_noalloc(const culip_bildlt* h, culip_side side,
    int32_t transpose, const double* b, double* x, void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(h && b && x, "bildlt: null argument");
        CULIP_REQUIRE(side == CULIP_SOLVE_LEFT || side == CULIP_SOLVE_MIDDLE || side == CULIP_SOLVE_RIGHT,
            "bildlt: unknown side");
        h->b.solve_noalloc(side, transpose != 0, b, x, (hipStream_t) stream);
        HIP_CHECK(hipGetLastError());
    });
}

int32_t culip_bildlt_n(const culip_bildlt* h) { return h ? h->b.n() : -1; }

culip_status culip_bildlt_export_factorized(culip_bildlt* h, const char* folder,
    double* fill_factor)
{
    return guarded([&] {
        CULIP_REQUIRE(h && folder, "bildlt: null argument");
        const double f = h->b.export_factorized(folder);
        if(fill_factor) *fill_factor = f;
    });
}

culip_status culip_bildlt_set_solve_algorithm(culip_bildlt* h, culip_solve_algorithm algorithm,
    int32_t num_jacobi_sweeps)
{
    return guarded([&] {
        CULIP_REQUIRE(h, "bildlt: null handle");
        CULIP_REQUIRE(algorithm == CULIP_TRSV_LEVELS || algorithm == CULIP_TRSV_SINGLE ||
                algorithm == CULIP_TRSV_JACOBI,
            "bildlt: unknown solve algorithm");
        CULIP_REQUIRE(algorithm != CULIP_TRSV_JACOBI || num_jacobi_sweeps >= 1,
            "bildlt: block-Jacobi solves need at least one sweep");
        if(algorithm == CULIP_TRSV_JACOBI) device_present();
        h->b.set_solve(algorithm, num_jacobi_sweeps);
    });
}

int64_t culip_bildlt_device_bytes(const culip_bildlt* h) { return h ? h->b.device_bytes() : 0; }

culip_status culip_bildlt_get_int(culip_bildlt* h, const char* name, const int32_t** ptr,
    int64_t* len)
{
    return guarded([&] {
        CULIP_REQUIRE(h && name && ptr && len, "bildlt: null argument");
        Bildlt& b = h->b;
        const Symbolic& s = b.sym;
        const std::string nm(name);
        auto put = [&](const std::vector<int32_t>& v) { *ptr = v.data(); *len = (int64_t) v.size(); };
        if(nm == "plan_sizes") {
            CULIP_REQUIRE(b.factored, "bildlt: factor has not been computed");
            const int32_t* ps = b.plan_sizes();
            b.host_int[nm] = std::vector<int32_t>(ps, ps + 4);
            put(b.host_int[nm]);
            return;
        }
        if(nm == "fblocks" || nm == "fix_store" || nm == "carry" || nm == "perm" ||
            nm == "fis_piv") {
            b.download_factor();
            put(b.host_int.at(nm));
            return;
        }
        CULIP_REQUIRE(b.have_sym, "bildlt: symbolic phase has not run");
        auto& cache = b.host_int;
        if(nm == "block_starts") put(s.block_starts);
        else if(nm == "coarse_csr_off") put(s.csr_off);
        else if(nm == "coarse_csr_col") put(s.csr_col);
        else if(nm == "coarse_csc_off") put(s.csc_off);
        else if(nm == "coarse_csc_idx") put(s.csc_idx);
        else if(nm == "a_block_nnz") put(s.a_block_nnz);
        else if(nm == "lvl_off") put(s.lvl_off);
        else if(nm == "lvl_rows") put(s.lvl_rows);
        else if(nm == "row_level") put(s.row_level);
        else if(nm == "lvlb_off") put(s.lvlb_off);
        else if(nm == "lvlb_idx") put(s.lvlb_idx);
        else if(nm == "block_in_lvl") put(s.block_in_lvl);
        else if(nm == "row_first_block_offset") put(s.row_first_block_offset);
        else if(nm == "solve_off") put(s.solve_off);
        else if(nm == "solve_idx") put(s.solve_idx);
        else if(nm == "is_piv") put(s.is_piv);
        else if(nm == "ix_store") put(s.ix_store);
        else if(nm == "blocks") {
            std::vector<int32_t> v((size_t) s.coarse_nnz * 9);
            std::memcpy(v.data(), s.blocks.data(), sizeof(BlockMeta) * s.blocks.size());
            cache[nm] = std::move(v);
            put(cache[nm]);
        } else if(nm == "upd_off") {
            std::vector<int32_t> v(s.upd_off.begin(), s.upd_off.end());
            cache[nm] = std::move(v);
            put(cache[nm]);
        } else if(nm == "upd_ki") put(s.upd_ki);
        else if(nm == "upd_kj") put(s.upd_kj);
        else if(nm == "scalars") {
            cache[nm] = {s.n, s.nb, s.fine_u_nnz, s.init_coarse_nnz, s.fill_in_nnz, s.coarse_nnz,
                s.num_levels, s.max_level_size, s.max_level_blocks, s.blks_sparse, s.blks_dense,
                (int32_t) s.ix_space, (int32_t) s.val_space};
            put(cache[nm]);
        } else fail(CULIP_ERR_INVALID, "bildlt: unknown structure '" + nm + "'");
    });
}

culip_status culip_bildlt_get_real(culip_bildlt* h, const char* name, const double** ptr,
    int64_t* len)
{
    return guarded([&] {
        CULIP_REQUIRE(h && name && ptr && len, "bildlt: null argument");
        Bildlt& b = h->b;
        const std::string nm(name);
        if(nm == "val_store") {
            CULIP_REQUIRE(b.have_sym, "bildlt: symbolic phase has not run");
            auto& v = b.host_real[nm];
            v.assign(b.sym.val_space, 0.0);
            symbolic_gather_values(b.sym, b.A.val.data(), v.data());
            if(b.f32()) for(double& q : v) q = (double) (float) q;
            *ptr = v.data();
            *len = (int64_t) v.size();
            return;
        }
        if(nm == "fval_store" || nm == "tri" || nm == "row_norms") {
            b.download_factor();
            const auto& v = b.host_real.at(nm);
            *ptr = v.data();
            *len = (int64_t) v.size();
            return;
        }
        fail(CULIP_ERR_INVALID, "bildlt: unknown structure '" + nm + "'");
    });
}

culip_status culip_bildlt_set_pivoting(culip_bildlt* h, culip_pivoting mode)
{
    return guarded([&] {
        CULIP_REQUIRE(h, "bildlt: null argument");
        CULIP_REQUIRE(mode == CULIP_PIVOT_STATIC || mode == CULIP_PIVOT_BUNCH_KAUFMAN ||
                mode == CULIP_PIVOT_ROOK, "bildlt: unknown pivoting");
        if(h->b.pivoting != mode) h->b.release_graphs();
        h->b.pivoting = mode;
    });
}

culip_status culip_bildlt_set_pivot_precision(culip_bildlt* h, culip_pivot_precision p)
{
    return guarded([&] {
        CULIP_REQUIRE(h, "bildlt: null argument");
        CULIP_REQUIRE(p == CULIP_PIVOT2_F64 || p == CULIP_PIVOT2_FACTOR,
            "bildlt: unknown pivot precision");
        if(h->b.pivot_precision != p) h->b.release_graphs();
        h->b.pivot_precision = p;
    });
}

culip_status culip_bildlt_set_graphs(culip_bildlt* h, int32_t enable)
{
    return guarded([&] {
        CULIP_REQUIRE(h, "bildlt: null argument");
        h->b.use_graphs = enable != 0;
        if(!h->b.use_graphs) h->b.release_graphs();
    });
}

culip_status culip_bildlt_inject_solve_error(culip_bildlt* h)
{
    return guarded([&] {
        CULIP_REQUIRE(h, "bildlt: null argument");
        if(!h->b.uploaded || !h->b.h_trsv_err.p) fail(CULIP_ERR_STATE, "bildlt: not uploaded");
        __atomic_store_n(h->b.h_trsv_err.p, 1, __ATOMIC_RELEASE);
    });
}

void culip_bildlt_destroy(culip_bildlt* h) { delete h; }

// ---- SPMV -----------------------------------------------------------------

static HostCsr transpose_csr(const HostCsr& a)
{
    HostCsr t;
    t.m = a.n;
    t.n = a.m;
    t.row_ptr.assign(a.n + 1, 0);
    for(int32_t c : a.col) t.row_ptr[c + 1]++;
    for(int32_t i = 0; i < a.n; ++i) t.row_ptr[i + 1] += t.row_ptr[i];
    t.col.resize(a.col.size());
    t.val.resize(a.col.size());
    std::vector<int32_t> fillp(t.row_ptr.begin(), t.row_ptr.end() - 1);
    for(int32_t r = 0; r < a.m; ++r)
        for(int32_t q = a.row_ptr[r]; q < a.row_ptr[r + 1]; ++q) {
            const int32_t p = fillp[a.col[q]]++;
            t.col[p] = r;
            t.val[p] = a.val[q];
        }
    return t;
}

culip_status culip_spmv_create_csr(const culip_csr* hA, culip_spmv** out)
{
    return guarded([&] {
        CULIP_REQUIRE(out, "spmv: null argument");
        device_present();
        HostCsr h = to_host(hA);
        auto* s = new culip_spmv;
        try {
            upload_csr(s->s.A, s->s.b_row, s->s.b_col, s->s.b_val, h);
            // explicit transpose for A' x (spmv.impl.cuh:57-66), unless A is
            // structurally and numerically symmetric
            HostCsr t = transpose_csr(h);
            const bool symm = t.row_ptr == h.row_ptr && t.col == h.col && t.val == h.val;
            if(!symm) {
                upload_csr(s->s.At, s->s.bt_row, s->s.bt_col, s->s.bt_val, t);
                s->s.have_at = true;
            }
            s->s.m = h.m;
            s->s.n = h.n;
        } catch(...) { delete s; throw; }
        *out = s;
    });
}

culip_status culip_spmv_create_kkt(const culip_csr* hA, const double* d, double delta,
    int32_t n_ix, const int32_t* ix, culip_spmv** out)
{
    return guarded([&] {
        CULIP_REQUIRE(out && d, "spmv: null argument");
        device_present();
        HostCsr a = to_host(hA);
        auto* s = new culip_spmv;
        try {
            Spmv& o = s->s;
            o.kkt = true;
            o.K.n = a.n;
            o.K.m = a.m;
            upload_csr(o.K.A, o.b_arow, o.b_acol, o.b_aval, a);
            upload_csr(o.K.At, o.b_atrow, o.b_atcol, o.b_atval, transpose_csr(a));
            upload(o.b_d, std::vector<double>(d, d + a.n));
            std::vector<double> dix(a.m, 0.0);
            for(int32_t q = 0; q < n_ix; ++q) {
                CULIP_REQUIRE(ix && ix[q] >= 0 && ix[q] < a.m, "spmv: ix out of range");
                dix[ix[q]] += delta;
            }
            upload(o.b_dix, dix);
            o.K.d = o.b_d.as<double>();
            o.K.dix = o.b_dix.as<double>();
            o.m = o.n = a.n + a.m;
        } catch(...) { delete s; throw; }
        *out = s;
    });
}

culip_status culip_spmv_create_kkt_csr(const culip_csr* hH, const culip_csr* hA, double delta,
    int32_t n_ix, const int32_t* ix, culip_spmv** out)
{
    return guarded([&] {
        CULIP_REQUIRE(out, "spmv: null argument");
        device_present();
        HostCsr h = to_host(hH);
        HostCsr a = to_host(hA);
        CULIP_REQUIRE(h.m == h.n && h.n == a.n, "spmv: H must be n x n with A m x n");
        auto* s = new culip_spmv;
        try {
            Spmv& o = s->s;
            o.kkt = true;
            o.K.n = a.n;
            o.K.m = a.m;
            upload_csr(o.K.H, o.b_hrow, o.b_hcol, o.b_hval, h);
            upload_csr(o.K.A, o.b_arow, o.b_acol, o.b_aval, a);
            upload_csr(o.K.At, o.b_atrow, o.b_atcol, o.b_atval, transpose_csr(a));
            std::vector<double> dix(a.m, 0.0);
            for(int32_t q = 0; q < n_ix; ++q) {
                CULIP_REQUIRE(ix && ix[q] >= 0 && ix[q] < a.m, "spmv: ix out of range");
                dix[ix[q]] += delta;
            }
            upload(o.b_dix, dix);
            o.K.dix = o.b_dix.as<double>();
            o.m = o.n = a.n + a.m;
        } catch(...) { delete s; throw; }
        *out = s;
    });
}

culip_status culip_spmv_create_normal(const culip_csr* hA, const double* d, double delta,
    int32_t n_ix, const int32_t* ix, culip_spmv** out)
{
    return guarded([&] {
        CULIP_REQUIRE(out && d, "spmv: null argument");
        HostCsr a = to_host(hA);
        std::vector<double> dix(a.m, 0.0);
        for(int32_t q = 0; q < n_ix; ++q) {
            CULIP_REQUIRE(ix && ix[q] >= 0 && ix[q] < a.m, "spmv: ix out of range");
            dix[ix[q]] += delta;
        }
        device_present();
        auto* s = new culip_spmv;
        try {
            Spmv& o = s->s;
            o.normal = true;
            o.N.m = a.m;
            o.N.n = a.n;
            upload_csr(o.N.A, o.b_arow, o.b_acol, o.b_aval, a);
            upload_csr(o.N.At, o.b_atrow, o.b_atcol, o.b_atval, transpose_csr(a));
            upload(o.b_d, std::vector<double>(d, d + a.n));
            upload(o.b_dix, dix);
            o.b_t.ensure(sizeof(double) * (size_t) std::max(a.n, 1));
            o.N.d = o.b_d.as<double>();
            o.N.dix = o.b_dix.as<double>();
            o.N.t = o.b_t.as<double>();
            o.m = o.n = a.m;
        } catch(...) { delete s; throw; }
        *out = s;
    });
}

culip_status culip_spmv_update_values(culip_spmv* op, const double* vals, int32_t on_device,
    void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(op && vals && !op->s.kkt && !op->s.normal,
            "spmv: values can be replaced on CSR operators only");
        HIP_CHECK(hipMemcpyAsync(op->s.A.val, vals, sizeof(double) * (size_t) op->s.A.nnz,
            on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, (hipStream_t) stream));
        CULIP_REQUIRE(!op->s.have_at, "spmv: value update of a non-symmetric operator");
    });
}

culip_status culip_spmv_multiply(const culip_spmv* op, const double* x, double* y, double alpha,
    double beta, int32_t transpose, void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(op && x && y, "spmv: null argument");
        op->s.multiply(x, y, alpha, beta, transpose != 0, (hipStream_t) stream);
        HIP_CHECK(hipGetLastError());
    });
}

culip_status culip_spmv_equilibrate(culip_spmv* op, int32_t sweeps, double* d, void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(op && d, "spmv: null argument");
        device_present();
        op->s.equilibrate(sweeps, d, (hipStream_t) stream);
    });
}

int32_t culip_spmv_m(const culip_spmv* op) { return op ? op->s.m : -1; }
int32_t culip_spmv_n(const culip_spmv* op) { return op ? op->s.n : -1; }
void culip_spmv_destroy(culip_spmv* op) { delete op; }

// ---- SQMR -----------------------------------------------------------------

culip_status culip_sqmr_create(const culip_spmv* A, const culip_bildlt* prec,
    const culip_sqmr_opts* opts, culip_sqmr** out)
{
    return guarded([&] {
        CULIP_REQUIRE(A && opts && out, "sqmr: null argument");
        CULIP_REQUIRE(A->s.m == A->s.n, "sqmr: operator must be square");
        CULIP_REQUIRE(!prec || prec->b.n() == A->s.m, "sqmr: preconditioner size mismatch");
        CULIP_REQUIRE(opts->max_it >= 0 && opts->tol >= 0, "sqmr: bad options");
        device_present();
        auto* s = new culip_sqmr;
        try {
            s->q.init(&A->s, prec ? &prec->b : nullptr, *opts);
        } catch(...) { delete s; throw; }
        *out = s;
    });
}

culip_status culip_sqmr_solve(culip_sqmr* s, const double* b, double* x, culip_sqmr_result* res,
    void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(s && b && x && res, "sqmr: null argument");
        s->q.run(b, x, *res, (hipStream_t) stream);
    });
}

culip_status culip_sqmr_set_residual_scale(culip_sqmr* s, const double* scale)
{
    return guarded([&] {
        CULIP_REQUIRE(s, "sqmr: null argument");
        s->q.res_scale = scale;
    });
}

void culip_sqmr_destroy(culip_sqmr* s) { delete s; }

// ---- imported factor -------------------------------------------------------

culip_status culip_ldu_create(const culip_csr* L, const culip_csr* U, const double* tri,
    const int32_t* is_piv, int32_t exact_solve, culip_prec prec, culip_ldu** out)
{
    return guarded([&] {
        CULIP_REQUIRE(L && out, "ldu: null argument");
        CULIP_REQUIRE(prec == CULIP_PREC_F64 || prec == CULIP_PREC_F32, "ldu: unknown precision");
        HostCsr l = to_host(L);
        HostCsr u;
        if(U) u = to_host(U);
        CULIP_REQUIRE(l.m == l.n, "ldu: L must be square");
        CULIP_REQUIRE(!U || (u.m == l.m && u.n == l.n), "ldu: U must match L");
        CULIP_REQUIRE((tri == nullptr) == (is_piv == nullptr),
            "ldu: D needs both the tridiagonal and the pivot flags");
        device_present();
        auto* h = new culip_ldu;
        try {
            h->l.create(l, U ? &u : nullptr, tri, is_piv, exact_solve != 0, prec);
        } catch(...) { delete h; throw; }
        *out = h;
    });
}

culip_status culip_ldu_set_permutation(culip_ldu* h, const int32_t* p_r, const int32_t* p_c)
{
    return guarded([&] {
        CULIP_REQUIRE(h, "ldu: null handle");
        h->l.set_permutation(p_r, p_c);
    });
}

culip_status culip_ldu_set_scaling(culip_ldu* h, const double* s_r, const double* s_c)
{
    return guarded([&] {
        CULIP_REQUIRE(h, "ldu: null handle");
        h->l.set_scaling(s_r, s_c);
    });
}

culip_status culip_ldu_solve(const culip_ldu* h, culip_side side, int32_t transpose,
    const double* b, double* x, void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(h && b && x, "ldu: null argument");
        CULIP_REQUIRE(side == CULIP_SOLVE_LEFT || side == CULIP_SOLVE_MIDDLE || side == CULIP_SOLVE_RIGHT,
            "ldu: unknown side");
        h->l.solve(side, transpose != 0, b, x, (hipStream_t) stream);
        HIP_CHECK(hipGetLastError());
    });
}

int32_t culip_ldu_n(const culip_ldu* h) { return h ? h->l.n : -1; }

int32_t culip_ldu_is_middle(const culip_ldu* h) { return h && h->l.has_d ? 1 : 0; }

void culip_ldu_destroy(culip_ldu* h) { delete h; }

culip_status culip_sqmr_create_ldu(const culip_spmv* A, const culip_ldu* prec,
    const culip_sqmr_opts* opts, culip_sqmr** out)
{
    return guarded([&] {
        CULIP_REQUIRE(A && prec && opts && out, "sqmr: null argument");
        CULIP_REQUIRE(A->s.m == A->s.n, "sqmr: operator must be square");
        CULIP_REQUIRE(prec->l.n == A->s.m, "sqmr: preconditioner size mismatch");
        CULIP_REQUIRE(opts->max_it >= 0 && opts->tol >= 0, "sqmr: bad options");
        device_present();
        auto* s = new culip_sqmr;
        try {
            s->q.init(&A->s, &prec->l, *opts);
        } catch(...) { delete s; throw; }
        *out = s;
    });
}

// ---- request loop --------------------------------------------------------

culip_status culip_server_create(const culip_server_opts* o, culip_server** out)
{
    return guarded([&] {
        CULIP_REQUIRE(o && out, "server: null argument");
        device_present();
        auto* h = new culip_server;
        try {
            h->s = std::make_unique<Server>(*o);
        } catch(...) { delete h; throw; }
        *out = h;
    });
}

culip_status culip_server_add_pattern(culip_server* s, const culip_csr* K, int32_t num_pivs,
    const int32_t* piv_starts, int32_t num_blocks, const int32_t* block_starts, int32_t* pattern)
{
    return guarded([&] {
        CULIP_REQUIRE(s && K && pattern, "server: null argument");
        CULIP_REQUIRE(num_pivs > 0 && piv_starts, "server: pivot starts required");
        CULIP_REQUIRE(num_blocks > 0 && block_starts, "server: block starts required");
        HostCsr h = to_host(K);
        CULIP_REQUIRE(h.m == h.n, "server: matrix must be square");
        *pattern = s->s->add_pattern(h, std::vector<int32_t>(piv_starts, piv_starts + num_pivs),
            std::vector<int32_t>(block_starts, block_starts + num_blocks));
    });
}

culip_status culip_server_set_pattern_solve(culip_server* s, int32_t pattern,
    culip_solve_algorithm algorithm, int32_t jacobi_sweeps)
{
    return guarded([&] {
        CULIP_REQUIRE(s, "server: null handle");
        s->s->set_pattern_solve(pattern, algorithm, jacobi_sweeps);
    });
}

culip_status culip_server_submit(culip_server* s, int32_t pattern, const double* vals,
    const double* rhs, int64_t* request)
{
    return guarded([&] {
        CULIP_REQUIRE(s && request, "server: null argument");
        *request = s->s->submit(pattern, vals, rhs);
    });
}

culip_status culip_server_wait(culip_server* s, int64_t request, double* x, culip_sqmr_result* res,
    double* timing, int32_t* slot)
{
    return guarded([&] {
        CULIP_REQUIRE(s, "server: null handle");
        s->s->wait(request, x, res, timing, slot);
    });
}

culip_status culip_server_stats(const culip_server* s, int64_t* out)
{
    return guarded([&] {
        CULIP_REQUIRE(s && out, "server: null argument");
        s->s->stats(out);
    });
}

void culip_server_destroy(culip_server* s) { delete s; }

culip_status culip_batch_create(const culip_batch_opts* o, culip_batch** out)
{
    return guarded([&] {
        CULIP_REQUIRE(o && out, "batch: null argument");
        device_present();
        BatchSolverOpts bo;
        bo.prec = o->prec;
        bo.fill_level = o->fill_level;
        bo.fill_factor = o->fill_factor;
        bo.threshold = o->threshold;
        bo.sqmr = o->sqmr;
        bo.jacobi_sweeps = o->jacobi_sweeps;
        auto* h = new culip_batch;
        try {
            h->b = std::make_unique<BatchSolver>(bo);
        } catch(...) { delete h; throw; }
        *out = h;
    });
}

culip_status culip_batch_add(culip_batch* b, const culip_csr* K, int32_t num_pivs,
    const int32_t* piv_starts, int32_t num_blocks, const int32_t* block_starts, int32_t* index)
{
    return guarded([&] {
        CULIP_REQUIRE(b && K && index, "batch: null argument");
        CULIP_REQUIRE(num_pivs > 0 && piv_starts, "batch: pivot starts required");
        CULIP_REQUIRE(num_blocks > 0 && block_starts, "batch: block starts required");
        HostCsr h = to_host(K);
        *index = b->b->add(h, std::vector<int32_t>(piv_starts, piv_starts + num_pivs),
            std::vector<int32_t>(block_starts, block_starts + num_blocks));
    });
}

culip_status culip_batch_solve(culip_batch* b, const double* const* vals,
    const double* const* rhs, double* const* x, culip_sqmr_result* res, int32_t device_arrays,
    void* stream)
{
    return guarded([&] {
        CULIP_REQUIRE(b, "batch: null handle");
        b->b->solve(vals, rhs, x, res, device_arrays != 0, (hipStream_t) stream);
    });
}

culip_status culip_batch_info(const culip_batch* b, int32_t* count, int64_t* device_bytes,
    double* seconds)
{
    return guarded([&] {
        CULIP_REQUIRE(b, "batch: null handle");
        if(count) *count = b->b->count();
        if(device_bytes) *device_bytes = b->b->device_bytes();
        if(seconds) b->b->last_seconds(seconds);
    });
}

void culip_batch_destroy(culip_batch* b) { delete b; }

}  // extern "C"
