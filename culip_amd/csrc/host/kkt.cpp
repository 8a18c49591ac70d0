// Synthetic KKT generator and input preparation (ordering, blocking, static
// pivot starts). The reference takes these as MatrixMarket inputs
// (apps/block_ildlt/block_ildlt_invariants.h:88-111); here they are produced
// deterministically from a seed.
//
// PRNG: xoshiro256** seeded by SplitMix64(seed). Uniform doubles use the top
// 53 bits; normals use Box-Muller (cosine branch only). Draw order:
//   1. H diagonal, i = 0..n-1
//   2. H off-diagonal pairs: repeat {i, j, value} until n*o/2 distinct pairs
//   3. A: for each column j: c distinct rows, then c values
//   4. rhs, N(0,1) for all n + m rows
//   5. barrier weights (if requested), i = 0..n-1
#include "kkt.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_set>

namespace culip {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t& x)
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for(auto& v : s) v = splitmix(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double u01() { return (double) (next() >> 11) * 0x1.0p-53; }
    int32_t below(int32_t k) { return (int32_t) (u01() * (double) k); }
    double normal()
    {
        const double u1 = 1.0 - u01();  // (0, 1]
        const double u2 = u01();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

static void build_csr(int32_t N, std::vector<std::vector<std::pair<int32_t, double>>>& rows,
    HostCsr& out)
{
    out.m = out.n = N;
    out.row_ptr.assign(N + 1, 0);
    for(int32_t i = 0; i < N; ++i) {
        auto& r = rows[i];
        std::sort(r.begin(), r.end(),
            [](const auto& a, const auto& b) { return a.first < b.first; });
        out.row_ptr[i + 1] = out.row_ptr[i] + (int32_t) r.size();
    }
    out.col.resize(out.row_ptr[N]);
    out.val.resize(out.row_ptr[N]);
    for(int32_t i = 0; i < N; ++i) {
        int64_t q = out.row_ptr[i];
        for(auto& e : rows[i]) { out.col[q] = e.first; out.val[q] = e.second; ++q; }
    }
}

Kkt* kkt_generate(const culip_kkt_params& p)
{
    CULIP_REQUIRE(p.n > 0 && p.m >= 0, "kkt: n must be > 0 and m >= 0");
    CULIP_REQUIRE(p.h_offdiag_per_row >= 0 && p.a_nnz_per_col >= 0, "kkt: negative counts");
    CULIP_REQUIRE(p.m == 0 || p.a_nnz_per_col <= p.m, "kkt: a_nnz_per_col > m");
    CULIP_REQUIRE((int64_t) p.n * p.h_offdiag_per_row / 2 <= (int64_t) p.n * (p.n - 1) / 2,
        "kkt: too many H off-diagonals");
    auto* k = new Kkt;
    k->p = p;
    const int32_t n = p.n, m = p.m, N = n + m;
    Rng rng(p.seed);
    std::vector<std::vector<std::pair<int32_t, double>>> rows(N);

    // 1. H diagonal
    std::vector<double> hd(n);
    for(int32_t i = 0; i < n; ++i) {
        const double u = rng.u01();
        hd[i] = p.h_diag_log ?
            std::exp(std::log(p.h_diag_lo) + u * (std::log(p.h_diag_hi) - std::log(p.h_diag_lo))) :
            p.h_diag_lo + u * (p.h_diag_hi - p.h_diag_lo);
    }
    // 2. H off-diagonal pairs (distinct, unordered)
    const int64_t npairs = (int64_t) n * p.h_offdiag_per_row / 2;
    std::unordered_set<uint64_t> seen;
    seen.reserve((size_t) npairs * 2 + 16);
    std::vector<std::pair<int32_t, int32_t>> pairs;
    std::vector<double> pv;
    pairs.reserve(npairs);
    pv.reserve(npairs);
    while((int64_t) pairs.size() < npairs) {
        int32_t i = rng.below(n), j = rng.below(n);
        const double v = rng.normal() * p.h_offdiag_std;
        if(i == j) continue;
        if(i > j) std::swap(i, j);
        const uint64_t key = ((uint64_t) i << 32) | (uint32_t) j;
        if(!seen.insert(key).second) continue;
        pairs.emplace_back(i, j);
        pv.push_back(v);
    }
    for(int32_t i = 0; i < n; ++i) rows[i].emplace_back(i, hd[i]);
    for(size_t q = 0; q < pairs.size(); ++q) {
        rows[pairs[q].first].emplace_back(pairs[q].second, pv[q]);
        rows[pairs[q].second].emplace_back(pairs[q].first, pv[q]);
    }
    // 3. A (m x n), c distinct rows per column
    for(int32_t j = 0; j < n && m > 0; ++j) {
        int32_t rsel[64];
        std::vector<int32_t> rbig;
        int32_t* rs = rsel;
        if(p.a_nnz_per_col > 64) { rbig.resize(p.a_nnz_per_col); rs = rbig.data(); }
        int32_t cnt = 0;
        while(cnt < p.a_nnz_per_col) {
            const int32_t r = rng.below(m);
            bool dup = false;
            for(int32_t q = 0; q < cnt; ++q) dup |= (rs[q] == r);
            if(!dup) rs[cnt++] = r;
        }
        for(int32_t q = 0; q < cnt; ++q) {
            const double v = rng.normal();
            rows[n + rs[q]].emplace_back(j, v);   // A(r, j)
            rows[j].emplace_back(n + rs[q], v);   // A'(j, r)
        }
    }
    for(int32_t i = 0; i < m; ++i) rows[n + i].emplace_back(n + i, -p.delta);
    // 4. rhs
    k->rhs.resize(N);
    for(int32_t i = 0; i < N; ++i) k->rhs[i] = rng.normal();

    build_csr(N, rows, k->K);
    k->diag_pos.resize(n);
    for(int32_t i = 0; i < n; ++i)
        for(int32_t q = k->K.row_ptr[i]; q < k->K.row_ptr[i + 1]; ++q)
            if(k->K.col[q] == i) k->diag_pos[i] = q;
    k->h0 = hd;
    // 5. barrier term
    if(p.barrier_hi > 0) {
        CULIP_REQUIRE(p.barrier_lo > 0 && p.barrier_hi >= p.barrier_lo,
            "kkt: barrier range must be positive");
        const double a = std::log(p.barrier_lo), b = std::log(p.barrier_hi);
        for(int32_t i = 0; i < n; ++i) k->K.val[k->diag_pos[i]] += std::exp(a + rng.u01() * (b - a));
    }
    return k;
}

void kkt_barrier_values(const Kkt& k, double wlo, double whi, uint64_t seed, double* out)
{
    CULIP_REQUIRE(wlo > 0 && whi >= wlo, "barrier weights must be positive");
    std::memcpy(out, k.K.val.data(), sizeof(double) * k.K.val.size());
    Rng rng(seed);
    const double a = std::log(wlo), b = std::log(whi);
    for(int32_t i = 0; i < k.p.n; ++i) out[k.diag_pos[i]] = k.h0[i] + std::exp(a + rng.u01() * (b - a));
}

// ---------------------------------------------------------------------------
// Input preparation

// Contiguous blocks of at most B rows over [0, n_primal) and [n_primal, N).
static std::vector<int32_t> split_blocks(int32_t N, int32_t n_primal, int32_t B)
{
    std::vector<int32_t> bs;
    for(int32_t s = 0; s < n_primal; s += B) bs.push_back(s);
    for(int32_t s = n_primal; s < N; s += B) bs.push_back(s);
    return bs;
}

Prep* prep_create(const HostCsr& K, int32_t n_primal, int32_t B, culip_ordering ord)
{
    const int32_t N = K.m;
    CULIP_REQUIRE(K.m == K.n, "prep: matrix must be square");
    CULIP_REQUIRE(n_primal >= 0 && n_primal <= N, "prep: bad n_primal");
    CULIP_REQUIRE(B >= 1 && B <= 32, "prep: max_block must be in [1, 32]");
    auto* p = new Prep;
    std::vector<int32_t> bs = split_blocks(N, n_primal, B);
    const int32_t nb = (int32_t) bs.size();
    const int32_t nbp = (int32_t) std::count_if(bs.begin(), bs.end(),
        [&](int32_t s) { return s < n_primal; });
    std::vector<int32_t> bend(nb);
    for(int32_t b = 0; b < nb; ++b) bend[b] = (b + 1 < nb) ? bs[b + 1] : N;

    // block order: natural, or greedy colouring of the primal block graph
    // (blocks joined by an entry of K between them) with primal blocks sorted
    // by (colour, index) and dual blocks kept after them in natural order.
    std::vector<int32_t> border(nb);
    std::iota(border.begin(), border.end(), 0);
    if(ord == CULIP_ORDER_BLOCK_COLOR && nbp > 1) {
        std::vector<int32_t> blk_of(N);
        for(int32_t b = 0; b < nb; ++b)
            for(int32_t r = bs[b]; r < bend[b]; ++r) blk_of[r] = b;
        std::vector<int32_t> color(nbp, -1), mark;
        std::vector<int32_t> stamp;
        for(int32_t b = 0; b < nbp; ++b) {
            // colours used by already-coloured neighbours
            for(int32_t r = bs[b]; r < bend[b]; ++r)
                for(int32_t q = K.row_ptr[r]; q < K.row_ptr[r + 1]; ++q) {
                    const int32_t nbk = blk_of[K.col[q]];
                    if(nbk < nbp && nbk != b && color[nbk] >= 0) {
                        if((int32_t) stamp.size() <= color[nbk]) stamp.resize(color[nbk] + 1, -1);
                        stamp[color[nbk]] = b;
                    }
                }
            int32_t c = 0;
            while(c < (int32_t) stamp.size() && stamp[c] == b) ++c;
            color[b] = c;
        }
        std::stable_sort(border.begin(), border.begin() + nbp,
            [&](int32_t a, int32_t b) { return color[a] < color[b]; });
    } else {
        CULIP_REQUIRE(ord == CULIP_ORDER_NATURAL || ord == CULIP_ORDER_BLOCK_COLOR,
            "prep: unknown ordering");
    }
    p->perm.reserve(N);
    p->block_starts.reserve(nb);
    for(int32_t b : border) {
        p->block_starts.push_back((int32_t) p->perm.size());
        for(int32_t r = bs[b]; r < bend[b]; ++r) p->perm.push_back(r);
    }
    // static 1x1 pivots: every row starts a pivot (block_starts subset)
    p->piv_starts.resize(N);
    std::iota(p->piv_starts.begin(), p->piv_starts.end(), 0);

    // K_perm(i, j) = K(perm[i], perm[j]), columns sorted per row
    std::vector<int32_t> inv(N);
    for(int32_t i = 0; i < N; ++i) inv[p->perm[i]] = i;
    p->K.m = p->K.n = N;
    p->K.row_ptr.assign(N + 1, 0);
    for(int32_t i = 0; i < N; ++i) {
        const int32_t o = p->perm[i];
        p->K.row_ptr[i + 1] = p->K.row_ptr[i] + (K.row_ptr[o + 1] - K.row_ptr[o]);
    }
    p->K.col.resize(K.nnz());
    p->K.val.resize(K.nnz());
    p->entry_src.resize(K.nnz());
    std::vector<std::pair<int32_t, int64_t>> tmp;
    for(int32_t i = 0; i < N; ++i) {
        const int32_t o = p->perm[i];
        tmp.clear();
        for(int32_t q = K.row_ptr[o]; q < K.row_ptr[o + 1]; ++q) tmp.emplace_back(inv[K.col[q]], q);
        std::sort(tmp.begin(), tmp.end());
        int64_t w = p->K.row_ptr[i];
        for(auto& e : tmp) {
            p->K.col[w] = e.first;
            p->K.val[w] = K.val[e.second];
            p->entry_src[w] = e.second;
            ++w;
        }
    }
    return p;
}

}  // namespace culip
