// Maximum-product matching by shortest augmenting paths on integer costs.
//
// Costs are the reference's (matching.impl.cuh:67-103): c_ij = log2(max_k
// |a_kj|) - log2 |a_ij| >= 0. Like the reference they are made integers so
// that reduced costs, zero tests and dual updates are exact; here on a
// fixed grid of 2^-20 in log2 units (int64), finer than the reference's
// global rescaling into [1, INT_MAX / (nnz + 1)], so the scaling is exact
// to about 1e-6 relative. Dual initialisation and the greedy start with a
// one-step pushback follow the reference (:180-280); the augmenting phase
// is a Dijkstra search over the alternating graph with a binary heap
// (ties to lower column indices) in place of the reference's label scans.
#include "matching.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <queue>

namespace culip {

namespace {

constexpr double GRID = 1048576.0;   // cost units per log2 unit
constexpr int64_t INF = std::numeric_limits<int64_t>::max() / 4;

struct HeapItem {
    int64_t d;
    int32_t col;
    bool operator>(const HeapItem& o) const { return d != o.d ? d > o.d : col > o.col; }
};

}  // namespace

MatchingResult weighted_matching(const HostCsr& A)
{
    CULIP_REQUIRE(A.m == A.n, "matching: the matrix must be square");
    const int32_t n = A.m;
    const int64_t nnz = A.nnz();
    MatchingResult R;
    R.row_match.assign(n, -1);
    R.col_match.assign(n, -1);
    R.u.assign(n, 0.0);
    R.v.assign(n, 0.0);
    R.row_scale.assign(n, 1.0);
    R.col_scale.assign(n, 1.0);
    if(n == 0) {
        R.feasible = true;
        return R;
    }
    // column maxima and integer costs; explicit zeros are not edges
    std::vector<double> cmax(n, 0.0);
    for(int64_t q = 0; q < nnz; ++q) cmax[A.col[q]] = std::max(cmax[A.col[q]], std::fabs(A.val[q]));
    std::vector<int64_t> cost(nnz, INF);
    for(int64_t q = 0; q < nnz; ++q) {
        const double a = std::fabs(A.val[q]);
        if(a > 0.0) cost[q] = std::llround((std::log2(cmax[A.col[q]]) - std::log2(a)) * GRID);
    }
    // dual initialisation (matching.impl.cuh:190-212): v_j = min_i c_ij,
    // u_i = min_j c_ij - v_j
    std::vector<int64_t> u(n, INF), v(n, INF);
    for(int32_t i = 0; i < n; ++i)
        for(int32_t q = A.row_ptr[i]; q < A.row_ptr[i + 1]; ++q)
            if(cost[q] < INF) v[A.col[q]] = std::min(v[A.col[q]], cost[q]);
    for(int32_t j = 0; j < n; ++j)
        if(v[j] == INF) return R;   // an empty column: no perfect matching
    for(int32_t i = 0; i < n; ++i)
        for(int32_t q = A.row_ptr[i]; q < A.row_ptr[i + 1]; ++q)
            if(cost[q] < INF) u[i] = std::min(u[i], cost[q] - v[A.col[q]]);
    for(int32_t i = 0; i < n; ++i)
        if(u[i] == INF) return R;   // an empty row
    auto slack = [&](int32_t i, int32_t q) { return cost[q] - u[i] - v[A.col[q]]; };
    std::vector<int32_t>& rm = R.row_match;
    std::vector<int32_t>& cm = R.col_match;
    int32_t matched = 0;
    // greedy start with a one-step pushback (:218-280): a row takes a free
    // zero-slack column, or one whose row can move to another free
    // zero-slack column
    for(int32_t i = 0; i < n; ++i) {
        bool found = false;
        for(int32_t q = A.row_ptr[i]; q < A.row_ptr[i + 1] && !found; ++q) {
            if(cost[q] == INF || slack(i, q) != 0) continue;
            const int32_t col = A.col[q];
            const int32_t other = cm[col];
            if(other < 0) {
                rm[i] = col;
                cm[col] = i;
                ++matched;
                found = true;
                break;
            }
            for(int32_t p = A.row_ptr[other]; p < A.row_ptr[other + 1]; ++p) {
                if(cost[p] == INF || slack(other, p) != 0 || cm[A.col[p]] >= 0) continue;
                rm[other] = A.col[p];
                cm[A.col[p]] = other;
                rm[i] = col;
                cm[col] = i;
                ++matched;
                found = true;
                break;
            }
        }
    }
    // shortest augmenting paths from every unmatched row
    std::vector<int64_t> dist(n, INF);
    std::vector<int32_t> pred(n, -1);
    std::vector<char> done(n, 0);
    std::vector<int32_t> touched, finalized;
    for(int32_t root = 0; root < n; ++root) {
        if(rm[root] >= 0) continue;
        std::priority_queue<HeapItem, std::vector<HeapItem>, std::greater<HeapItem>> heap;
        touched.clear();
        finalized.clear();
        auto relax = [&](int32_t i, int64_t base) {
            for(int32_t q = A.row_ptr[i]; q < A.row_ptr[i + 1]; ++q) {
                if(cost[q] == INF) continue;
                const int32_t k = A.col[q];
                if(done[k]) continue;
                const int64_t nd = base + std::max<int64_t>(0, slack(i, q));
                if(nd < dist[k]) {
                    if(dist[k] == INF) touched.push_back(k);
                    dist[k] = nd;
                    pred[k] = i;
                    heap.push({nd, k});
                }
            }
        };
        relax(root, 0);
        int32_t end = -1;
        int64_t L = 0;
        while(!heap.empty()) {
            const HeapItem top = heap.top();
            heap.pop();
            if(done[top.col] || top.d != dist[top.col]) continue;
            if(cm[top.col] < 0) {
                end = top.col;
                L = top.d;
                break;
            }
            done[top.col] = 1;
            finalized.push_back(top.col);
            relax(cm[top.col], top.d);
        }
        if(end < 0) {
            for(int32_t k : touched) { dist[k] = INF; pred[k] = -1; done[k] = 0; }
            return R;   // structurally singular: no perfect matching
        }
        // duals: D = min(dist, L) on the reached nodes; u += L - D(row),
        // v -= L - D(col) keeps every reduced cost >= 0 and zeroes them on
        // the shortest path (rows share the distance of their matched column)
        u[root] += L;
        for(int32_t j : finalized) {
            const int64_t delta = L - dist[j];
            v[j] -= delta;
            u[cm[j]] += delta;
        }
        // augment along the predecessors
        int32_t j = end;
        for(;;) {
            const int32_t i = pred[j];
            const int32_t next = rm[i];
            rm[i] = j;
            cm[j] = i;
            if(i == root) break;
            j = next;
        }
        ++matched;
        for(int32_t k : touched) { dist[k] = INF; pred[k] = -1; done[k] = 0; }
    }
    R.feasible = matched == n;
    int64_t obj = 0;
    for(int32_t i = 0; i < n; ++i)
        for(int32_t q = A.row_ptr[i]; q < A.row_ptr[i + 1]; ++q)
            if(A.col[q] == rm[i] && cost[q] < INF) obj += cost[q];
    R.objective = (double) obj / GRID;
    for(int32_t i = 0; i < n; ++i) {
        R.u[i] = (double) u[i] / GRID;
        R.row_scale[i] = std::exp2(R.u[i]);
    }
    for(int32_t j = 0; j < n; ++j) {
        R.v[j] = (double) v[j] / GRID;
        R.col_scale[j] = std::exp2(R.v[j]) / cmax[j];
    }
    return R;
}

std::vector<double> symmetric_matching_scaling(const MatchingResult& m)
{
    const size_t n = m.row_scale.size();
    std::vector<double> d(n, 1.0);
    for(size_t i = 0; i < n; ++i) {
        const double s = std::sqrt(m.row_scale[i] * m.col_scale[i]);
        if(std::isfinite(s) && s > 0.0) d[i] = s;
    }
    return d;
}

}  // namespace culip

namespace culip {

int32_t structural_rank(int32_t m, int32_t n, const int32_t* row_ptr, const int32_t* col,
    std::vector<int32_t>* row_match)
{
    std::vector<int32_t> rm(m, -1), cm(n, -1);
    int32_t matched = 0;
    for(int32_t i = 0; i < m; ++i)
        for(int32_t q = row_ptr[i]; q < row_ptr[i + 1]; ++q)
            if(cm[col[q]] < 0) {
                rm[i] = col[q];
                cm[col[q]] = i;
                ++matched;
                break;
            }
    // breadth-first search over alternating paths from each unmatched row
    std::vector<int32_t> col_from(n, -1), seen(n, -1), queue;
    for(int32_t root = 0; root < m; ++root) {
        if(rm[root] >= 0) continue;
        queue.assign(1, root);
        int32_t end = -1;
        for(size_t h = 0; h < queue.size() && end < 0; ++h) {
            const int32_t i = queue[h];
            for(int32_t q = row_ptr[i]; q < row_ptr[i + 1]; ++q) {
                const int32_t j = col[q];
                if(seen[j] == root) continue;
                seen[j] = root;
                col_from[j] = i;
                if(cm[j] < 0) {
                    end = j;
                    break;
                }
                queue.push_back(cm[j]);
            }
        }
        if(end < 0) continue;
        for(int32_t j = end;;) {
            const int32_t i = col_from[j];
            const int32_t next = rm[i];
            rm[i] = j;
            cm[j] = i;
            if(i == root) break;
            j = next;
        }
        ++matched;
    }
    if(row_match) *row_match = rm;
    return matched;
}

namespace {

// Binary min-heap of keys 0..K-1 by integer value with updatable values.
// Ties behave as in the reference's heap (b_kvheap.impl.h): an element
// moves up only past a strictly larger parent, sifts down to the smaller
// child (the right one when both are equal) only when that child is
// strictly smaller, and removal moves the last element into the hole and
// sifts it up or down by the sign of its value's change.
class DegreeHeap {
public:
    explicit DegreeHeap(int32_t keys) : pos_(keys, -1), val_(keys, 0) {}
    void push(int32_t key, int32_t v)
    {
        val_[key] = v;
        pos_[key] = (int32_t) data_.size();
        data_.push_back(key);
        sift_up(key);
    }
    bool empty() const { return data_.empty(); }
    int32_t top_key() const { return data_[0]; }
    int32_t top_value() const { return val_[data_[0]]; }
    int32_t value(int32_t key) const { return pos_[key] >= 0 ? val_[key] : 0; }
    void pop() { remove(data_[0]); }
    void update(int32_t key, int32_t v)
    {
        if(pos_[key] < 0 || v == val_[key]) return;
        const int32_t old = val_[key];
        val_[key] = v;
        if(v < old) sift_up(key);
        else sift_down(key);
    }
    void remove(int32_t key)
    {
        const int32_t p = pos_[key];
        if(p < 0) return;
        const int32_t last = data_.back();
        const int32_t delta = val_[last] - val_[key];
        pos_[key] = -1;
        pos_[last] = p;
        data_[p] = last;
        data_.pop_back();
        if(last == key) return;
        if(delta < 0) sift_up(last);
        else if(delta > 0) sift_down(last);
    }

private:
    void swap_at(int32_t a, int32_t b)
    {
        std::swap(data_[a], data_[b]);
        pos_[data_[a]] = a;
        pos_[data_[b]] = b;
    }
    void sift_up(int32_t key)
    {
        int32_t me = pos_[key];
        while(me > 0) {
            const int32_t par = (me - 1) >> 1;
            if(!(val_[data_[par]] > val_[data_[me]])) break;
            swap_at(par, me);
            me = par;
        }
    }
    void sift_down(int32_t key)
    {
        int32_t me = pos_[key];
        const int32_t sz = (int32_t) data_.size();
        if(me >= sz - 1) return;
        for(;;) {
            const int32_t l = 2 * me + 1, r = 2 * me + 2;
            int32_t c;
            if(r < sz) c = (val_[data_[l]] < val_[data_[r]]) ? l : r;
            else if(l < sz) c = l;
            else break;
            if(!(val_[data_[c]] < val_[data_[me]])) break;
            swap_at(c, me);
            me = c;
        }
    }
    std::vector<int32_t> pos_, val_, data_;
};

}  // namespace

int32_t btf_rank_bound(int32_t m, int32_t n, const int32_t* csc_col_ptr, const int32_t* csc_row,
    const int32_t* csr_row_ptr, const int32_t* csr_col, int32_t* row_order, int32_t* col_order)
{
    std::vector<char> row_active(m, 1), col_active(n, 1);
    DegreeHeap ch(n), rh(m);
    for(int32_t j = 0; j < n; ++j) ch.push(j, csc_col_ptr[j + 1] - csc_col_ptr[j]);
    for(int32_t i = 0; i < m; ++i) rh.push(i, csr_row_ptr[i + 1] - csr_row_ptr[i]);
    std::vector<int32_t> r_rest, c_rest;
    int32_t it = 0, cur_row = -1;
    while(it < std::min(m, n) && !ch.empty() && !rh.empty()) {
        bool upd_rows = false, upd_cols = false;
        const int32_t cur_col = ch.top_key();
        const int32_t deg = ch.top_value();
        if(deg == 1) {
            // a column with one active row: peel the pair
            ch.pop();
            col_order[it] = cur_col + 1;
            col_active[cur_col] = 0;
            for(int32_t q = csc_col_ptr[cur_col]; q < csc_col_ptr[cur_col + 1]; ++q)
                if(row_active[csc_row[q]]) {
                    cur_row = csc_row[q];
                    break;
                }
            row_order[it] = cur_row + 1;
            row_active[cur_row] = 0;
            rh.remove(cur_row);
            upd_rows = upd_cols = true;
            ++it;
        } else if(deg == 0) {
            c_rest.push_back(cur_col);
            col_active[cur_col] = 0;
            ch.pop();
        } else {
            // no column of degree 1: remove the row of smallest degree
            cur_row = rh.top_key();
            rh.pop();
            row_active[cur_row] = 0;
            r_rest.push_back(cur_row);
            upd_cols = true;
        }
        if(upd_cols)
            for(int32_t q = csr_row_ptr[cur_row]; q < csr_row_ptr[cur_row + 1]; ++q) {
                const int32_t c = csr_col[q];
                if(col_active[c]) ch.update(c, ch.value(c) - 1);
            }
        if(upd_rows)
            for(int32_t q = csc_col_ptr[cur_col]; q < csc_col_ptr[cur_col + 1]; ++q) {
                const int32_t r = csc_row[q];
                if(row_active[r]) rh.update(r, rh.value(r) - 1);
            }
    }
    const int32_t peeled = it;
    int32_t ci = it, ri = it;
    for(int32_t j = 0; j < n; ++j)
        if(col_active[j]) col_order[ci++] = j + 1;
    for(int32_t c : c_rest) col_order[ci++] = c + 1;
    for(int32_t i = 0; i < m; ++i)
        if(row_active[i]) row_order[ri++] = i + 1;
    for(int32_t r : r_rest) row_order[ri++] = r + 1;
    return peeled;
}

}  // namespace culip
