// Maximum-product weighted bipartite matching and the scaling derived from
// its dual variables (the MC64 idea; WeightedBipartiteMatching,
// libs/algorithms/matching.cuh:69-125, matching.impl.cuh:55-160), for input
// preparation of indefinite systems.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace culip {

struct MatchingResult {
    bool feasible = false;           // a perfect matching exists
    std::vector<int32_t> row_match;  // column matched to each row
    std::vector<int32_t> col_match;  // row matched to each column
    std::vector<double> u, v;        // dual variables (log2 scale)
    std::vector<double> row_scale;   // r_i = 2^u_i
    std::vector<double> col_scale;   // c_j = 2^v_j / max_k |a_kj|
    double objective = 0.0;          // sum of costs on the matching
};

// Square n x n matrix A (CSR, sorted or not). Costs c_ij = log2(max_k
// |a_kj|) - log2 |a_ij| on the entries with a_ij != 0; the matching
// minimises their sum over perfect matchings (maximises prod |a_ij|).
// At the optimum u_i + v_j <= c_ij, with equality on matched entries, so
// |r_i a_ij c_j| <= 1 everywhere and = 1 on the matching. Ties are broken
// towards lower indices, so the result is deterministic.
MatchingResult weighted_matching(const HostCsr& A);

// Symmetric scaling of a symmetric A from its matching (Duff and Pralet):
// d_i = sqrt(r_i c_i), so |d_i a_ij d_j| <= 1. Rows without a finite scale
// (structurally singular A) get 1.
std::vector<double> symmetric_matching_scaling(const MatchingResult& m);

// Structural rank of an m x n pattern (sprank, libs/algorithms/sprank.cuh:
// 15-36): the size of a maximum matching of rows to columns, by breadth-
// first augmenting paths from a greedy start. row_match (if not null)
// receives each row's column or -1.
int32_t structural_rank(int32_t m, int32_t n, const int32_t* row_ptr, const int32_t* col,
    std::vector<int32_t>* row_match);

// Lower bound on the rank of an m x n pattern by greedy peeling into a
// triangular form (BTF::permute, libs/algorithms/permute_btf_instances.cc:
// 35-229): while a column of degree 1 remains it is taken with its one
// active row; a column of degree 0 goes to the rest; otherwise the row of
// smallest degree is removed. Columns and rows are kept in binary min-heaps
// keyed by degree whose tie order follows the reference's heap
// (libs/data_structures/b_kvheap.impl.h), so the bound is the reference's.
// row_order / col_order (m and n entries, 1-based as the reference writes
// them) receive the triangular part first, then the remaining and removed
// rows and columns. Returns the number of peeled pairs.
int32_t btf_rank_bound(int32_t m, int32_t n, const int32_t* csc_col_ptr, const int32_t* csc_row,
    const int32_t* csr_row_ptr, const int32_t* csr_col, int32_t* row_order, int32_t* col_order);

}  // namespace culip
