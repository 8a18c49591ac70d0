// Synthetic KKT systems and input preparation (host).
#pragma once

#include "common.hpp"

namespace culip {

struct Kkt {
    culip_kkt_params p;
    HostCsr K;                       // full symmetric, sorted columns
    std::vector<double> rhs;
    std::vector<int64_t> diag_pos;   // entry of H_ii in K
    std::vector<double> h0;          // H diagonal before any barrier term
};

struct Prep {
    HostCsr K;                       // K(perm[i], perm[j])
    std::vector<int32_t> perm;       // new -> old
    std::vector<int64_t> entry_src;  // K_perm entry -> K entry
    std::vector<int32_t> block_starts, piv_starts;
};

Kkt* kkt_generate(const culip_kkt_params& p);
void kkt_barrier_values(const Kkt& k, double wlo, double whi, uint64_t seed, double* out);
Prep* prep_create(const HostCsr& K, int32_t n_primal, int32_t B, culip_ordering ord);

}  // namespace culip
