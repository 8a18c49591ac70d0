// Shared host-side definitions for the culip_amd library.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "culip_amd.h"

namespace culip {

// Error raised inside the library; the C-ABI turns it into a status code and
// keeps the message for culip_last_error().
struct Error : std::runtime_error {
    culip_status status;
    Error(culip_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

[[noreturn]] inline void fail(culip_status s, const std::string& msg) { throw Error(s, msg); }

#define CULIP_REQUIRE(cond, msg)                                                \
    do {                                                                        \
        if(!(cond)) ::culip::fail(CULIP_ERR_INVALID, std::string(msg));         \
    } while(0)

// Host CSR with int32 indices (mat_int_t in the reference, defines.h:56-62).
struct HostCsr {
    int32_t m = 0, n = 0;
    std::vector<int32_t> row_ptr;  // m + 1
    std::vector<int32_t> col;      // nnz, sorted per row
    std::vector<double> val;       // nnz
    int64_t nnz() const { return (int64_t) col.size(); }
};

}  // namespace culip
