// Static device scratch of the allocation-free entry points.
#pragma once

#include <cstdint>

namespace culip {

// Largest system (rows) an allocation-free call takes.
constexpr int64_t NOALLOC_MAX_ROWS = int64_t(1) << 20;
// Vectors of NOALLOC_MAX_ROWS doubles in the scratch array.
constexpr int64_t NOALLOC_SCRATCH_VECTORS = 3;

// Device address of the current device's scratch array (looked up on the
// first call per device), or nullptr when the lookup fails.
double* noalloc_scratch();

}  // namespace culip
