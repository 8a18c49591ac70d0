/* Oracle BILDLt, T = double. TEST INFRASTRUCTURE ONLY. */
#include "oracle.h"
#define REAL double
#define OR(x) or_##x##_d
#include "bildlt_oracle.inc"
