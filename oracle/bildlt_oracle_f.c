/* Oracle BILDLt, T = float (BlockiLDLt<float> semantics). TEST
 * INFRASTRUCTURE ONLY. */
#include "oracle.h"
#define REAL float
#define OR(x) or_##x##_f
#include "bildlt_oracle.inc"
