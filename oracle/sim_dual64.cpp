// ORACLE — test infrastructure only (see oracle_math.h).
// The dual-number build whose value part is the FP64 oracle: the gradient of
// the same physics in double precision.
#define DUAL_V double
#define DUAL_COS(v) std::cos(v)
#define DUAL_ATAN2(y, x) std::atan2(y, x)
#define FN(name) d64_##name
#include "sim_dual.hpp"
