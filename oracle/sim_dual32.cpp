// ORACLE — test infrastructure only (see oracle_math.h).
// The dual-number build whose value part is the FP32 oracle bit for bit:
// the reference gradient of the FP32 trajectory the GPU follows.
#define DUAL_V float
#define DUAL_COS(v) orc_cos_0_pi(v)
#define DUAL_ATAN2(y, x) orc_atan2(y, x)
#define FN(name) d32_##name
#include "sim_dual.hpp"
