/* ORACLE — test infrastructure only (see oracle_math.h).
 * The FP32 build of sim_impl.h: the bit-exact checker of the HIP kernel. */
#include <math.h>
#include <stddef.h>
#include "oracle_math.h"
#define R float
#define FN(name) f32_##name
#define FMA fmaf
#define SQRT sqrtf
#define FLOOR floorf
#define ABS fabsf
#define MIN fminf
#define MAX fmaxf
#define COPYSIGN copysignf
#define COS_0_PI orc_cos_0_pi
#define ATAN2 orc_atan2
#define LIT(x) x##f
#define SIM_FRAME_T float
#define SIM_FRAME_BITS(p, lo, hi) do { (p)[0] = orc_as_float(lo); (p)[1] = orc_as_float(hi); } while (0)
#include "sim_impl.h"
#include "oracle_api.h"
ORACLE_DEFINE_RUNNER(f32, float)
