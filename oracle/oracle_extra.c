/* ORACLE — test infrastructure only (see oracle_math.h).
 * Exported wrappers of the elementary routines, for their error-bound and
 * round-trip tests. */
#include <stdint.h>
#include "oracle_math.h"

float oracle_cos_0_pi(float x) { return orc_cos_0_pi(x); }
float oracle_atan2(float y, float x) { return orc_atan2(y, x); }
uint32_t oracle_f2h(float v) { return orc_f2h(v); }
float oracle_h2f(uint32_t h) { return orc_h2f(h); }
void oracle_cos_0_pi_n(const float* x, float* out, long n) { for (long i = 0; i < n; i++) out[i] = orc_cos_0_pi(x[i]); }
void oracle_atan2_n(const float* y, const float* x, float* out, long n) { for (long i = 0; i < n; i++) out[i] = orc_atan2(y[i], x[i]); }
void oracle_f2h_n(const float* v, uint32_t* out, long n) { for (long i = 0; i < n; i++) out[i] = orc_f2h(v[i]); }
void oracle_h2f_n(const uint32_t* h, float* out, long n) { for (long i = 0; i < n; i++) out[i] = orc_h2f(h[i]); }

/* The ground and the rhythm of the FP32 and FP64 builds, for their tests. */
void f32_ground_export(uint32_t flags, const void* params, float x, float amp, float qphase, float* h, float* s);
void f64_ground_export(uint32_t flags, const void* params, double x, double amp, double qphase, double* h, double* s);
float f32_rhythm_export(float t, float inv_period, float phase, float offset, float duty, float inv_duty,
                        float inv_complement);
