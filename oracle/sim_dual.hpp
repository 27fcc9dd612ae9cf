// ORACLE — test infrastructure only (see oracle_math.h).
//
// The oracle compiled with a forward-mode dual number: every real value
// carries, beside its value, its derivatives along ORACLE_DIRS directions in
// double precision. The value part is computed exactly as the plain build
// computes it (float with the oracle's own cos/atan2 for the FP32 build,
// double with libm for the FP64 build), so the trial follows the same
// branches, and the derivative part is the exact derivative of that
// piecewise-smooth program: the reference gradient of the adjoint kernel.
//
// The including file defines DUAL_V (float or double), FN(name) and the
// elementary functions DUAL_COS(v) and DUAL_ATAN2(y, x) on values.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <atomic>
#include <thread>
#include <type_traits>
#include <vector>

#include "oracle_math.h"

#ifndef ORACLE_DIRS
#define ORACLE_DIRS 8
#endif

struct DV {
    DUAL_V v;
    double d[ORACLE_DIRS];
    DV() : v(0) { std::memset(d, 0, sizeof(d)); }
    template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
    DV(T x) : v(static_cast<DUAL_V>(x)) { std::memset(d, 0, sizeof(d)); }
    static DV lit(double x) { return DV(static_cast<DUAL_V>(x)); }
    explicit operator float() const { return static_cast<float>(v); }
    explicit operator double() const { return static_cast<double>(v); }
    explicit operator uint32_t() const { return static_cast<uint32_t>(v); }
};

static inline DV dv(DUAL_V v) { DV r; r.v = v; return r; }

static inline DV operator+(const DV& a, const DV& b) {
    DV r = dv(a.v + b.v);
    for (int i = 0; i < ORACLE_DIRS; i++) { r.d[i] = a.d[i] + b.d[i]; }
    return r;
}
static inline DV operator-(const DV& a, const DV& b) {
    DV r = dv(a.v - b.v);
    for (int i = 0; i < ORACLE_DIRS; i++) { r.d[i] = a.d[i] - b.d[i]; }
    return r;
}
static inline DV operator*(const DV& a, const DV& b) {
    DV r = dv(a.v * b.v);
    const double av = a.v, bv = b.v;
    for (int i = 0; i < ORACLE_DIRS; i++) { r.d[i] = a.d[i] * bv + av * b.d[i]; }
    return r;
}
static inline DV operator/(const DV& a, const DV& b) {
    DV r = dv(a.v / b.v);
    const double av = a.v, bv = b.v;
    for (int i = 0; i < ORACLE_DIRS; i++) { r.d[i] = (a.d[i] * bv - av * b.d[i]) / (bv * bv); }
    return r;
}
static inline DV operator-(const DV& a) {
    DV r = dv(-a.v);
    for (int i = 0; i < ORACLE_DIRS; i++) { r.d[i] = -a.d[i]; }
    return r;
}
static inline DV& operator+=(DV& a, const DV& b) { a = a + b; return a; }
static inline DV& operator-=(DV& a, const DV& b) { a = a - b; return a; }
static inline DV& operator*=(DV& a, const DV& b) { a = a * b; return a; }
static inline DV& operator/=(DV& a, const DV& b) { a = a / b; return a; }
static inline bool operator<(const DV& a, const DV& b) { return a.v < b.v; }
static inline bool operator>(const DV& a, const DV& b) { return a.v > b.v; }
static inline bool operator<=(const DV& a, const DV& b) { return a.v <= b.v; }
static inline bool operator>=(const DV& a, const DV& b) { return a.v >= b.v; }
static inline bool operator==(const DV& a, const DV& b) { return a.v == b.v; }
static inline bool operator!=(const DV& a, const DV& b) { return a.v != b.v; }

static inline DV d_fma(const DV& a, const DV& b, const DV& c) {
    DV r = dv(std::fma(a.v, b.v, c.v));
    const double av = a.v, bv = b.v;
    for (int i = 0; i < ORACLE_DIRS; i++) { r.d[i] = a.d[i] * bv + av * b.d[i] + c.d[i]; }
    return r;
}
static inline DV d_sqrt(const DV& a) {
    DV r = dv(std::sqrt(a.v));
    const double s = r.v;
    for (int i = 0; i < ORACLE_DIRS; i++) { r.d[i] = s > 0.0 ? a.d[i] / (2.0 * s) : 0.0; }
    return r;
}
static inline DV d_floor(const DV& a) { return dv(std::floor(a.v)); }
static inline DV d_abs(const DV& a) { return a.v < 0 ? -a : a; }
// fmin / fmax: the derivative of the operand whose value is returned
static inline DV d_min(const DV& a, const DV& b) { return (b.v < a.v || a.v != a.v) ? b : a; }
static inline DV d_max(const DV& a, const DV& b) { return (b.v > a.v || a.v != a.v) ? b : a; }
static inline DV d_copysign(const DV& a, const DV& b) {
    const bool flip = std::signbit(a.v) != std::signbit(b.v);
    return flip ? -a : a;
}
static inline DV d_cos(const DV& a) {
    DV r = dv(DUAL_COS(a.v));
    const double s = -std::sin(static_cast<double>(a.v));
    for (int i = 0; i < ORACLE_DIRS; i++) { r.d[i] = s * a.d[i]; }
    return r;
}
static inline DV d_atan2(const DV& y, const DV& x) {
    DV r = dv(DUAL_ATAN2(y.v, x.v));
    const double yv = y.v, xv = x.v, q = xv * xv + yv * yv;
    for (int i = 0; i < ORACLE_DIRS; i++) { r.d[i] = q > 0.0 ? (xv * y.d[i] - yv * x.d[i]) / q : 0.0; }
    return r;
}

#define R DV
#define FMA d_fma
#define SQRT d_sqrt
#define FLOOR d_floor
#define ABS d_abs
#define MIN d_min
#define MAX d_max
#define COPYSIGN d_copysign
#define COS_0_PI d_cos
#define ATAN2 d_atan2
#define LIT(x) DV::lit(x)
#define SIM_FRAME_T float
#define SIM_FRAME_BITS(p, lo, hi) do { (p)[0] = orc_as_float(lo); (p)[1] = orc_as_float(hi); } while (0)
#pragma GCC diagnostic ignored "-Wclass-memaccess"
#include "sim_impl.h"

// The parameters a gradient is taken with respect to, in the order of
// evo_adjoint_device: per muscle k, six values (rhythm amplitude, inverse
// period, phase, Hill factor, force cap, inverse energy capacity) at
// 6k .. 6k+5; then per node i its friction at 6M + i.
static DV* param_slot(FN(Body)* b, uint32_t p) {
    const uint32_t m = b->muscles;
    if (p < 6u * m) {
        const uint32_t k = p / 6u;
        switch (p % 6u) {
            case 0: return &b->mamp[k];
            case 1: return &b->minvp[k];
            case 2: return &b->mphase[k];
            case 3: return &b->mhill[k];
            case 4: return &b->mcap[k];
            default: return &b->minvcap[k];
        }
    }
    return &b->nf[p - 6u * m];
}

extern "C" {

// The trial of creature cidx with derivatives of its fitness along
// `ndirs` (<= ORACLE_DIRS) directions in parameter space: dirs[j * nparams + p]
// is direction j's component on parameter p. Writes the fitness, the
// derivatives and the steps run. Returns 0, or -1 for a record the
// simulator cannot run or a parameter count that does not match the body.
int FN(gradient)(const uint32_t* records, const float* muscles, const uint32_t* heads, const evo_params* p,
                 const evo_run* run, uint32_t cidx, const double* dirs, uint32_t ndirs, uint32_t nparams,
                 double* fitness, double* derivs, long* steps_out) {
    if (ndirs > ORACLE_DIRS) { return -1; }
    FN(World) w;
    FN(make_world)(&w, p, run);
    std::vector<FN(Body)> bv(1);
    std::vector<FN(Trial)> tv(1);
    FN(Body)* b = bv.data();
    FN(Trial)* tr = tv.data();
    if (FN(begin_creature)(&w, b, tr, cidx, records, muscles, heads, nullptr) != 0) { return -1; }
    if (nparams != 6u * b->muscles + b->nodes) { return -1; }
    for (uint32_t q = 0; q < nparams; q++) {
        DV* slot = param_slot(b, q);
        for (uint32_t j = 0; j < ndirs; j++) { slot->d[j] = dirs[j * nparams + q]; }
    }
    evo_result out;
    long steps = 0;
    for (;;) {
        steps++;
        if (FN(run_step)(&w, b, tr, muscles, &out, nullptr)) { break; }
    }
    *fitness = static_cast<double>(tr->fitness.v);
    for (uint32_t j = 0; j < ndirs; j++) { derivs[j] = tr->fitness.d[j]; }
    if (steps_out) { *steps_out = steps; }
    return 0;
}

// The trials of creatures list[0..n) with one derivative direction each
// (dirs holds each creature's nparams[i] components one after the other),
// on `threads` host threads: fitness[i] and derivs[i]. Returns 0, or -1 if a
// creature cannot be differentiated.
int FN(gradient_batch)(const uint32_t* records, const float* muscles, const uint32_t* heads, const evo_params* p,
                       const evo_run* run, const uint32_t* list, uint32_t n, const double* dirs,
                       const uint32_t* offsets, int threads, double* fitness, double* derivs) {
    std::atomic<uint32_t> next{0};
    std::atomic<int> bad{0};
    auto work = [&]() {
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= n) { break; }
            const uint32_t c = list[i];
            const uint32_t np_ = 6u * heads[8u * c + 1u] + (heads[8u * c] & 255u);
            if (FN(gradient)(records, muscles, heads, p, run, c, dirs + offsets[i], 1u, np_, fitness + i, derivs + i,
                             nullptr) != 0) {
                bad = 1;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) { pool.emplace_back(work); }
    work();
    for (auto& t : pool) { t.join(); }
    return bad ? -1 : 0;
}

}  // extern "C"
