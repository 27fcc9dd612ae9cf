// The creature trial on gfx950: one lane simulates one creature from its
// start pose to the end of its trial (shaders/creature.cu, whole file).
//
// Layout on the chip:
//   - A creature's constants (node masses, radii and friction; bone pivots,
//     lengths, inverse masses and drag; joint topology and ranges; muscle
//     anchors, Hill factor, caps, tendon; muscle energy and demand) live in
//     VGPRs. Every loop over nodes, bones, joints and muscles is unrolled to
//     the body-size class's compile-time maximum and exits per lane at the
//     creature's own count, so each array element has a fixed register.
//   - The moving state (position and velocity of each node as one float4,
//     the muscle force / previous position as one float2, and each muscle's
//     rhythm offset) lives in LDS, one column per lane. A lane reads the
//     nodes that its bones, joints and muscles name; with the column layout
//     the lanes of a wave never meet in a bank whatever nodes they ask for.
//   - The rhythm constants of the muscles stay in the packed records in HBM
//     and are read once a step.
//   - The waves of a launch take creatures from a shared counter in chunks
//     and refill lanes whose creature has ended (shaders/creature.cu:1250-1394).
// The arithmetic follows the oracle's specification expression for
// expression (explicit fused multiply-adds, correctly rounded division and
// square root, the elementary functions of devmath.hpp), so a creature's
// result is the same bit pattern on the GPU as on the CPU oracle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "../../include/evo_types.h"
#include "devmath.hpp"

namespace evo {
namespace dev {

constexpr uint32_t kAnyWorld = 0xffffffffu;

// Calls f(integral_constant<int, I>) for I = 0 .. N-1, each call a separate
// copy of the body, so array indices inside it are compile-time constants.
template <int I, int N, class F>
__device__ __forceinline__ void unroll_from(F& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>());
        unroll_from<I + 1, N>(f);
    }
}
template <int N, class F>
__device__ __forceinline__ void unroll(F&& f) {
    unroll_from<0, N>(f);
}
template <int I, int N, class F>
__device__ __forceinline__ void unroll_guarded(uint32_t count, F& f) {
    if constexpr (I < N) {
        if (static_cast<uint32_t>(I) < count) { f(std::integral_constant<int, I>()); }
        unroll_guarded<I + 1, N>(count, f);
    }
}
// The loops over a body's nodes, bones and muscles: items START .. count-1
// of at most N. A small body-size class unrolls them to N guarded copies
// (constants in registers); a large class runs them as loops (constants in
// per-lane scratch memory, which a wave reads with one coalesced access per
// item).
template <class C, int N, int START, class F>
__device__ __forceinline__ void loop(uint32_t count, F&& f) {
    if constexpr (C::unrolled) {
        unroll_guarded<START, N>(count, f);
    } else {
        const uint32_t end = count < static_cast<uint32_t>(N) ? count : static_cast<uint32_t>(N);
        for (uint32_t i = START; i < end; i++) { f(static_cast<int>(i)); }
    }
}

template <uint32_t RATE, uint32_t SUB, uint32_t WORLD, int MAXN_, int MAXM_, bool RECORD, int BLOCK_, bool UNROLLED>
struct Cfg {
    static constexpr bool unrolled = UNROLLED;
    static constexpr int MAXN = MAXN_;
    static constexpr int MAXB = MAXN_ - 1;
    static constexpr int MAXM = MAXM_;
    static constexpr int BLOCK = BLOCK_;
    static constexpr bool record = RECORD;
    static constexpr uint32_t world = WORLD;
    static constexpr uint32_t rate_u = RATE;
    static constexpr uint32_t substeps = SUB;
    static constexpr float rate = static_cast<float>(RATE);
    static constexpr float dt = 1.0f / rate;
    static constexpr float inv_h = rate * static_cast<float>(SUB);
    static constexpr float h = 1.0f / inv_h;
    static constexpr float half_h = 0.5f * h;
    static constexpr float damp_share = h * 10.0f < 1.0f ? h * 10.0f : 1.0f;
    static constexpr uint32_t settle = (200u * RATE + 119u) / 120u;
    static constexpr uint32_t sample = RATE / 30u > 1u ? RATE / 30u : 1u;
    static constexpr uint32_t head_sample =
        static_cast<uint32_t>(rate / 60.0f + 0.5f) == 0u ? 1u : static_cast<uint32_t>(rate / 60.0f + 0.5f);
    static constexpr uint32_t rung1 = RATE - 1u;
    static constexpr uint32_t rung2 = static_cast<uint32_t>(2.5f * rate) - 1u;
    static constexpr uint32_t rung3 = 5u * RATE - 1u;
    static constexpr uint32_t rung4 = 10u * RATE - 1u;
    static constexpr uint32_t half_s = RATE / 2u;
};

// Whether effect F is on: fixed for a world-specialized kernel, read from
// the launch's flag word (a scalar register) for the any-world kernel.
template <class C, uint32_t F>
__device__ __forceinline__ bool has(uint32_t flags) {
    if constexpr (C::world != kAnyWorld) {
        return (C::world & F) != 0u;
    } else {
        return (flags & F) != 0u;
    }
}

__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ float fracf(float x) { return x - floorf(x); }
__device__ __forceinline__ float wrapf(float a) {
    const float tau = 6.28318530718f;
    return __builtin_fmaf(-tau, floorf((a + 3.14159265359f) * (1.0f / tau)), a);
}

// shaders/creature.cu:153-159
__device__ __forceinline__ float rhythm(float t, float inv_period, float phase, float offset, float duty,
                                        float inv_duty, float inv_complement) {
    const float ph = fracf(__builtin_fmaf(t, inv_period, phase) + offset);
    const float pi = 3.14159265359f;
    if (ph < duty) { return __builtin_fmaf(-0.5f, cos_0_pi(pi * ph * inv_duty), 0.5f); }
    return __builtin_fmaf(0.5f, cos_0_pi(pi * (ph - duty) * inv_complement), 0.5f);
}

// shaders/creature.cu:166-226
template <class C>
__device__ __forceinline__ float2 ground(float x, float amp, float qphase, const evo_params& p, uint32_t flags) {
    float height = 0.0f, slope = 0.0f;
    if (has<C, EVO_TERRAIN>(flags)) {
        const float t0 = __builtin_fmaf(x, 1.0f / 1.1f, qphase);
        const float u0 = fracf(t0);
        const float w0 = u0 * (1.0f - u0);
        const float t1 = __builtin_fmaf(x, 1.0f / 0.43f, 0.3f) + qphase;
        const float u1 = fracf(t1);
        const float w1 = u1 * (1.0f - u1);
        height = amp * __builtin_fmaf(0.65f * 16.0f * w0, w0, 0.35f * 16.0f * w1 * w1);
        slope = amp * __builtin_fmaf(0.65f * 32.0f * w0 * (1.0f - 2.0f * u0), 1.0f / 1.1f,
                                     0.35f * 32.0f * w1 * (1.0f - 2.0f * u1) * (1.0f / 0.43f));
    }
    if (has<C, EVO_SLOPE>(flags)) {
        height = __builtin_fmaf(p.slope, x, height);
        slope += p.slope;
    }
    if (has<C, EVO_GAPS>(flags)) {
        const float spacing = __builtin_fmaf(4.0f, p.gaps, 2.0f);
        const float center = spacing * 0.5f;
        const float r = x - floorf(x / spacing) * spacing;
        const float distance = fabsf(r - center);
        const float half_w = 0.5f * p.gaps;
        const float run = fmaxf(fminf(0.15f, half_w), 1e-6f);
        float factor = clampf((half_w - distance) / run, 0.0f, 1.0f);
        if (distance <= half_w - run) { factor = 1.0f; } else if (distance >= half_w) { factor = 0.0f; }
        height = __builtin_fmaf(-2.0f, factor, height);
        if (distance > half_w - run && distance < half_w) { slope -= 2.0f * ((r < center ? 1.0f : -1.0f) / run); }
    }
    if (has<C, EVO_HURDLES>(flags)) {
        const float spacing = 3.0f;
        const float center = spacing * 0.5f;
        const float r = x - floorf(x / spacing) * spacing;
        const float distance = fabsf(r - center);
        const float half_w = 0.5f * 1.2f;
        const float run = 0.2f;
        float factor = clampf((half_w + run - distance) / run, 0.0f, 1.0f);
        if (distance <= half_w) { factor = 1.0f; } else if (distance >= half_w + run) { factor = 0.0f; }
        height = __builtin_fmaf(p.hurdles, factor, height);
        if (distance > half_w && distance < half_w + run) { slope += p.hurdles * ((r < center ? 1.0f : -1.0f) / run); }
    }
    return make_float2(height, slope);
}

// shaders/creature.cu:230-235
__device__ __forceinline__ float ice_at(float x) {
    const float w = fracf(x * (1.0f / 6.0f));
    const float t = fabsf(w - 0.5f) * 2.0f;
    const float s = clampf((0.7f - t) * 2.5f, 0.0f, 1.0f);
    return s * s * __builtin_fmaf(-2.0f, s, 3.0f);
}

// One creature's constants and per-substep muscle state, in registers.
template <class C>
struct Body {
    uint32_t nodes, bones, muscles, feet;
    float inv_mass, amp, qphase;
    float nm[C::MAXN], nw[C::MAXN], nr[C::MAXN], nf[C::MAXN];
    uint32_t bi0[C::MAXB];
    float blen[C::MAXB], bw0[C::MAXB], bk[C::MAXB], blim[C::MAXB], br0[C::MAXB];
    uint32_t jtopo[C::MAXB];
    float jlo[C::MAXB], jhi[C::MAXB], jwf[C::MAXB];
    uint32_t mnodes[C::MAXM];
    float maa[C::MAXM], mab[C::MAXM], mhill[C::MAXM], mcap[C::MAXM], minvcap[C::MAXM], mtk[C::MAXM],
        mslack[C::MAXM];
    float mstore[C::MAXM], mdemand[C::MAXM];
    float mpull[C::record ? C::MAXM : 1];
    float normal_force[C::record ? C::MAXN : 1], friction_force[C::record ? C::MAXN : 1];
};

// The lane's columns of the LDS tables: node i's position and velocity, its
// force-then-previous-position pair, and muscle k's rhythm offset.
template <class C>
struct Cols {
    float4* pv;
    float2* fp;
    float* off;
    __device__ __forceinline__ float4& p(uint32_t i) const { return pv[i * C::BLOCK]; }
    __device__ __forceinline__ float2& f(uint32_t i) const { return fp[i * C::BLOCK]; }
    __device__ __forceinline__ float& o(uint32_t k) const { return off[k * C::BLOCK]; }
};

// The trial so far (shaders/creature.cu:783-826).
struct Trial {
    float fitness, ground_contact, vmin, vmax, previous_center_y, vertical_extremum, vertical_trend, gait_turns;
    float height_sum, fall_time, head_shake_out, screen_x, screened;
    uint32_t contact_lo, contact_hi, lift_lo, lift_hi, ground_lo, ground_hi;
    uint32_t contact_bits, lift_bits, ground_bits;
    float head_shake, head_at_x, head_at_y, head_vel_x, head_vel_y;
    uint32_t rung_speed, rung_early, rung_late, rung_bits;
    uint32_t flags, step, mus_off, cidx;
    float rung_x, period_f;
    bool limp, scoring, live;
    evo_result kept;
};

// shaders/creature.cu:320-356
struct Joint {
    uint32_t tip, far_, hub;
    float gtx, gty, gfx, gfy, ghx, ghy, angle, lo, hi, wt, wf, wh;
};

template <class C>
__device__ __forceinline__ Joint joint(const Body<C>& b, const Cols<C>& s, int J, bool with_angle) {
    Joint k;
    const uint32_t topo = b.jtopo[J];
    k.tip = static_cast<uint32_t>(J) + 1u;
    k.hub = topo & 31u;
    const uint32_t ptip = (topo >> 5u) & 31u, ppiv = (topo >> 10u) & 31u;
    k.far_ = (topo >> 15u) & 31u;
    const bool at_head = ((topo >> 20u) & 1u) != 0u;
    k.lo = b.jlo[J];
    k.hi = b.jhi[J];
    k.wt = b.nw[J + 1];
    k.wf = b.jwf[J];
    k.wh = b.bw0[J];
    const float4 pt = s.p(ptip), pp = s.p(ppiv), tt = s.p(k.tip), hh = s.p(k.hub);
    const float ux = pt.x - pp.x, uy = pt.y - pp.y;
    const float vx = tt.x - hh.x, vy = tt.y - hh.y;
    const float uu = fmaxf(__builtin_fmaf(ux, ux, uy * uy), 1e-12f);
    const float vv = fmaxf(__builtin_fmaf(vx, vx, vy * vy), 1e-12f);
    k.angle = 0.0f;
    if (with_angle) {
        const float turn = atan2f_evo(__builtin_fmaf(ux, vy, -(uy * vx)), __builtin_fmaf(ux, vx, uy * vy));
        const float mid = 0.5f * (k.lo + k.hi);
        k.angle = mid + wrapf(turn - mid);
    }
    const float ivv = 1.0f / vv, iuu = 1.0f / uu;
    k.gtx = -vy * ivv;
    k.gty = vx * ivv;
    const float gux = -uy * iuu, guy = ux * iuu;
    k.gfx = at_head ? -gux : gux;
    k.gfy = at_head ? -guy : guy;
    k.ghx = -(k.gtx + k.gfx);
    k.ghy = -(k.gty + k.gfy);
    return k;
}

__device__ __forceinline__ float joint_weight(const Joint& k) {
    const float nt = __builtin_fmaf(k.gtx, k.gtx, k.gty * k.gty);
    const float nf = __builtin_fmaf(k.gfx, k.gfx, k.gfy * k.gfy);
    const float nh = __builtin_fmaf(k.ghx, k.ghx, k.ghy * k.ghy);
    return __builtin_fmaf(k.wt, nt, __builtin_fmaf(k.wf, nf, k.wh * nh));
}

// Pushes the positions (VEL false) or velocities (VEL true) of the joint's
// three nodes, tip first, then far, then hub, each a read-modify-write so that
// a node named twice takes both pushes in order.
template <class C, bool VEL>
__device__ __forceinline__ void joint_push(const Cols<C>& s, const Joint& k, float lambda) {
    const float st = k.wt * lambda, sf = k.wf * lambda, sh = k.wh * lambda;
    {
        float4& q = s.p(k.tip);
        float4 v = q;
        if (VEL) { v.z = __builtin_fmaf(st, k.gtx, v.z); v.w = __builtin_fmaf(st, k.gty, v.w); }
        else { v.x = __builtin_fmaf(st, k.gtx, v.x); v.y = __builtin_fmaf(st, k.gty, v.y); }
        q = v;
    }
    {
        float4& q = s.p(k.far_);
        float4 v = q;
        if (VEL) { v.z = __builtin_fmaf(sf, k.gfx, v.z); v.w = __builtin_fmaf(sf, k.gfy, v.w); }
        else { v.x = __builtin_fmaf(sf, k.gfx, v.x); v.y = __builtin_fmaf(sf, k.gfy, v.y); }
        q = v;
    }
    {
        float4& q = s.p(k.hub);
        float4 v = q;
        if (VEL) { v.z = __builtin_fmaf(sh, k.ghx, v.z); v.w = __builtin_fmaf(sh, k.ghy, v.w); }
        else { v.x = __builtin_fmaf(sh, k.ghx, v.x); v.y = __builtin_fmaf(sh, k.ghy, v.y); }
        q = v;
    }
}

// shaders/creature.cu:442-451 — the step's demand of every muscle.
template <class C>
__device__ __forceinline__ void muscle_demands(Body<C>& b, const Cols<C>& s, const float* __restrict__ mus, float t,
                                               bool limp) {
    loop<C, C::MAXM, 0>(b.muscles, [&](auto kc_) {
        const int k = kc_;
        const float4* q = reinterpret_cast<const float4*>(mus + k * EVO_MUSCLE_WORDS);
        const float4 r2 = q[2], r3 = q[3];
        const float offset = s.o(k);
        const float before = rhythm(t, r2.y, r2.z, offset, r2.w, r3.x, r3.y);
        const float after = rhythm(t + C::dt, r2.y, r2.z, offset, r2.w, r3.x, r3.y);
        const float shortening = r2.x * (after - before) * C::rate;
        b.mdemand[k] = limp ? 0.0f : fmaxf(shortening * r3.z * 0.25f, 0.0f);
    });
}

// Adds (gx, gy) to node i's force pair.
template <class C>
__device__ __forceinline__ void add_force(const Cols<C>& s, uint32_t i, float gx, float gy, float share) {
    float2& q = s.f(i);
    float2 v = q;
    v.x = __builtin_fmaf(gx, share, v.x);
    v.y = __builtin_fmaf(gy, share, v.y);
    q = v;
}

// shaders/creature.cu:455-508 — the muscles' pull and their energy stores.
template <class C>
__device__ __forceinline__ void muscle_forces(Body<C>& b, const Cols<C>& s, float rec_h, float inv_me) {
    loop<C, C::MAXN, 0>(b.nodes, [&](auto ic_) {
        const int i = ic_;
        s.f(i) = make_float2(0.0f, 0.0f);
    });
    loop<C, C::MAXM, 0>(b.muscles, [&](auto kc_) {
        const int k = kc_;
        const uint32_t nodes = b.mnodes[k];
        const uint32_t a0 = nodes & 255u, a1 = (nodes >> 8u) & 255u, b0 = (nodes >> 16u) & 255u, b1 = nodes >> 24u;
        const float aa = b.maa[k], ab = b.mab[k];
        const float4 qa0 = s.p(a0), qa1 = s.p(a1), qb0 = s.p(b0), qb1 = s.p(b1);
        const float ax = __builtin_fmaf(qa1.x - qa0.x, aa, qa0.x);
        const float ay = __builtin_fmaf(qa1.y - qa0.y, aa, qa0.y);
        const float avx = __builtin_fmaf(qa1.z - qa0.z, aa, qa0.z);
        const float avy = __builtin_fmaf(qa1.w - qa0.w, aa, qa0.w);
        const float bx = __builtin_fmaf(qb1.x - qb0.x, ab, qb0.x);
        const float by = __builtin_fmaf(qb1.y - qb0.y, ab, qb0.y);
        const float bvx = __builtin_fmaf(qb1.z - qb0.z, ab, qb0.z);
        const float bvy = __builtin_fmaf(qb1.w - qb0.w, ab, qb0.w);
        const float dx = bx - ax, dy = by - ay;
        const float len = fmaxf(sqrtf(__builtin_fmaf(dx, dx, dy * dy)), 1e-6f);
        const float inv_len = 1.0f / len;
        const float ex = dx * inv_len, ey = dy * inv_len;
        const float lengthening = __builtin_fmaf(bvx - avx, ex, (bvy - avy) * ey);
        const float store = b.mstore[k];
        float drive = b.mdemand[k] * store;
        if (b.mhill[k] > 0.0f) { drive *= clampf(__builtin_fmaf(lengthening, b.mhill[k], 1.0f), 0.0f, 1.0f); }
        const float cap = b.mcap[k];
        const float active = clampf(__builtin_fmaf(lengthening, 0.15f, drive), -cap, cap);
        const float work = fminf(drive, cap) * fmaxf(-lengthening, 0.0f) * C::h;
        b.mstore[k] = clampf(__builtin_fmaf(rec_h, 1.0f - store, store - work * b.minvcap[k] * inv_me), 0.0f, 1.0f);
        const float pull = __builtin_fmaf(b.mtk[k], fmaxf(len - b.mslack[k], 0.0f), active);
        if constexpr (C::record) { b.mpull[k] = pull; }
        const float fx = ex * pull, fy = ey * pull;
        add_force<C>(s, a0, fx, fy, 1.0f - aa);
        add_force<C>(s, a1, fx, fy, aa);
        add_force<C>(s, b0, -fx, -fy, 1.0f - ab);
        add_force<C>(s, b1, -fx, -fy, ab);
    });
}

// shaders/creature.cu:511-546
template <class C>
__device__ __forceinline__ float2 node_force(const Body<C>& b, int I, const float4& q, const evo_params& p,
                                             uint32_t flags) {
    const float m = b.nm[I];
    float fx = 0.0f, fy = -p.gravity * m;
    if (has<C, EVO_WIND>(flags)) { fx = __builtin_fmaf(p.wind, m, fx); }
    if (has<C, EVO_MUD>(flags) || has<C, EVO_BRAMBLES>(flags)) {
        const float2 g = ground<C>(q.x, b.amp, b.qphase, p, flags);
        const float dry = (q.y - g.x) / sqrtf(__builtin_fmaf(g.y, g.y, 1.0f)) - b.nr[I];
        if (has<C, EVO_MUD>(flags)) { fx -= m * 2.0f * (clampf(-dry, 0.0f, p.mud) * (1.0f / 0.1f)) * q.z; }
        if (has<C, EVO_BRAMBLES>(flags)) {
            if (((b.feet >> I) & 1u) == 0u && dry < 0.01f) { fx -= m * p.brambles * q.z; }
        }
    }
    if (has<C, EVO_WATER>(flags)) {
        const float wet = clampf((p.water - (q.y - b.nr[I])) / (2.0f * b.nr[I]), 0.0f, 1.0f);
        fy = __builtin_fmaf(0.7f * m * p.gravity, wet, fy);
    }
    return make_float2(fx, fy);
}

// shaders/creature.cu:553-593
template <class C>
__device__ __forceinline__ void bone_drag(const Body<C>& b, const Cols<C>& s, const evo_params& p, uint32_t flags) {
    loop<C, C::MAXB, 0>(b.bones, [&](auto jc_) {
        const int j = jc_;
        const uint32_t i0 = b.bi0[j];
        float4& q0r = s.p(i0);
        float4& q1r = s.p(j + 1);
        float4 q0 = q0r, q1 = q1r;
        const float wx = 0.5f * (q0.z + q1.z), wy = 0.5f * (q0.w + q1.w);
        const float speed = sqrtf(__builtin_fmaf(wx, wx, wy * wy));
        const float limit = b.blim[j];
        const float k_air = fminf(b.bk[j] * speed, limit);
        float dfx = -wx * k_air, dfy = -wy * k_air;
        if (has<C, EVO_WATER>(flags)) {
            const float r0 = b.br0[j], r1 = b.nr[j + 1];
            const float width = r0 + r1;
            const float wet0 = clampf((p.water - (q0.y - r0)) / (2.0f * r0), 0.0f, 1.0f);
            const float wet1 = clampf((p.water - (q1.y - r1)) / (2.0f * r1), 0.0f, 1.0f);
            const float wet = 0.5f * (wet0 + wet1);
            const float ax = (q1.x - q0.x) / b.blen[j], ay = (q1.y - q0.y) / b.blen[j];
            const float along = __builtin_fmaf(wx, ax, wy * ay);
            const float lx = ax * along, ly = ay * along;
            const float sx = wx - lx, sy = wy - ly;
            const float k_water = fminf(100.0f * wet * b.blen[j] * width * speed, limit);
            dfx -= __builtin_fmaf(sx, k_water, lx * k_water * 0.25f);
            dfy -= __builtin_fmaf(sy, k_water, ly * k_water * 0.25f);
        }
        const float s0 = C::half_h * b.bw0[j], s1 = C::half_h * b.nw[j + 1];
        q0.z = __builtin_fmaf(s0, dfx, q0.z);
        q0.w = __builtin_fmaf(s0, dfy, q0.w);
        q0r = q0;
        q1.z = __builtin_fmaf(s1, dfx, q1.z);
        q1.w = __builtin_fmaf(s1, dfy, q1.w);
        q1r = q1;
    });
}

// shaders/creature.cu:597-613
template <class C>
__device__ __forceinline__ void solve_bones(const Body<C>& b, const Cols<C>& s) {
    loop<C, C::MAXB, 0>(b.bones, [&](auto jc_) {
        const int j = jc_;
        const uint32_t i0 = b.bi0[j];
        float4& q0r = s.p(i0);
        float4& q1r = s.p(j + 1);
        float4 q0 = q0r, q1 = q1r;
        const float dx = q1.x - q0.x, dy = q1.y - q0.y;
        const float len = fmaxf(sqrtf(__builtin_fmaf(dx, dx, dy * dy)), 1e-9f);
        const float w0 = b.bw0[j], w1 = b.nw[j + 1];
        const float sc = (len - b.blen[j]) / ((w0 + w1) * len);
        const float s0 = w0 * sc, s1 = w1 * sc;
        q0.x = __builtin_fmaf(s0, dx, q0.x);
        q0.y = __builtin_fmaf(s0, dy, q0.y);
        q1.x = __builtin_fmaf(-s1, dx, q1.x);
        q1.y = __builtin_fmaf(-s1, dy, q1.y);
        q0r = q0;
        q1r = q1;
    });
}

// shaders/creature.cu:618-634
template <class C>
__device__ __forceinline__ void solve_joints(const Body<C>& b, const Cols<C>& s) {
    loop<C, C::MAXB, 1>(b.bones, [&](auto jc_) {
        const int j = jc_;
        {
            if (b.jtopo[j] == EVO_NONE) { return; }
            const Joint k = joint<C>(b, s, j, true);
            float error = 0.0f;
            if (k.angle < k.lo) { error = k.angle - k.lo; }
            if (k.angle > k.hi) { error = k.angle - k.hi; }
            if (error == 0.0f) { return; }
            const float w = joint_weight(k);
            if (w > 0.0f) { joint_push<C, false>(s, k, -error / w); }
        }
    });
}

// shaders/creature.cu:683-702
template <class C>
__device__ __forceinline__ void damp_joints(const Body<C>& b, const Cols<C>& s) {
    loop<C, C::MAXB, 1>(b.bones, [&](auto jc_) {
        const int j = jc_;
        {
            if (b.jtopo[j] == EVO_NONE) { return; }
            const Joint k = joint<C>(b, s, j, false);
            const float4 vt = s.p(k.tip), vf = s.p(k.far_), vh = s.p(k.hub);
            float rate = k.gtx * vt.z;
            rate = __builtin_fmaf(k.gty, vt.w, rate);
            rate = __builtin_fmaf(k.gfx, vf.z, rate);
            rate = __builtin_fmaf(k.gfy, vf.w, rate);
            rate = __builtin_fmaf(k.ghx, vh.z, rate);
            rate = __builtin_fmaf(k.ghy, vh.w, rate);
            const float w = joint_weight(k);
            if (w > 0.0f) { joint_push<C, true>(s, k, (-C::damp_share * rate) / w); }
        }
    });
}

// shaders/creature.cu:746-759
template <class C>
__device__ __forceinline__ uint64_t broken_joints(const Body<C>& b, const Cols<C>& s) {
    uint64_t bits = 0u;
    loop<C, C::MAXB, 1>(b.bones, [&](auto jc_) {
        const int j = jc_;
        {
            if (b.jtopo[j] == EVO_NONE) { return; }
            const Joint k = joint<C>(b, s, j, true);
            if (k.angle < k.lo - 0.5f || k.angle > k.hi + 0.5f) { bits |= 1ull << j; }
        }
    });
    return bits;
}

// shaders/creature.cu:640-679 plus the velocity of :733-740, for node I.
template <class C>
__device__ __forceinline__ void ground_and_velocity(Body<C>& b, const Cols<C>& s, int I, const evo_params& p,
                                                    uint32_t flags) {
    float4& qr = s.p(I);
    float4 q = qr;
    const float2 prev = s.f(I);
    if (has<C, EVO_GROUND>(flags)) {
        const float2 g = ground<C>(q.x, b.amp, b.qphase, p, flags);
        const float secant = sqrtf(__builtin_fmaf(g.y, g.y, 1.0f));
        const float inv_sec = 1.0f / secant;
        const float nx = -g.y * inv_sec, ny = inv_sec;
        const float dry = __builtin_fmaf(q.y - g.x, inv_sec, -b.nr[I]);
        float depth, mu;
        if (has<C, EVO_MUD>(flags)) {
            depth = -(dry + p.mud);
            const float sink = clampf(-dry, 0.0f, p.mud) * (1.0f / 0.1f);
            mu = b.nf[I] * p.friction * __builtin_fmaf(2.0f, sink, 1.0f) * __builtin_fmaf(2.0f, sink, 1.0f);
        } else {
            depth = -dry;
            mu = b.nf[I] * p.friction;
        }
        if (depth > 0.0f) {
            if (has<C, EVO_ICE>(flags)) { mu *= __builtin_fmaf(-p.patches, ice_at(q.x), 1.0f); }
            q.x = __builtin_fmaf(nx, depth, q.x);
            q.y = __builtin_fmaf(ny, depth, q.y);
            const float tx = ny, ty = -nx;
            const float slide = __builtin_fmaf(q.x - prev.x, tx, (q.y - prev.y) * ty);
            const float budget = mu * depth;
            const float back = fabsf(slide) <= budget ? slide : copysignf(budget, slide);
            q.x = __builtin_fmaf(-tx, back, q.x);
            q.y = __builtin_fmaf(-ty, back, q.y);
            if constexpr (C::record) {
                b.normal_force[I] += b.nm[I] * depth * C::inv_h * C::inv_h;
                b.friction_force[I] += b.nm[I] * back * C::inv_h * C::inv_h;
            }
        }
    }
    q.z = (q.x - prev.x) * C::inv_h;
    q.w = (q.y - prev.y) * C::inv_h;
    qr = q;
}

template <class C>
__device__ __forceinline__ void integrate(const Body<C>& b, const Cols<C>& s, int I, const evo_params& p,
                                          uint32_t flags) {
    float4& qr = s.p(I);
    float4 q = qr;
    const float2 fm = s.f(I);
    const float2 f = node_force<C>(b, I, q, p, flags);
    const float wi = b.nw[I];
    q.z = __builtin_fmaf(C::h * (f.x + fm.x), wi, q.z);
    q.w = __builtin_fmaf(C::h * (f.y + fm.y), wi, q.w);
    if (has<C, EVO_AIR>(flags)) {
        q.z *= p.air_sub;
        q.w *= p.air_sub;
    }
    s.f(I) = make_float2(q.x, q.y);
    q.x = __builtin_fmaf(C::h, q.z, q.x);
    q.y = __builtin_fmaf(C::h, q.w, q.y);
    qr = q;
}

// shaders/creature.cu:705-742 — one substep.
template <class C>
__device__ __forceinline__ void substep(Body<C>& b, const Cols<C>& s, const evo_params& p, uint32_t flags,
                                       float rec_h) {
    muscle_forces<C>(b, s, rec_h, p.inv_muscle_energy);
    bone_drag<C>(b, s, p, flags);
    loop<C, C::MAXN, 0>(b.nodes, [&](auto ic_) { integrate<C>(b, s, ic_, p, flags); });
    solve_bones<C>(b, s);
    solve_joints<C>(b, s);
    loop<C, C::MAXN, 0>(b.nodes, [&](auto ic_) { ground_and_velocity<C>(b, s, ic_, p, flags); });
    damp_joints<C>(b, s);
}

// shaders/creature.cu:767-779 — frame t of a recording.
template <class C>
__device__ __forceinline__ void record_frame(const Body<C>& b, const Cols<C>& s, float2* __restrict__ frames,
                                             uint32_t t, uint32_t stride) {
    if constexpr (C::record) {
        float2* f = frames + static_cast<size_t>(t) * stride;
        loop<C, C::MAXN, 0>(b.nodes, [&](auto ic_) {
            const int i = ic_;
            const float4 q = s.p(i);
            f[i] = make_float2(q.x, q.y);
            f[EVO_MAX_NODES + b.muscles + i] = make_float2(b.normal_force[i], b.friction_force[i]);
        });
        loop<C, C::MAXM, 0>(b.muscles, [&](auto kc_) {
            const int k = kc_;
            f[EVO_MAX_NODES + k] = make_float2(b.mstore[k], b.mpull[k]);
        });
        const uint64_t broken = broken_joints<C>(b, s);
        f[stride - 1u] = make_float2(bits_f(static_cast<uint32_t>(broken)), bits_f(static_cast<uint32_t>(broken >> 32)));
    }
}

// shaders/creature.cu:845-950 — the lane's new creature.
template <class C>
__device__ __forceinline__ void begin_creature(Body<C>& b, Trial& tr, const Cols<C>& s, uint32_t cidx,
                                               const uint32_t* __restrict__ records,
                                               const float* __restrict__ muscles, const uint4* __restrict__ heads,
                                               const evo_params& p, uint32_t flags, float2* __restrict__ frames) {
    tr.cidx = cidx;
    const uint4 h0 = heads[2u * cidx];
    const uint4 h1 = heads[2u * cidx + 1u];
    b.nodes = h0.x & 255u;
    b.bones = b.nodes - 1u;
    b.muscles = h0.y;
    const uint32_t* rec = records + h0.w;
    tr.mus_off = h1.x;
    b.inv_mass = bits_f(h1.z);
    tr.flags = h1.w >> 16u;
    tr.period_f = h2f(h1.w & 0xffffu);
    if (has<C, EVO_QUAKE>(flags)) {
        b.qphase = static_cast<float>(h0.z & 0xffffu) * (1.0f / 65536.0f);
        b.amp = __builtin_fmaf(
            p.quake, __builtin_fmaf(static_cast<float>((h0.z >> 16u) & 0xffffu), 0.8f / 65536.0f, 0.6f), p.terrain);
    } else {
        b.qphase = 0.0f;
        b.amp = p.terrain;
    }
    b.feet = 0u;
    loop<C, C::MAXN, 0>(b.nodes, [&](auto ic_) {
        const int i = ic_;
        const uint32_t* r = rec + i * EVO_NODE_WORDS;
        b.nm[i] = bits_f(r[0]);
        b.nw[i] = 1.0f / b.nm[i];
        b.nr[i] = bits_f(r[1]);
        b.nf[i] = bits_f(r[2]);
        s.p(i) = make_float4(bits_f(r[3]), bits_f(r[4]), 0.0f, 0.0f);
        b.feet |= (r[5] & 1u) << i;
        if constexpr (C::record) {
            b.normal_force[i] = 0.0f;
            b.friction_force[i] = 0.0f;
        }
    });
    const uint32_t* bone_rec = rec + b.nodes * EVO_NODE_WORDS;
    loop<C, C::MAXB, 0>(b.bones, [&](auto jc_) {
        const int j = jc_;
        const uint32_t* r = bone_rec + j * EVO_BONE_WORDS;
        const uint32_t i0 = r[0];
        // Constants of the pivot node come from its record: the registers
        // of the node arrays cannot be indexed by a per-lane number.
        const float m0 = bits_f(rec[i0 * EVO_NODE_WORDS]);
        const float r0 = bits_f(rec[i0 * EVO_NODE_WORDS + 1]);
        const float length = bits_f(r[1]);
        const float width = r0 + b.nr[j + 1];
        const float bone_mass = m0 + b.nm[j + 1];
        b.bi0[j] = i0;
        b.blen[j] = length;
        b.bw0[j] = 1.0f / m0;
        b.bk[j] = 0.6f * length * width;
        b.blim[j] = 0.5f * bone_mass * C::inv_h;
        b.br0[j] = r0;
        b.jtopo[j] = EVO_NONE;
        const uint32_t pj = r[2];
        if (j > 0 && pj != EVO_NONE) {
            const uint32_t hub = i0, parent_pivot = bone_rec[pj * EVO_BONE_WORDS];
            const bool at_head = hub == parent_pivot;
            const uint32_t far_ = at_head ? pj + 1u : parent_pivot;
            b.jtopo[j] = hub | ((pj + 1u) << 5u) | (parent_pivot << 10u) | (far_ << 15u) | ((at_head ? 1u : 0u) << 20u);
            b.jlo[j] = bits_f(r[3]);
            b.jhi[j] = bits_f(r[4]);
            b.jwf[j] = 1.0f / bits_f(rec[far_ * EVO_NODE_WORDS]);
        } else {
            b.jlo[j] = 0.0f;
            b.jhi[j] = 0.0f;
            b.jwf[j] = 0.0f;
        }
    });
    loop<C, C::MAXM, 0>(b.muscles, [&](auto kc_) {
        const int k = kc_;
        const float4* q = reinterpret_cast<const float4*>(muscles + tr.mus_off + k * EVO_MUSCLE_WORDS);
        const float4 r0 = q[0], r1 = q[1];
        b.mnodes[k] = f_bits(r0.x);
        b.maa[k] = r0.y;
        b.mab[k] = r0.z;
        b.mhill[k] = r0.w;
        b.mcap[k] = r1.x;
        b.minvcap[k] = r1.y;
        b.mtk[k] = r1.z;
        b.mslack[k] = r1.w;
        b.mstore[k] = 1.0f;
        b.mdemand[k] = 0.0f;
        if constexpr (C::record) { b.mpull[k] = 0.0f; }
        s.o(k) = 0.0f;
    });
    tr.fitness = 0.0f;
    tr.ground_contact = 0.0f;
    tr.vmin = 1e20f;
    tr.vmax = -1e20f;
    tr.previous_center_y = 0.0f;
    tr.vertical_extremum = 0.0f;
    tr.vertical_trend = 0.0f;
    tr.gait_turns = 0.0f;
    tr.height_sum = 0.0f;
    tr.fall_time = 0.0f;
    tr.head_shake_out = 0.0f;
    tr.screen_x = 0.0f;
    tr.screened = 0.0f;
    tr.contact_lo = tr.contact_hi = tr.lift_lo = tr.lift_hi = tr.ground_lo = tr.ground_hi = 0u;
    tr.contact_bits = tr.lift_bits = tr.ground_bits = 0u;
    tr.head_shake = 0.0f;
    const float4 head = s.p(0);
    tr.head_at_x = head.x;
    tr.head_at_y = head.y;
    tr.head_vel_x = 0.0f;
    tr.head_vel_y = 0.0f;
    tr.rung_speed = tr.rung_early = tr.rung_late = tr.rung_bits = 0u;
    tr.rung_x = 0.0f;
    tr.limp = false;
    tr.scoring = true;
    tr.step = 0u;
    if constexpr (C::record) {
        for (uint32_t t = 0u; t <= C::settle; t++) { record_frame<C>(b, s, frames, t, p.stride); }
    }
}

__device__ __forceinline__ evo_result pack_result(const Trial& tr) {
    evo_result o;
    o.fitness = tr.fitness;
    o.ground_contact = tr.ground_contact;
    o.vertical_oscillation = tr.vmin;
    o.gait_frequency = tr.vmax;
    o.previous_center_y = tr.previous_center_y;
    o.vertical_extremum = tr.vertical_extremum;
    o.vertical_trend = tr.vertical_trend;
    o.gait_turns = tr.gait_turns;
    o.height_sum = tr.height_sum;
    o.contact_lo = bits_f(tr.contact_lo);
    o.contact_hi = bits_f(tr.contact_hi);
    o.lift_lo = bits_f(tr.lift_lo);
    o.lift_hi = bits_f(tr.lift_hi);
    o.ground_lo = bits_f(tr.ground_lo);
    o.ground_hi = bits_f(tr.ground_hi);
    o.fall_time = tr.fall_time;
    o.head_shake = tr.head_shake_out;
    o.screen_x = tr.screen_x;
    o.screened = tr.screened;
    return o;
}

// Stores a result with plain dword stores.
__device__ __forceinline__ void store_result(evo_result* __restrict__ out, const evo_result& r) {
    float* o = reinterpret_cast<float*>(out);
    const float* v = reinterpret_cast<const float*>(&r);
#pragma unroll
    for (int i = 0; i < 19; i++) { o[i] = v[i]; }
}

template <class C>
__device__ __forceinline__ void step_metrics(const Body<C>& b, const Cols<C>& s, int I, const evo_params& p,
                                             uint32_t flags, bool& failed, float& center_y, float& com_x, float& low,
                                             float& high, uint32_t& touching, uint32_t& lifted) {
    const float4 q = s.p(I);
    failed = failed || !(fabsf(q.x) <= 1e6f && fabsf(q.y) <= 1e6f);
    center_y += q.y;
    com_x = __builtin_fmaf(b.nm[I], q.x, com_x);
    low = fminf(low, q.y - b.nr[I]);
    high = fmaxf(high, q.y + b.nr[I]);
    if (has<C, EVO_GROUND>(flags)) {
        const float2 g = ground<C>(q.x, b.amp, b.qphase, p, flags);
        const float floor_y = __builtin_fmaf(b.nr[I], sqrtf(__builtin_fmaf(g.y, g.y, 1.0f)), g.x);
        if (q.y <= floor_y + 0.002f) { touching |= 1u << I; }
        if (q.y > floor_y + 0.01f) { lifted |= 1u << I; }
    }
}

// shaders/creature.cu:956-1248 — one step of the lane's creature.
template <class C>
__device__ __forceinline__ void run_step(Body<C>& b, Trial& tr, const Cols<C>& s, const float* __restrict__ muscles,
                                         evo_result* __restrict__ results, const evo_params& p, uint32_t flags,
                                         float rec_h, float2* __restrict__ frames) {
    const uint32_t step = tr.step;
    const float* mus = muscles + tr.mus_off;
    const uint32_t cflags = tr.flags;
    const float t_now = static_cast<float>(step) * C::dt;
    if constexpr (C::record) {
        loop<C, C::MAXN, 0>(static_cast<uint32_t>(C::MAXN), [&](auto ic_) {
            const int i = ic_;
            b.normal_force[i] = 0.0f;
            b.friction_force[i] = 0.0f;
        });
    }
    muscle_demands<C>(b, s, mus, t_now, tr.limp);
    for (uint32_t sub = 0u; sub < C::substeps; sub++) { substep<C>(b, s, p, flags, rec_h); }
    if constexpr (C::record) {
        const float inv_sub = 1.0f / static_cast<float>(C::substeps);
        loop<C, C::MAXN, 0>(static_cast<uint32_t>(C::MAXN), [&](auto ic_) {
            const int i = ic_;
            b.normal_force[i] *= inv_sub;
            b.friction_force[i] *= inv_sub;
        });
    }
    bool failed = false;
    float center_y = 0.0f, com_x = 0.0f, low = 1e20f, high = -1e20f;
    uint32_t touching = 0u, lifted = 0u;
    loop<C, C::MAXN, 0>(b.nodes, [&](auto ic_) {
        step_metrics<C>(b, s, ic_, p, flags, failed, center_y, com_x, low, high, touching, lifted);
    });
    center_y /= static_cast<float>(b.nodes);
    com_x *= b.inv_mass;
    tr.contact_bits |= touching;
    tr.lift_bits |= tr.contact_bits & lifted;
    const uint32_t down = touching & ~tr.ground_bits;
    tr.ground_bits = touching;
    if (down != 0u && step > 0u) {
        const float next = t_now + C::dt;
        loop<C, C::MAXM, 0>(b.muscles, [&](auto kc_) {
            const int k = kc_;
            const float4* q = reinterpret_cast<const float4*>(mus + k * EVO_MUSCLE_WORDS);
            const uint32_t sensor = f_bits(q[4].x);
            if (sensor != EVO_NONE && sensor < 32u && ((down >> sensor) & 1u)) {
                const float4 r2 = q[2], r3 = q[3];
                s.o(k) = fracf(r3.w - __builtin_fmaf(next, r2.y, r2.z));
            }
        });
    }
    if ((step + 1u) % C::head_sample == 0u) {
        const float4 head = s.p(0);
        const float vx = (head.x - tr.head_at_x) * 60.0f, vy = (head.y - tr.head_at_y) * 60.0f;
        const float ax = (vx - tr.head_vel_x) * 60.0f, ay = (vy - tr.head_vel_y) * 60.0f;
        if (t_now >= 0.1f) {
            const float accel = sqrtf(__builtin_fmaf(ax, ax, ay * ay));
            tr.head_shake =
                __builtin_fmaf(accel - tr.head_shake, fminf(1.0f / (0.1f * 60.0f), 1.0f), tr.head_shake);
        }
        tr.head_at_x = head.x;
        tr.head_at_y = head.y;
        tr.head_vel_x = vx;
        tr.head_vel_y = vy;
    }
    const bool broken = broken_joints<C>(b, s) != 0u;
    if (tr.scoring) {
        tr.contact_lo = tr.contact_bits;
        tr.lift_lo = tr.lift_bits;
        tr.ground_lo = tr.ground_bits;
        tr.head_shake_out = tr.head_shake;
        const float y0 = s.p(0).y, y1 = s.p(1).y;
        const bool fell = y0 < y1 || broken || tr.head_shake > 8.0f * 9.8f || failed;
        bool ended = false;
        if (fell) {
            tr.fall_time = t_now + C::dt;
            tr.fitness = failed ? -1e20f : com_x;
            if (step <= p.screen_step) { tr.screen_x = tr.fitness; }
            ended = true;
        }
        tr.ground_contact += static_cast<float>(__builtin_popcount(touching));
        tr.height_sum += high - low;
        tr.vmin = fminf(tr.vmin, center_y);
        tr.vmax = fmaxf(tr.vmax, center_y);
        if (step == 0u) {
            tr.previous_center_y = center_y;
            tr.vertical_extremum = center_y;
            tr.vertical_trend = 0.0f;
            tr.gait_turns = 0.0f;
        } else if (step % C::sample == 0u) {
            const float delta = center_y - tr.previous_center_y;
            if (tr.vertical_trend == 0.0f) {
                if (fabsf(delta) > 0.0005f) {
                    tr.vertical_trend = delta > 0.0f ? 1.0f : -1.0f;
                    tr.vertical_extremum = center_y;
                }
            } else if (tr.vertical_trend > 0.0f) {
                if (center_y > tr.vertical_extremum) {
                    tr.vertical_extremum = center_y;
                } else if (tr.vertical_extremum - center_y > 0.005f) {
                    tr.gait_turns += 1.0f;
                    tr.vertical_trend = -1.0f;
                    tr.vertical_extremum = center_y;
                }
            } else {
                if (center_y < tr.vertical_extremum) {
                    tr.vertical_extremum = center_y;
                } else if (center_y - tr.vertical_extremum > 0.005f) {
                    tr.gait_turns += 1.0f;
                    tr.vertical_trend = 1.0f;
                    tr.vertical_extremum = center_y;
                }
            }
            tr.previous_center_y = center_y;
        }
        if (step + C::half_s == C::rung1 || step + C::half_s == C::rung2) { tr.rung_x = com_x; }
        if (step == C::rung1 || step == C::rung2) {
            const bool second = step == C::rung2;
            const uint32_t sh = second ? 16u : 0u;
            const uint32_t keep = second ? 0x0000ffffu : 0xffff0000u;
            const float speed = (com_x - tr.rung_x) * (C::rate / static_cast<float>(C::half_s));
            const float touched = static_cast<float>(__builtin_popcount(tr.contact_bits)) / static_cast<float>(b.nodes);
            float en_mean = 0.0f;
            loop<C, C::MAXM, 0>(b.muscles, [&](auto kc_) {
                const int k = kc_;
                en_mean += b.mstore[k];
            });
            en_mean /= static_cast<float>(b.muscles > 1u ? b.muscles : 1u);
            tr.contact_hi = (tr.contact_hi & keep) | (f2h(com_x) << sh);
            tr.ground_hi = (tr.ground_hi & keep) | (f2h(tr.head_shake) << sh);
            tr.rung_speed = (tr.rung_speed & keep) | (f2h(speed) << sh);
            const uint32_t pair = f2h(touched) | (f2h(en_mean) << 16u);
            if (second) { tr.rung_late = pair; } else { tr.rung_early = pair; }
            const uint32_t r = second ? 1u : 0u;
            const float gait_now = tr.gait_turns * (0.5f / (t_now + C::dt));
            const float bandf = gait_now * (8.0f / 6.0f);
            uint32_t band = bandf >= 7.0f ? 7u : (bandf > 0.0f ? static_cast<uint32_t>(bandf) : 0u);
            if (bandf != bandf) { band = 0u; }
            tr.rung_bits |= band << (8u + 3u * r);
            if (!ended && (cflags & (second ? 0xbu : 0x7u)) == 0u) {
                const evo_rung& rp = second ? p.r2 : p.r1;
                const float fv[6] = {h2f(f2h(com_x)), h2f(f2h(speed)), h2f(f2h(touched)),
                                     h2f(f2h(tr.head_shake)), h2f(f2h(en_mean)), tr.period_f};
                float score = 0.0f;
                bool finite = true;
                unroll<6>([&](auto qc_) {
                    constexpr int q = decltype(qc_)::value;
                    score = __builtin_fmaf(rp.w[q], fv[q], score);
                    finite = finite && __builtin_isfinite(fv[q]);
                });
                if (finite && score < rp.bias && ((rp.off >> band) & 1u) == 0u) {
                    tr.screened = t_now + C::dt;
                    tr.screen_x = com_x;
                    tr.fitness = com_x;
                    tr.rung_bits |= (r + 1u) << 6u;
                    ended = true;
                }
            }
        }
        if (step == C::rung3 || step == C::rung4) {
            const bool second = step == C::rung4;
            const uint32_t keep = second ? 0x0000ffffu : 0xffff0000u;
            tr.lift_hi = (tr.lift_hi & keep) | (f2h(com_x) << (second ? 16u : 0u));
        }
        if (step == p.screen_step && !ended) {
            tr.screen_x = com_x;
            const float bar = (cflags & 32u) != 0u ? p.screen_bar_reshaped
                            : (cflags & 16u) != 0u ? p.screen_bar_young : p.screen_bar;
            if (com_x < bar && (cflags & 1u) == 0u) {
                tr.screened = t_now + C::dt;
                tr.fitness = com_x;
                ended = true;
            }
        }
        bool last = false;
        if (ended) {
            tr.vmin = fmaxf(tr.vmax - tr.vmin, 0.0f);
            tr.vmax = tr.gait_turns * 0.5f / (t_now + C::dt);
        } else if (step + 1u == p.steps) {
            tr.fitness = com_x;
            tr.vmin = fmaxf(tr.vmax - tr.vmin, 0.0f);
            tr.vmax = tr.gait_turns * 0.5f / fmaxf(static_cast<float>(step + 1u) * C::dt, C::dt);
            last = true;
        }
        if (ended || last) {
            const uint32_t fin = f2h(tr.fitness);
            uint32_t d = tr.contact_hi;
            if (step < C::rung1) { d = (d & 0xffff0000u) | fin; }
            if (step < C::rung2) { d = (d & 0x0000ffffu) | (fin << 16u); }
            tr.contact_hi = d;
            d = tr.lift_hi;
            if (step < C::rung3) { d = (d & 0xffff0000u) | fin; }
            if (step < C::rung4) { d = (d & 0x0000ffffu) | (fin << 16u); }
            tr.lift_hi = d;
            const uint32_t code = (fell ? 16u : 0u) | (failed ? 32u : 0u) | (tr.screened > 0.0f ? 3u : 0u) |
                                  (tr.rung_bits & 0x3fc0u) | ((cflags & 1u) << 14u);
            evo_result res = pack_result(tr);
            res.previous_center_y = bits_f(tr.rung_speed);
            res.vertical_extremum = bits_f(tr.rung_early);
            res.vertical_trend = bits_f(tr.rung_late);
            res.gait_turns = bits_f(tr.ground_hi);
            res.ground_hi = bits_f(code | ((step + 1u < 65535u ? step + 1u : 65535u) << 16u));
            if constexpr (C::record) {
                tr.kept = res;
                tr.scoring = false;
                tr.limp = true;
            } else {
                store_result(results + tr.cidx, res);
                tr.live = false;
                return;
            }
        }
    }
    if constexpr (C::record) {
        record_frame<C>(b, s, frames, C::settle + step + 1u, p.stride);
        if (step + 1u >= p.steps) {
            store_result(results + tr.cidx, tr.kept);
            tr.live = false;
            return;
        }
    }
    tr.step = step + 1u;
}

// Most creatures a wave claims from the launch's counter at a time, and the
// divisor that shrinks claims near the end (shaders/creature.cu:828-836).
constexpr uint32_t kChunk = 64u;
constexpr uint32_t kChunkK = 4u;

// shaders/creature.cu:1266-1394 — the persistent launch. Each wave claims
// runs of creatures from `counter`, hands them to its free lanes in lane
// order, and steps every live lane until no creature is left.
template <class C>
__global__ void __launch_bounds__(C::BLOCK) advance_kernel(const uint32_t* __restrict__ records,
                                                          const float* __restrict__ muscles,
                                                          const uint4* __restrict__ heads,
                                                          evo_result* __restrict__ results,
                                                          uint32_t* __restrict__ counter, const evo_params p,
                                                          uint32_t flags, float2* __restrict__ frames,
                                                          uint64_t frames_per_creature) {
    __shared__ float4 s_pv[C::MAXN * C::BLOCK];
    __shared__ float2 s_fp[C::MAXN * C::BLOCK];
    __shared__ float s_off[(C::MAXM > 0 ? C::MAXM : 1) * C::BLOCK];
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    Cols<C> s;
    s.pv = s_pv + threadIdx.x;
    s.fp = s_fp + threadIdx.x;
    s.off = s_off + threadIdx.x;
    const float rec_h = 0.5f * p.muscle_recovery * C::h;
    const uint32_t waves = gridDim.x * (C::BLOCK / 64u);
    Body<C> b;
    Trial tr;
    tr.live = false;
    uint32_t win_cur = 0u, win_end = 0u;
    bool claimed_all = false;
    float2* my_frames = nullptr;
    for (;;) {
        uint64_t need = ~__ballot(tr.live);
        uint32_t mine_at = 0u;
        bool fresh = false;
        while (need != 0ull) {
            if (win_cur < win_end) {
                const uint32_t rank = static_cast<uint32_t>(__popcll(need & below));
                const uint32_t take = min(win_end - win_cur, static_cast<uint32_t>(__popcll(need)));
                const bool mine = ((need >> lane) & 1ull) != 0ull && rank < take;
                if (mine) {
                    mine_at = win_cur + rank;
                    fresh = true;
                }
                need &= ~__ballot(mine);
                win_cur += take;
                continue;
            }
            if (claimed_all) { break; }
            uint32_t c0 = 0u, want = kChunk;
            if (lane == 0u) {
                const uint32_t seen = __atomic_load_n(counter, __ATOMIC_RELAXED);
                const uint32_t left = seen < p.count ? p.count - seen : 0u;
                want = min(max(left / (kChunkK * waves), 1u), kChunk);
                c0 = atomicAdd(counter, want);
            }
            c0 = __shfl(c0, 0);
            want = __shfl(want, 0);
            if (c0 >= p.count) {
                claimed_all = true;
                break;
            }
            win_cur = c0;
            win_end = min(c0 + want, p.count);
        }
        if (fresh) {
            tr.live = true;
            if constexpr (C::record) {
                my_frames = frames + static_cast<uint64_t>(mine_at) * frames_per_creature * p.stride;
            }
            begin_creature<C>(b, tr, s, p.base + mine_at, records, muscles, heads, p, flags, my_frames);
        }
        if (__ballot(tr.live) == 0ull) { break; }
        if (tr.live) { run_step<C>(b, tr, s, muscles, results, p, flags, rec_h, my_frames); }
    }
}

// Clears n words with vector stores (a fill kernel instead of a memset node,
// so a captured graph replays it faithfully).
template <int = 0>
__global__ void clear_words(uint32_t* __restrict__ words, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { words[i] = 0u; }
}

}  // namespace dev
}  // namespace evo
