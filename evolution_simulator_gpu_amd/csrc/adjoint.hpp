// Reverse-mode derivatives of a creature's trial: d(fitness)/d(parameters).
//
// The reference has no adjoint; this is the differentiation of its trial
// physics (shaders/creature.cu:442-742): muscle demands, muscle forces,
// bone drag, integration, bone and joint projections, ground contact with
// friction, and joint damping. The forward expressions are those of the
// simulator kernel, operation for operation, so a recomputed trajectory is
// the simulated trajectory bit for bit; the reverse functions are the
// transposed Jacobians of each stage, derived by hand, evaluated in FP32.
//
// Discrete decisions (contact or not, a joint outside its range or not, the
// clamps of the Hill factor, force cap and energy store, the Coulomb
// friction cone) are taken as in the forward pass, so the result is the
// derivative of the piecewise-smooth trajectory the forward pass followed.
//
// Parameters, in the order of the gradient vector: per muscle k six values
// at 6k .. 6k+5 (rhythm amplitude, inverse period, phase, Hill factor, force
// cap, inverse energy capacity: muscle record words 8, 9, 10, 3, 4, 5), then
// per node i its friction (node record word 2) at 6M + i.
//
// Everything here is host-and-device code, so the same functions run in the
// HIP kernel and in a CPU build used for checking.
#pragma once

#include <cmath>
#include <cstdint>

#include "../../include/evo_types.h"
#include "devmath.hpp"

namespace evo {
namespace adj {

constexpr uint32_t kNone = 0xffffffffu;

// The world of an adjoint launch: the effects the reverse pass handles are
// ground, terrain and quake bumps, slope, gaps, hurdles, ice, wind and air.
struct World {
    uint32_t flags;
    float h, inv_h, dt, rate;
    float gravity, friction, terrain, quake, slope, wind, gaps, hurdles, patches, air_sub;
    float rec_h, inv_energy, damp_share;
};

constexpr uint32_t kAdjointFlags = EVO_GROUND | EVO_TERRAIN | EVO_SLOPE | EVO_GAPS | EVO_HURDLES | EVO_QUAKE |
                                   EVO_ICE | EVO_WIND | EVO_AIR;

EVO_HD World make_world(const evo_params& p, uint32_t flags, uint32_t rate, uint32_t substeps) {
    World w;
    w.flags = flags;
    w.rate = static_cast<float>(rate);
    w.dt = 1.0f / w.rate;
    w.inv_h = w.rate * static_cast<float>(substeps);
    w.h = 1.0f / w.inv_h;
    w.gravity = p.gravity;
    w.friction = p.friction;
    w.terrain = p.terrain;
    w.quake = p.quake;
    w.slope = p.slope;
    w.wind = p.wind;
    w.gaps = p.gaps;
    w.hurdles = p.hurdles;
    w.patches = p.patches;
    w.air_sub = p.air_sub;
    w.rec_h = 0.5f * p.muscle_recovery * w.h;
    w.inv_energy = p.inv_muscle_energy;
    w.damp_share = fminf(w.h * 10.0f, 1.0f);
    return w;
}

EVO_HD float fracf_(float x) { return x - floorf(x); }
EVO_HD float clampf_(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
EVO_HD float wrapf_(float a) {
    const float tau = 6.28318530718f;
    return __builtin_fmaf(-tau, floorf((a + 3.14159265359f) * (1.0f / tau)), a);
}

template <int MN, int MM>
struct Body {
    uint32_t nodes, bones, muscles;
    float inv_mass, amp, qphase;
    float nm[MN], nw[MN], nr[MN], nf[MN];
    uint32_t bi0[MN];
    float blen[MN], bk[MN], blim[MN], bs0[MN], bs1[MN];
    uint32_t jtopo[MN];  // hub | ptip << 6 | ppiv << 12 | far << 18 | at_head << 24, or kNone
    float jlo[MN], jhi[MN];
    uint32_t mnodes[MM];
    float maa[MM], mab[MM], mhill[MM], mcap[MM], minvcap[MM], mtk[MM], mslack[MM];
    float mamp[MM], minvp[MM], mphase[MM], mduty[MM], midu[MM], micomp[MM], mstiff[MM], moff[MM];
    // state at the start of a substep, and the step's demands
    float px[MN], py[MN], vx[MN], vy[MN], store[MM], demand[MM];
};

// The derivatives carried backwards, and the parameter gradient.
template <int MN, int MM>
struct Adj {
    float px[MN], py[MN], vx[MN], vy[MN], store[MM], demand[MM];
    float g_amp[MM], g_invp[MM], g_phase[MM], g_hill[MM], g_cap[MM], g_invcap[MM], g_fric[MN];
};

// What one substep's reverse needs beyond its start state.
template <int MN>
struct Tape {
    float drag_v[MN][4];   // bone j: velocities of its two nodes before its drag
    float bone_x[MN][4];   // bone j: positions of its two nodes before its projection
    float joint_x[MN][6];  // joint j: positions of tip, far, hub before its push
    float ground_x[MN][2]; // node i: position before the ground
    float damp_v[MN][6];   // joint j: velocities of tip, far, hub before its damping
    float fx[MN], fy[MN];  // the force table (for checks)
    float end_x[MN], end_y[MN];
};

// ---------------------------------------------------------------- set-up

// Loads creature cidx (shaders/creature.cu:845-950). Returns false for a
// body larger than MN nodes or MM muscles.
template <int MN, int MM>
EVO_HD bool load(Body<MN, MM>& b, const World& w, uint32_t cidx, const uint32_t* records, const float* muscles,
                 const uint32_t* heads) {
    const uint32_t* h0 = heads + 8u * static_cast<uint64_t>(cidx);
    const uint32_t* h1 = h0 + 4;
    b.nodes = h0[0] & 255u;
    b.muscles = h0[1];
    if (b.nodes < 2u || b.nodes > static_cast<uint32_t>(MN) || b.muscles > static_cast<uint32_t>(MM)) {
        return false;
    }
    b.bones = b.nodes - 1u;
    b.inv_mass = bits_f(h1[2]);
    if ((w.flags & EVO_QUAKE) != 0u) {
        b.qphase = static_cast<float>(h0[2] & 0xffffu) * (1.0f / 65536.0f);
        b.amp = __builtin_fmaf(w.quake,
                               __builtin_fmaf(static_cast<float>((h0[2] >> 16u) & 0xffffu), 0.8f / 65536.0f, 0.6f),
                               w.terrain);
    } else {
        b.qphase = 0.0f;
        b.amp = w.terrain;
    }
    const uint32_t* rec = records + h0[3];
    for (uint32_t i = 0; i < b.nodes; i++) {
        const uint32_t* r = rec + i * EVO_NODE_WORDS;
        b.nm[i] = bits_f(r[0]);
        b.nw[i] = 1.0f / b.nm[i];
        b.nr[i] = bits_f(r[1]);
        b.nf[i] = bits_f(r[2]);
        b.px[i] = bits_f(r[3]);
        b.py[i] = bits_f(r[4]);
        b.vx[i] = 0.0f;
        b.vy[i] = 0.0f;
    }
    const uint32_t* bone_rec = rec + b.nodes * EVO_NODE_WORDS;
    for (uint32_t j = 0; j < b.bones; j++) {
        const uint32_t* r = bone_rec + j * EVO_BONE_WORDS;
        const uint32_t i0 = r[0], i1 = j + 1u;
        const float length = bits_f(r[1]);
        const float width = b.nr[i0] + b.nr[i1];
        const float bone_mass = b.nm[i0] + b.nm[i1];
        b.bi0[j] = i0;
        b.blen[j] = length;
        b.bk[j] = 0.6f * length * width;
        b.blim[j] = 0.5f * bone_mass * w.inv_h;
        b.bs0[j] = 0.5f * w.h * b.nw[i0];
        b.bs1[j] = 0.5f * w.h * b.nw[i1];
        b.jtopo[j] = kNone;
        const uint32_t pj = r[2];
        if (j > 0u && pj != EVO_NONE) {
            const uint32_t hub = i0, parent_pivot = bone_rec[pj * EVO_BONE_WORDS];
            const bool at_head = hub == parent_pivot;
            const uint32_t far_ = at_head ? pj + 1u : parent_pivot;
            b.jtopo[j] = hub | ((pj + 1u) << 6u) | (parent_pivot << 12u) | (far_ << 18u) | ((at_head ? 1u : 0u) << 24u);
            b.jlo[j] = bits_f(r[3]);
            b.jhi[j] = bits_f(r[4]);
        }
    }
    const float* mus = muscles + h1[0];
    for (uint32_t k = 0; k < b.muscles; k++) {
        const float* q = mus + k * EVO_MUSCLE_WORDS;
        b.mnodes[k] = f_bits(q[0]);
        b.maa[k] = q[1];
        b.mab[k] = q[2];
        b.mhill[k] = q[3];
        b.mcap[k] = q[4];
        b.minvcap[k] = q[5];
        b.mtk[k] = q[6];
        b.mslack[k] = q[7];
        b.mamp[k] = q[8];
        b.minvp[k] = q[9];
        b.mphase[k] = q[10];
        b.mduty[k] = q[11];
        b.midu[k] = q[12];
        b.micomp[k] = q[13];
        b.mstiff[k] = q[14];
        b.moff[k] = 0.0f;
        b.store[k] = 1.0f;
        b.demand[k] = 0.0f;
    }
    return true;
}

// ---------------------------------------------------------------- ground

struct Ground {
    float h, s, ds;  // height, slope (its derivative) and the slope's derivative
};

// shaders/creature.cu:166-226, with the derivative of the slope.
EVO_HD Ground ground(const World& w, float x, float amp, float qphase) {
    float height = 0.0f, slope = 0.0f, ds = 0.0f;
    if ((w.flags & EVO_TERRAIN) != 0u) {
        const float t0 = __builtin_fmaf(x, 1.0f / 1.1f, qphase);
        const float u0 = fracf_(t0);
        const float w0 = u0 * (1.0f - u0);
        const float t1 = __builtin_fmaf(x, 1.0f / 0.43f, 0.3f) + qphase;
        const float u1 = fracf_(t1);
        const float w1 = u1 * (1.0f - u1);
        height = amp * __builtin_fmaf(0.65f * 16.0f * w0, w0, 0.35f * 16.0f * w1 * w1);
        slope = amp * __builtin_fmaf(0.65f * 32.0f * w0 * (1.0f - 2.0f * u0), 1.0f / 1.1f,
                                     0.35f * 32.0f * w1 * (1.0f - 2.0f * u1) * (1.0f / 0.43f));
        const float c0 = (1.0f - 2.0f * u0) * (1.0f - 2.0f * u0) - 2.0f * w0;
        const float c1 = (1.0f - 2.0f * u1) * (1.0f - 2.0f * u1) - 2.0f * w1;
        ds = amp * (0.65f * 32.0f * c0 * (1.0f / (1.1f * 1.1f)) + 0.35f * 32.0f * c1 * (1.0f / (0.43f * 0.43f)));
    }
    if ((w.flags & EVO_SLOPE) != 0u) {
        height = __builtin_fmaf(w.slope, x, height);
        slope += w.slope;
    }
    if ((w.flags & EVO_GAPS) != 0u) {
        const float spacing = __builtin_fmaf(4.0f, w.gaps, 2.0f);
        const float center = spacing * 0.5f;
        const float r = x - floorf(x / spacing) * spacing;
        const float distance = fabsf(r - center);
        const float half_w = 0.5f * w.gaps;
        const float run = fmaxf(fminf(0.15f, half_w), 1e-6f);
        float factor = clampf_((half_w - distance) / run, 0.0f, 1.0f);
        if (distance <= half_w - run) {
            factor = 1.0f;
        } else if (distance >= half_w) {
            factor = 0.0f;
        }
        height = __builtin_fmaf(-2.0f, factor, height);
        if (distance > half_w - run && distance < half_w) { slope -= 2.0f * ((r < center ? 1.0f : -1.0f) / run); }
    }
    if ((w.flags & EVO_HURDLES) != 0u) {
        const float spacing = 3.0f;
        const float center = spacing * 0.5f;
        const float r = x - floorf(x / spacing) * spacing;
        const float distance = fabsf(r - center);
        const float half_w = 0.5f * 1.2f;
        const float run = 0.2f;
        float factor = clampf_((half_w + run - distance) / run, 0.0f, 1.0f);
        if (distance <= half_w) {
            factor = 1.0f;
        } else if (distance >= half_w + run) {
            factor = 0.0f;
        }
        height = __builtin_fmaf(w.hurdles, factor, height);
        if (distance > half_w && distance < half_w + run) { slope += w.hurdles * ((r < center ? 1.0f : -1.0f) / run); }
    }
    return Ground{height, slope, ds};
}

// shaders/creature.cu:230-235, and its derivative.
EVO_HD float ice_at(float x, float* dice) {
    const float wv = fracf_(x * (1.0f / 6.0f));
    const float t = fabsf(wv - 0.5f) * 2.0f;
    const float raw = (0.7f - t) * 2.5f;
    const float s = clampf_(raw, 0.0f, 1.0f);
    const float ds = (raw > 0.0f && raw < 1.0f) ? -2.5f * 2.0f * (wv < 0.5f ? -1.0f : 1.0f) * (1.0f / 6.0f) : 0.0f;
    *dice = 6.0f * s * (1.0f - s) * ds;
    return s * s * __builtin_fmaf(-2.0f, s, 3.0f);
}

// ---------------------------------------------------------------- forward

// shaders/creature.cu:153-159
EVO_HD float rhythm(float t, float inv_period, float phase, float offset, float duty, float inv_duty,
                    float inv_complement) {
    const float ph = fracf_(__builtin_fmaf(t, inv_period, phase) + offset);
    const float pi = 3.14159265359f;
    if (ph < duty) { return __builtin_fmaf(-0.5f, cos_0_pi(pi * ph * inv_duty), 0.5f); }
    return __builtin_fmaf(0.5f, cos_0_pi(pi * (ph - duty) * inv_complement), 0.5f);
}

// d rhythm / d phase-argument at t (the derivative of frac is 1).
EVO_HD float rhythm_slope(float t, float inv_period, float phase, float offset, float duty, float inv_duty,
                          float inv_complement) {
    const float ph = fracf_(__builtin_fmaf(t, inv_period, phase) + offset);
    const float pi = 3.14159265359f;
    if (ph < duty) {
        const float a = pi * ph * inv_duty;
        return 0.5f * cos_0_pi(fabsf(a - 1.57079632679f)) * pi * inv_duty;
    }
    const float a = pi * (ph - duty) * inv_complement;
    return -0.5f * cos_0_pi(fabsf(a - 1.57079632679f)) * pi * inv_complement;
}

// shaders/creature.cu:442-451
template <int MN, int MM>
EVO_HD void demands(const World& w, Body<MN, MM>& b, float t) {
    for (uint32_t k = 0; k < b.muscles; k++) {
        const float before = rhythm(t, b.minvp[k], b.mphase[k], b.moff[k], b.mduty[k], b.midu[k], b.micomp[k]);
        const float after = rhythm(t + w.dt, b.minvp[k], b.mphase[k], b.moff[k], b.mduty[k], b.midu[k], b.micomp[k]);
        const float shortening = b.mamp[k] * (after - before) * w.rate;
        b.demand[k] = fmaxf(shortening * b.mstiff[k] * 0.25f, 0.0f);
    }
}

struct JointGeo {
    uint32_t tip, far_, hub;
    bool at_head;
    float ux, uy, vx, vy, uu, vv;
    float gtx, gty, gfx, gfy, ghx, ghy, wt, wf, wh;
};

// shaders/creature.cu:320-356 on positions xs, ys.
template <int MN, int MM>
EVO_HD JointGeo joint_geo(const Body<MN, MM>& b, uint32_t j, const float* xs, const float* ys) {
    JointGeo k;
    const uint32_t topo = b.jtopo[j];
    k.tip = j + 1u;
    k.hub = topo & 63u;
    const uint32_t ptip = (topo >> 6u) & 63u, ppiv = (topo >> 12u) & 63u;
    k.far_ = (topo >> 18u) & 63u;
    k.at_head = ((topo >> 24u) & 1u) != 0u;
    k.wt = b.nw[k.tip];
    k.wf = b.nw[k.far_];
    k.wh = b.nw[k.hub];
    k.ux = xs[ptip] - xs[ppiv];
    k.uy = ys[ptip] - ys[ppiv];
    k.vx = xs[k.tip] - xs[k.hub];
    k.vy = ys[k.tip] - ys[k.hub];
    k.uu = fmaxf(__builtin_fmaf(k.ux, k.ux, k.uy * k.uy), 1e-12f);
    k.vv = fmaxf(__builtin_fmaf(k.vx, k.vx, k.vy * k.vy), 1e-12f);
    const float inv = 1.0f / (k.uu * k.vv);
    const float ivv = k.uu * inv, iuu = k.vv * inv;
    k.gtx = -k.vy * ivv;
    k.gty = k.vx * ivv;
    const float gux = -k.uy * iuu, guy = k.ux * iuu;
    k.gfx = k.at_head ? -gux : gux;
    k.gfy = k.at_head ? -guy : guy;
    k.ghx = -(k.gtx + k.gfx);
    k.ghy = -(k.gty + k.gfy);
    return k;
}

EVO_HD float joint_angle(const JointGeo& k, float lo, float hi) {
    const float turn = atan2f_evo(__builtin_fmaf(k.ux, k.vy, -(k.uy * k.vx)), __builtin_fmaf(k.ux, k.vx, k.uy * k.vy));
    const float mid = 0.5f * (lo + hi);
    return mid + wrapf_(turn - mid);
}

EVO_HD float joint_weight(const JointGeo& k) {
    const float nt = __builtin_fmaf(k.gtx, k.gtx, k.gty * k.gty);
    const float nf = __builtin_fmaf(k.gfx, k.gfx, k.gfy * k.gfy);
    const float nh = __builtin_fmaf(k.ghx, k.ghx, k.ghy * k.ghy);
    return __builtin_fmaf(k.wt, nt, __builtin_fmaf(k.wf, nf, k.wh * nh));
}

EVO_HD void joint_push(float* xs, float* ys, const JointGeo& k, float lambda) {
    const float st = k.wt * lambda, sf = k.wf * lambda, sh = k.wh * lambda;
    xs[k.tip] = __builtin_fmaf(st, k.gtx, xs[k.tip]);
    ys[k.tip] = __builtin_fmaf(st, k.gty, ys[k.tip]);
    xs[k.far_] = __builtin_fmaf(sf, k.gfx, xs[k.far_]);
    ys[k.far_] = __builtin_fmaf(sf, k.gfy, ys[k.far_]);
    xs[k.hub] = __builtin_fmaf(sh, k.ghx, xs[k.hub]);
    ys[k.hub] = __builtin_fmaf(sh, k.ghy, ys[k.hub]);
}

// One substep from the state in b (shaders/creature.cu:705-742). With a
// tape, records what the reverse needs. The state in b becomes the state at
// the substep's end.
template <int MN, int MM, bool TAPE>
EVO_HD void substep(const World& w, Body<MN, MM>& b, Tape<MN>* t) {
    float fx[MN], fy[MN], qx[MN], qy[MN];
    // muscle forces (shaders/creature.cu:455-508)
    for (uint32_t i = 0; i < b.nodes; i++) {
        fx[i] = 0.0f;
        fy[i] = 0.0f;
    }
    for (uint32_t k = 0; k < b.muscles; k++) {
        const uint32_t nodes = b.mnodes[k];
        const uint32_t a0 = nodes & 255u, a1 = (nodes >> 8u) & 255u, b0 = (nodes >> 16u) & 255u, b1 = nodes >> 24u;
        const float aa = b.maa[k], ab = b.mab[k];
        const float ax = __builtin_fmaf(b.px[a1] - b.px[a0], aa, b.px[a0]);
        const float ay = __builtin_fmaf(b.py[a1] - b.py[a0], aa, b.py[a0]);
        const float avx = __builtin_fmaf(b.vx[a1] - b.vx[a0], aa, b.vx[a0]);
        const float avy = __builtin_fmaf(b.vy[a1] - b.vy[a0], aa, b.vy[a0]);
        const float bx = __builtin_fmaf(b.px[b1] - b.px[b0], ab, b.px[b0]);
        const float by = __builtin_fmaf(b.py[b1] - b.py[b0], ab, b.py[b0]);
        const float bvx = __builtin_fmaf(b.vx[b1] - b.vx[b0], ab, b.vx[b0]);
        const float bvy = __builtin_fmaf(b.vy[b1] - b.vy[b0], ab, b.vy[b0]);
        const float dx = bx - ax, dy = by - ay;
        const float len = fmaxf(sqrtf(__builtin_fmaf(dx, dx, dy * dy)), 1e-6f);
        const float inv_len = 1.0f / len;
        const float ex = dx * inv_len, ey = dy * inv_len;
        const float lengthening = __builtin_fmaf(bvx - avx, ex, (bvy - avy) * ey);
        const float store = b.store[k];
        float drive = b.demand[k] * store;
        if (b.mhill[k] > 0.0f) { drive *= clampf_(__builtin_fmaf(lengthening, b.mhill[k], 1.0f), 0.0f, 1.0f); }
        const float cap = b.mcap[k];
        const float active = clampf_(__builtin_fmaf(lengthening, 0.15f, drive), -cap, cap);
        const float work = fminf(drive, cap) * fmaxf(-lengthening, 0.0f) * w.h;
        b.store[k] = clampf_(__builtin_fmaf(w.rec_h, 1.0f - store, store - work * b.minvcap[k] * w.inv_energy), 0.0f,
                             1.0f);
        const float pull = __builtin_fmaf(b.mtk[k], fmaxf(len - b.mslack[k], 0.0f), active);
        const float fxk = ex * pull, fyk = ey * pull;
        const float ra = 1.0f - aa, rb = 1.0f - ab;
        fx[a0] += fxk * ra;
        fy[a0] += fyk * ra;
        fx[a1] += fxk * aa;
        fy[a1] += fyk * aa;
        fx[b0] += -(fxk * rb);
        fy[b0] += -(fyk * rb);
        fx[b1] += -(fxk * ab);
        fy[b1] += -(fyk * ab);
    }
    if constexpr (TAPE) {
        for (uint32_t i = 0; i < b.nodes; i++) {
            t->fx[i] = fx[i];
            t->fy[i] = fy[i];
        }
    }
    // bone drag (shaders/creature.cu:553-593), air only
    for (uint32_t j = 0; j < b.bones; j++) {
        const uint32_t i0 = b.bi0[j], i1 = j + 1u;
        if constexpr (TAPE) {
            t->drag_v[j][0] = b.vx[i0];
            t->drag_v[j][1] = b.vy[i0];
            t->drag_v[j][2] = b.vx[i1];
            t->drag_v[j][3] = b.vy[i1];
        }
        const float wx = 0.5f * (b.vx[i0] + b.vx[i1]), wy = 0.5f * (b.vy[i0] + b.vy[i1]);
        const float speed = sqrtf(__builtin_fmaf(wx, wx, wy * wy));
        const float k_air = fminf(b.bk[j] * speed, b.blim[j]);
        const float dfx = -wx * k_air, dfy = -wy * k_air;
        b.vx[i0] = __builtin_fmaf(b.bs0[j], dfx, b.vx[i0]);
        b.vy[i0] = __builtin_fmaf(b.bs0[j], dfy, b.vy[i0]);
        b.vx[i1] = __builtin_fmaf(b.bs1[j], dfx, b.vx[i1]);
        b.vy[i1] = __builtin_fmaf(b.bs1[j], dfy, b.vy[i1]);
    }
    // integrate (shaders/creature.cu:714-728)
    for (uint32_t i = 0; i < b.nodes; i++) {
        const float m = b.nm[i];
        float gx = 0.0f;
        const float gy = -w.gravity * m;
        if ((w.flags & EVO_WIND) != 0u) { gx = __builtin_fmaf(w.wind, m, gx); }
        const float wi = b.nw[i];
        b.vx[i] = __builtin_fmaf(w.h * (gx + fx[i]), wi, b.vx[i]);
        b.vy[i] = __builtin_fmaf(w.h * (gy + fy[i]), wi, b.vy[i]);
        if ((w.flags & EVO_AIR) != 0u) {
            b.vx[i] *= w.air_sub;
            b.vy[i] *= w.air_sub;
        }
        qx[i] = b.px[i];
        qy[i] = b.py[i];
        b.px[i] = __builtin_fmaf(w.h, b.vx[i], b.px[i]);
        b.py[i] = __builtin_fmaf(w.h, b.vy[i], b.py[i]);
    }
    // bones (shaders/creature.cu:597-613)
    for (uint32_t j = 0; j < b.bones; j++) {
        const uint32_t i0 = b.bi0[j], i1 = j + 1u;
        if constexpr (TAPE) {
            t->bone_x[j][0] = b.px[i0];
            t->bone_x[j][1] = b.py[i0];
            t->bone_x[j][2] = b.px[i1];
            t->bone_x[j][3] = b.py[i1];
        }
        const float dx = b.px[i1] - b.px[i0], dy = b.py[i1] - b.py[i0];
        const float len = fmaxf(sqrtf(__builtin_fmaf(dx, dx, dy * dy)), 1e-9f);
        const float w0 = b.nw[i0], w1 = b.nw[i1];
        const float s = (len - b.blen[j]) / ((w0 + w1) * len);
        const float s0 = w0 * s, s1 = w1 * s;
        b.px[i0] = __builtin_fmaf(s0, dx, b.px[i0]);
        b.py[i0] = __builtin_fmaf(s0, dy, b.py[i0]);
        b.px[i1] = __builtin_fmaf(-s1, dx, b.px[i1]);
        b.py[i1] = __builtin_fmaf(-s1, dy, b.py[i1]);
    }
    // joints (shaders/creature.cu:618-634)
    for (uint32_t j = 1; j < b.bones; j++) {
        if (b.jtopo[j] == kNone) { continue; }
        const JointGeo k = joint_geo(b, j, b.px, b.py);
        if constexpr (TAPE) {
            t->joint_x[j][0] = b.px[k.tip];
            t->joint_x[j][1] = b.py[k.tip];
            t->joint_x[j][2] = b.px[k.far_];
            t->joint_x[j][3] = b.py[k.far_];
            t->joint_x[j][4] = b.px[k.hub];
            t->joint_x[j][5] = b.py[k.hub];
        }
        const float angle = joint_angle(k, b.jlo[j], b.jhi[j]);
        float error = 0.0f;
        if (angle < b.jlo[j]) { error = angle - b.jlo[j]; }
        if (angle > b.jhi[j]) { error = angle - b.jhi[j]; }
        if (error == 0.0f) { continue; }
        const float wgt = joint_weight(k);
        if (wgt > 0.0f) { joint_push(b.px, b.py, k, -error / wgt); }
    }
    // ground and velocity (shaders/creature.cu:640-679, :733-740)
    for (uint32_t i = 0; i < b.nodes; i++) {
        if constexpr (TAPE) {
            t->ground_x[i][0] = b.px[i];
            t->ground_x[i][1] = b.py[i];
        }
        if ((w.flags & EVO_GROUND) != 0u) {
            const Ground g = ground(w, b.px[i], b.amp, b.qphase);
            const float secant = sqrtf(__builtin_fmaf(g.s, g.s, 1.0f));
            const float inv_sec = 1.0f / secant;
            const float nx = -g.s * inv_sec, ny = inv_sec;
            const float dry = __builtin_fmaf(b.py[i] - g.h, inv_sec, -b.nr[i]);
            const float depth = -dry;
            float mu = b.nf[i] * w.friction;
            if (depth > 0.0f) {
                if ((w.flags & EVO_ICE) != 0u) {
                    float dice;
                    mu *= __builtin_fmaf(-w.patches, ice_at(b.px[i], &dice), 1.0f);
                }
                b.px[i] = __builtin_fmaf(nx, depth, b.px[i]);
                b.py[i] = __builtin_fmaf(ny, depth, b.py[i]);
                const float tx = ny, ty = -nx;
                const float slide = __builtin_fmaf(b.px[i] - qx[i], tx, (b.py[i] - qy[i]) * ty);
                const float budget = mu * depth;
                const float back = fabsf(slide) <= budget ? slide : copysignf(budget, slide);
                b.px[i] = __builtin_fmaf(-tx, back, b.px[i]);
                b.py[i] = __builtin_fmaf(-ty, back, b.py[i]);
            }
        }
        b.vx[i] = (b.px[i] - qx[i]) * w.inv_h;
        b.vy[i] = (b.py[i] - qy[i]) * w.inv_h;
    }
    // joint damping (shaders/creature.cu:683-702)
    for (uint32_t j = 1; j < b.bones; j++) {
        if (b.jtopo[j] == kNone) { continue; }
        const JointGeo k = joint_geo(b, j, b.px, b.py);
        if constexpr (TAPE) {
            t->damp_v[j][0] = b.vx[k.tip];
            t->damp_v[j][1] = b.vy[k.tip];
            t->damp_v[j][2] = b.vx[k.far_];
            t->damp_v[j][3] = b.vy[k.far_];
            t->damp_v[j][4] = b.vx[k.hub];
            t->damp_v[j][5] = b.vy[k.hub];
        }
        float rate = k.gtx * b.vx[k.tip];
        rate = __builtin_fmaf(k.gty, b.vy[k.tip], rate);
        rate = __builtin_fmaf(k.gfx, b.vx[k.far_], rate);
        rate = __builtin_fmaf(k.gfy, b.vy[k.far_], rate);
        rate = __builtin_fmaf(k.ghx, b.vx[k.hub], rate);
        rate = __builtin_fmaf(k.ghy, b.vy[k.hub], rate);
        const float wgt = joint_weight(k);
        if (wgt > 0.0f) { joint_push(b.vx, b.vy, k, (-w.damp_share * rate) / wgt); }
    }
    if constexpr (TAPE) {
        for (uint32_t i = 0; i < b.nodes; i++) {
            t->end_x[i] = b.px[i];
            t->end_y[i] = b.py[i];
        }
    }
}

// The centre of mass x: the fitness at the end of a trial that does not
// fail (shaders/creature.cu:997-1003).
template <int MN, int MM>
EVO_HD float com_x(const Body<MN, MM>& b) {
    float c = 0.0f;
    for (uint32_t i = 0; i < b.nodes; i++) { c = __builtin_fmaf(b.nm[i], b.px[i], c); }
    return c * b.inv_mass;
}

// ---------------------------------------------------------------- reverse

// The transposed Jacobian of a joint's gradient vectors and angle. Given
// adjoints of gt, gu (the parent bone's raw gradient, before the at-head
// sign) and of the angle, adds d/du and d/dv into (dux, duy, dvx, dvy).
// gt = (-vy, vx) / |v|^2, gu = (-uy, ux) / |u|^2, and the angle is the turn
// from u to v, whose gradient is (-vy, vx) / |v|^2 in v and (uy, -ux) / |u|^2
// in u.
EVO_HD void joint_geo_back(const JointGeo& k, float bgtx, float bgty, float bgux, float bguy, float bangle,
                           float& dux, float& duy, float& dvx, float& dvy) {
    const float ivv = 1.0f / k.vv, iuu = 1.0f / k.uu;
    // gt = (-vy, vx) * ivv
    {
        const float vx = k.vx, vy = k.vy;
        const float i2 = ivv * ivv;
        // d gtx / d vx = 2 vx vy / vv^2, d gtx / d vy = -1/vv + 2 vy^2 / vv^2
        // d gty / d vx = 1/vv - 2 vx^2 / vv^2, d gty / d vy = -2 vx vy / vv^2
        dvx += bgtx * (2.0f * vx * vy * i2) + bgty * (ivv - 2.0f * vx * vx * i2);
        dvy += bgtx * (-ivv + 2.0f * vy * vy * i2) + bgty * (-2.0f * vx * vy * i2);
        dvx += bangle * (-vy * ivv);
        dvy += bangle * (vx * ivv);
    }
    {
        const float ux = k.ux, uy = k.uy;
        const float i2 = iuu * iuu;
        dux += bgux * (2.0f * ux * uy * i2) + bguy * (iuu - 2.0f * ux * ux * i2);
        duy += bgux * (-iuu + 2.0f * uy * uy * i2) + bguy * (-2.0f * ux * uy * i2);
        dux += bangle * (uy * iuu);
        duy += bangle * (-ux * iuu);
    }
}

// Adds (dux, duy) to the parent bone's nodes and (dvx, dvy) to the bone's.
template <int MN, int MM>
EVO_HD void joint_uv_back(const Body<MN, MM>& b, uint32_t j, const JointGeo& k, float dux, float duy, float dvx,
                          float dvy, float* ax, float* ay) {
    const uint32_t topo = b.jtopo[j];
    const uint32_t ptip = (topo >> 6u) & 63u, ppiv = (topo >> 12u) & 63u;
    ax[k.tip] += dvx;
    ay[k.tip] += dvy;
    ax[k.hub] -= dvx;
    ay[k.hub] -= dvy;
    ax[ptip] += dux;
    ay[ptip] += duy;
    ax[ppiv] -= dux;
    ay[ppiv] -= duy;
}

// Reverse of joint_push(xs, ys, k, lambda) for lambda = scale * num / wgt,
// where the caller differentiates num. Reads the adjoints of the pushed
// values in (ax, ay) and leaves them there (the push adds to its input).
// Returns the adjoint of num and adds the gradient-vector adjoints.
struct PushBack {
    float bnum, bgtx, bgty, bgux, bguy;
};
EVO_HD PushBack push_back(const JointGeo& k, float lambda, float num, float wgt, const float* ax, const float* ay) {
    const float st = k.wt * lambda, sf = k.wf * lambda, sh = k.wh * lambda;
    const float bst = ax[k.tip] * k.gtx + ay[k.tip] * k.gty;
    const float bsf = ax[k.far_] * k.gfx + ay[k.far_] * k.gfy;
    const float bsh = ax[k.hub] * k.ghx + ay[k.hub] * k.ghy;
    float bgtx = st * ax[k.tip], bgty = st * ay[k.tip];
    float bgfx = sf * ax[k.far_], bgfy = sf * ay[k.far_];
    const float bghx = sh * ax[k.hub], bghy = sh * ay[k.hub];
    const float blambda = k.wt * bst + k.wf * bsf + k.wh * bsh;
    // lambda = num / wgt (with the scale folded into num by the caller)
    const float bnum = blambda / wgt;
    const float bwgt = -blambda * num / (wgt * wgt);
    // wgt = wt |gt|^2 + wf |gf|^2 + wh |gh|^2, gh = -(gt + gf)
    float bghx2 = bghx + 2.0f * k.wh * bwgt * k.ghx, bghy2 = bghy + 2.0f * k.wh * bwgt * k.ghy;
    bgtx += 2.0f * k.wt * bwgt * k.gtx - bghx2;
    bgty += 2.0f * k.wt * bwgt * k.gty - bghy2;
    bgfx += 2.0f * k.wf * bwgt * k.gfx - bghx2;
    bgfy += 2.0f * k.wf * bwgt * k.gfy - bghy2;
    PushBack r;
    r.bnum = bnum;
    r.bgtx = bgtx;
    r.bgty = bgty;
    r.bgux = k.at_head ? -bgfx : bgfx;
    r.bguy = k.at_head ? -bgfy : bgfy;
    return r;
}

// Reverses one substep. `b` holds the substep's start state (positions,
// velocities, stores) and the step's demands; `t` the tape of its forward
// recomputation. `a` holds the adjoints of the end state on entry and of
// the start state on return; parameter gradients accumulate in `a`.
template <int MN, int MM>
EVO_HD void substep_back(const World& w, const Body<MN, MM>& b, const Tape<MN>& t, Adj<MN, MM>& a) {
    float ex_[MN], ey_[MN];  // positions at the end of the substep
    for (uint32_t i = 0; i < b.nodes; i++) {
        ex_[i] = t.end_x[i];
        ey_[i] = t.end_y[i];
    }
    // --- joint damping, last joint first. Positions are the end positions.
    for (uint32_t j = b.bones - 1u; j >= 1u && j < b.bones; j--) {
        if (b.jtopo[j] == kNone) { continue; }
        const JointGeo k = joint_geo(b, j, ex_, ey_);
        const float wgt = joint_weight(k);
        if (!(wgt > 0.0f)) { continue; }
        const float v0x = t.damp_v[j][0], v0y = t.damp_v[j][1], v1x = t.damp_v[j][2], v1y = t.damp_v[j][3];
        const float v2x = t.damp_v[j][4], v2y = t.damp_v[j][5];
        float rate = k.gtx * v0x;
        rate = __builtin_fmaf(k.gty, v0y, rate);
        rate = __builtin_fmaf(k.gfx, v1x, rate);
        rate = __builtin_fmaf(k.gfy, v1y, rate);
        rate = __builtin_fmaf(k.ghx, v2x, rate);
        rate = __builtin_fmaf(k.ghy, v2y, rate);
        const float num = -w.damp_share * rate;
        const float lambda = num / wgt;
        PushBack p = push_back(k, lambda, num, wgt, a.vx, a.vy);
        const float brate = -w.damp_share * p.bnum;
        a.vx[k.tip] += brate * k.gtx;
        a.vy[k.tip] += brate * k.gty;
        a.vx[k.far_] += brate * k.gfx;
        a.vy[k.far_] += brate * k.gfy;
        a.vx[k.hub] += brate * k.ghx;
        a.vy[k.hub] += brate * k.ghy;
        // rate's dependence on the gradient vectors
        float bgtx = p.bgtx + brate * v0x, bgty = p.bgty + brate * v0y;
        const float bgfx_r = brate * v1x - brate * v2x, bgfy_r = brate * v1y - brate * v2y;
        bgtx -= brate * v2x;
        bgty -= brate * v2y;
        const float bgux = p.bgux + (k.at_head ? -bgfx_r : bgfx_r);
        const float bguy = p.bguy + (k.at_head ? -bgfy_r : bgfy_r);
        float dux = 0.0f, duy = 0.0f, dvx = 0.0f, dvy = 0.0f;
        joint_geo_back(k, bgtx, bgty, bgux, bguy, 0.0f, dux, duy, dvx, dvy);
        joint_uv_back(b, j, k, dux, duy, dvx, dvy, a.px, a.py);
    }
    // --- velocity from positions: v = (x - q) / h, q = the start positions
    float aqx[MN], aqy[MN];
    for (uint32_t i = 0; i < b.nodes; i++) {
        a.px[i] += a.vx[i] * w.inv_h;
        a.py[i] += a.vy[i] * w.inv_h;
        aqx[i] = -a.vx[i] * w.inv_h;
        aqy[i] = -a.vy[i] * w.inv_h;
        a.vx[i] = 0.0f;
        a.vy[i] = 0.0f;
    }
    // --- ground, node by node
    if ((w.flags & EVO_GROUND) != 0u) {
        for (uint32_t i = 0; i < b.nodes; i++) {
            const float x0 = t.ground_x[i][0], y0 = t.ground_x[i][1];
            const Ground g = ground(w, x0, b.amp, b.qphase);
            const float secant = sqrtf(__builtin_fmaf(g.s, g.s, 1.0f));
            const float inv_sec = 1.0f / secant;
            const float nx = -g.s * inv_sec, ny = inv_sec;
            const float dry = __builtin_fmaf(y0 - g.h, inv_sec, -b.nr[i]);
            const float depth = -dry;
            if (!(depth > 0.0f)) { continue; }
            const float mu0 = b.nf[i] * w.friction;
            float mu = mu0, ice = 0.0f, dice = 0.0f, ice_f = 1.0f;
            if ((w.flags & EVO_ICE) != 0u) {
                ice = ice_at(x0, &dice);
                ice_f = __builtin_fmaf(-w.patches, ice, 1.0f);
                mu = mu0 * ice_f;
            }
            const float x1 = __builtin_fmaf(nx, depth, x0), y1 = __builtin_fmaf(ny, depth, y0);
            const float tx = ny, ty = -nx;
            const float qx = b.px[i], qy = b.py[i];
            const float slide = __builtin_fmaf(x1 - qx, tx, (y1 - qy) * ty);
            const float budget = mu * depth;
            const bool stick = fabsf(slide) <= budget;
            const float back = stick ? slide : copysignf(budget, slide);
            // x2 = x1 - tx back, y2 = y1 - ty back
            const float ax2 = a.px[i], ay2 = a.py[i];
            const float bback = -(ax2 * tx + ay2 * ty);
            float btx = -ax2 * back, bty = -ay2 * back;
            float ax1 = ax2, ay1 = ay2;
            float bslide = 0.0f, bbudget = 0.0f;
            if (stick) {
                bslide = bback;
            } else {
                bbudget = slide < 0.0f ? -bback : bback;
            }
            const float bmu = bbudget * depth;
            float bdepth = bbudget * mu;
            // slide = (x1 - qx) tx + (y1 - qy) ty
            ax1 += bslide * tx;
            ay1 += bslide * ty;
            aqx[i] -= bslide * tx;
            aqy[i] -= bslide * ty;
            btx += bslide * (x1 - qx);
            bty += bslide * (y1 - qy);
            // tx = ny, ty = -nx
            float bny = btx, bnx = -bty;
            // x1 = x0 + nx depth, y1 = y0 + ny depth
            float ax0 = ax1, ay0 = ay1;
            bnx += ax1 * depth;
            bny += ay1 * depth;
            bdepth += ax1 * nx + ay1 * ny;
            // depth = -((y0 - gh) inv_sec - r)
            ay0 += -bdepth * inv_sec;
            const float bgh = bdepth * inv_sec;
            float binv = -bdepth * (y0 - g.h);
            // nx = -gs inv_sec, ny = inv_sec
            float bgs = -bnx * inv_sec;
            binv += -bnx * g.s + bny;
            // inv_sec = 1 / sqrt(1 + gs^2)
            bgs += binv * (-g.s * inv_sec * inv_sec * inv_sec);
            // mu = f friction (1 - patches ice(x0))
            a.g_fric[i] += bmu * w.friction * ice_f;
            ax0 += bmu * mu0 * (-w.patches) * dice;
            // gh(x0), gs(x0)
            ax0 += bgh * g.s + bgs * g.ds;
            a.px[i] = ax0;
            a.py[i] = ay0;
        }
    }
    // --- joints, last first: positions before each push from the tape
    for (uint32_t j = b.bones - 1u; j >= 1u && j < b.bones; j--) {
        if (b.jtopo[j] == kNone) { continue; }
        // positions of the joint's nodes before its push; the other nodes'
        // positions do not enter the joint
        float xs[MN], ys[MN];
        const uint32_t topo = b.jtopo[j];
        const uint32_t tip = j + 1u, hub = topo & 63u, far_ = (topo >> 18u) & 63u;
        xs[tip] = t.joint_x[j][0];
        ys[tip] = t.joint_x[j][1];
        xs[far_] = t.joint_x[j][2];
        ys[far_] = t.joint_x[j][3];
        xs[hub] = t.joint_x[j][4];
        ys[hub] = t.joint_x[j][5];
        const JointGeo k = joint_geo(b, j, xs, ys);
        const float angle = joint_angle(k, b.jlo[j], b.jhi[j]);
        float error = 0.0f;
        if (angle < b.jlo[j]) { error = angle - b.jlo[j]; }
        if (angle > b.jhi[j]) { error = angle - b.jhi[j]; }
        if (error == 0.0f) { continue; }
        const float wgt = joint_weight(k);
        if (!(wgt > 0.0f)) { continue; }
        const float num = -error;
        const float lambda = num / wgt;
        PushBack p = push_back(k, lambda, num, wgt, a.px, a.py);
        const float bangle = -p.bnum;
        float dux = 0.0f, duy = 0.0f, dvx = 0.0f, dvy = 0.0f;
        joint_geo_back(k, p.bgtx, p.bgty, p.bgux, p.bguy, bangle, dux, duy, dvx, dvy);
        joint_uv_back(b, j, k, dux, duy, dvx, dvy, a.px, a.py);
    }
    // --- bones, last first
    for (uint32_t j = b.bones; j-- > 0u;) {
        const uint32_t i0 = b.bi0[j], i1 = j + 1u;
        const float x0 = t.bone_x[j][0], y0 = t.bone_x[j][1], x1 = t.bone_x[j][2], y1 = t.bone_x[j][3];
        const float dx = x1 - x0, dy = y1 - y0;
        const float raw = sqrtf(__builtin_fmaf(dx, dx, dy * dy));
        const float len = fmaxf(raw, 1e-9f);
        const float w0 = b.nw[i0], w1 = b.nw[i1];
        const float W = w0 + w1;
        const float s = (len - b.blen[j]) / (W * len);
        const float s0 = w0 * s, s1 = w1 * s;
        const float a0x = a.px[i0], a0y = a.py[i0], a1x = a.px[i1], a1y = a.py[i1];
        const float bs0 = a0x * dx + a0y * dy;
        const float bs1 = -(a1x * dx + a1y * dy);
        float bdx = a0x * s0 - a1x * s1, bdy = a0y * s0 - a1y * s1;
        const float bs = bs0 * w0 + bs1 * w1;
        // s = 1/W - L / (W len)
        const float blen = bs * b.blen[j] / (W * len * len);
        if (raw > 1e-9f) {
            bdx += blen * dx / len;
            bdy += blen * dy / len;
        }
        a.px[i1] += bdx;
        a.py[i1] += bdy;
        a.px[i0] -= bdx;
        a.py[i0] -= bdy;
    }
    // --- integrate: x1 = q + h va, va = (vd + h (g + F) wi) air
    float afx[MN], afy[MN];
    for (uint32_t i = 0; i < b.nodes; i++) {
        aqx[i] += a.px[i];
        aqy[i] += a.py[i];
        float avx = a.px[i] * w.h, avy = a.py[i] * w.h;
        if ((w.flags & EVO_AIR) != 0u) {
            avx *= w.air_sub;
            avy *= w.air_sub;
        }
        afx[i] = avx * w.h * b.nw[i];
        afy[i] = avy * w.h * b.nw[i];
        a.vx[i] = avx;
        a.vy[i] = avy;
        // the start positions' adjoint: through q (previous position)
        a.px[i] = aqx[i];
        a.py[i] = aqy[i];
    }
    // --- bone drag, last bone first, from the velocities before each drag
    for (uint32_t j = b.bones; j-- > 0u;) {
        const uint32_t i0 = b.bi0[j], i1 = j + 1u;
        const float v0x = t.drag_v[j][0], v0y = t.drag_v[j][1], v1x = t.drag_v[j][2], v1y = t.drag_v[j][3];
        const float wx = 0.5f * (v0x + v1x), wy = 0.5f * (v0y + v1y);
        const float speed = sqrtf(__builtin_fmaf(wx, wx, wy * wy));
        const float lin = b.bk[j] * speed;
        const float k_air = fminf(lin, b.blim[j]);
        const float bdfx = b.bs0[j] * a.vx[i0] + b.bs1[j] * a.vx[i1];
        const float bdfy = b.bs0[j] * a.vy[i0] + b.bs1[j] * a.vy[i1];
        float bwx = -bdfx * k_air, bwy = -bdfy * k_air;
        const float bk = -(bdfx * wx + bdfy * wy);
        if (lin < b.blim[j] && speed > 0.0f) {
            const float bspeed = bk * b.bk[j];
            bwx += bspeed * wx / speed;
            bwy += bspeed * wy / speed;
        }
        a.vx[i0] += 0.5f * bwx;
        a.vx[i1] += 0.5f * bwx;
        a.vy[i0] += 0.5f * bwy;
        a.vy[i1] += 0.5f * bwy;
    }
    // --- muscles, from the start state
    for (uint32_t k = 0; k < b.muscles; k++) {
        const uint32_t nodes = b.mnodes[k];
        const uint32_t a0 = nodes & 255u, a1 = (nodes >> 8u) & 255u, b0 = (nodes >> 16u) & 255u, b1 = nodes >> 24u;
        const float aa = b.maa[k], ab = b.mab[k];
        const float ra = 1.0f - aa, rb = 1.0f - ab;
        const float ax = __builtin_fmaf(b.px[a1] - b.px[a0], aa, b.px[a0]);
        const float ay = __builtin_fmaf(b.py[a1] - b.py[a0], aa, b.py[a0]);
        const float avx = __builtin_fmaf(b.vx[a1] - b.vx[a0], aa, b.vx[a0]);
        const float avy = __builtin_fmaf(b.vy[a1] - b.vy[a0], aa, b.vy[a0]);
        const float bx = __builtin_fmaf(b.px[b1] - b.px[b0], ab, b.px[b0]);
        const float by = __builtin_fmaf(b.py[b1] - b.py[b0], ab, b.py[b0]);
        const float bvx = __builtin_fmaf(b.vx[b1] - b.vx[b0], ab, b.vx[b0]);
        const float bvy = __builtin_fmaf(b.vy[b1] - b.vy[b0], ab, b.vy[b0]);
        const float dx = bx - ax, dy = by - ay;
        const float raw = sqrtf(__builtin_fmaf(dx, dx, dy * dy));
        const float len = fmaxf(raw, 1e-6f);
        const float inv_len = 1.0f / len;
        const float ex = dx * inv_len, ey = dy * inv_len;
        const float dvx = bvx - avx, dvy = bvy - avy;
        const float lengthening = __builtin_fmaf(dvx, ex, dvy * ey);
        const float store = b.store[k];
        const float drive0 = b.demand[k] * store;
        float hc = 1.0f, hraw = 1.0f;
        const bool hill = b.mhill[k] > 0.0f;
        if (hill) {
            hraw = __builtin_fmaf(lengthening, b.mhill[k], 1.0f);
            hc = clampf_(hraw, 0.0f, 1.0f);
        }
        const float drive = hill ? drive0 * hc : drive0;
        const float cap = b.mcap[k];
        const float q = __builtin_fmaf(lengthening, 0.15f, drive);
        const float active = clampf_(q, -cap, cap);
        const float mn = fminf(drive, cap), pos = fmaxf(-lengthening, 0.0f);
        const float work = mn * pos * w.h;
        const float zin = __builtin_fmaf(w.rec_h, 1.0f - store, store - work * b.minvcap[k] * w.inv_energy);
        const float pull = __builtin_fmaf(b.mtk[k], fmaxf(len - b.mslack[k], 0.0f), active);
        // shares -> pull, direction
        const float bfx = afx[a0] * ra + afx[a1] * aa - afx[b0] * rb - afx[b1] * ab;
        const float bfy = afy[a0] * ra + afy[a1] * aa - afy[b0] * rb - afy[b1] * ab;
        const float bpull = bfx * ex + bfy * ey;
        float bex = bfx * pull, bey = bfy * pull;
        float blen = 0.0f;
        if (b.mtk[k] != 0.0f && len - b.mslack[k] > 0.0f) { blen += bpull * b.mtk[k]; }
        const float bactive = bpull;
        // the store update
        const float bz = (zin > 0.0f && zin < 1.0f) ? a.store[k] : 0.0f;
        float bstore = bz * (1.0f - w.rec_h);
        const float bwork = -bz * b.minvcap[k] * w.inv_energy;
        a.g_invcap[k] += -bz * work * w.inv_energy;
        // active = clamp(q, -cap, cap)
        float bq = 0.0f, bcap = 0.0f;
        if (q < -cap) {
            bcap -= bactive;
        } else if (q > cap) {
            bcap += bactive;
        } else {
            bq = bactive;
        }
        float blength = bq * 0.15f;
        float bdrive = bq;
        // work = min(drive, cap) pos h
        const float bmn = bwork * pos * w.h;
        const float bpos = bwork * mn * w.h;
        if (drive < cap) {
            bdrive += bmn;
        } else {
            bcap += bmn;
        }
        if (-lengthening > 0.0f) { blength -= bpos; }
        // drive = drive0 hc
        float bdrive0 = bdrive;
        if (hill) {
            bdrive0 = bdrive * hc;
            const float bhc = bdrive * drive0;
            if (hraw > 0.0f && hraw < 1.0f) {
                blength += bhc * b.mhill[k];
                a.g_hill[k] += bhc * lengthening;
            }
        }
        a.demand[k] += bdrive0 * store;
        bstore += bdrive0 * b.demand[k];
        a.store[k] = bstore;
        a.g_cap[k] += bcap;
        // lengthening = dv . e
        const float bdvx = blength * ex, bdvy = blength * ey;
        bex += blength * dvx;
        bey += blength * dvy;
        // e = d / len
        float bdx = bex * inv_len, bdy = bey * inv_len;
        const float binv = bex * dx + bey * dy;
        blen += -binv * inv_len * inv_len;
        if (raw > 1e-6f) {
            bdx += blen * dx / len;
            bdy += blen * dy / len;
        }
        // anchors
        a.px[a1] -= bdx * aa;
        a.px[a0] -= bdx * ra;
        a.py[a1] -= bdy * aa;
        a.py[a0] -= bdy * ra;
        a.px[b1] += bdx * ab;
        a.px[b0] += bdx * rb;
        a.py[b1] += bdy * ab;
        a.py[b0] += bdy * rb;
        a.vx[a1] -= bdvx * aa;
        a.vx[a0] -= bdvx * ra;
        a.vy[a1] -= bdvy * aa;
        a.vy[a0] -= bdvy * ra;
        a.vx[b1] += bdvx * ab;
        a.vx[b0] += bdvx * rb;
        a.vy[b1] += bdvy * ab;
        a.vy[b0] += bdvy * rb;
    }
}

// Adds the adjoints of a step's demands (computed at time t) to the rhythm
// parameters' gradients and clears them.
template <int MN, int MM>
EVO_HD void demands_back(const World& w, const Body<MN, MM>& b, float t, Adj<MN, MM>& a) {
    for (uint32_t k = 0; k < b.muscles; k++) {
        const float bd = a.demand[k];
        a.demand[k] = 0.0f;
        const float before = rhythm(t, b.minvp[k], b.mphase[k], b.moff[k], b.mduty[k], b.midu[k], b.micomp[k]);
        const float after = rhythm(t + w.dt, b.minvp[k], b.mphase[k], b.moff[k], b.mduty[k], b.midu[k], b.micomp[k]);
        const float shortening = b.mamp[k] * (after - before) * w.rate;
        if (!(shortening * b.mstiff[k] * 0.25f > 0.0f)) { continue; }
        const float bsh = bd * b.mstiff[k] * 0.25f;
        a.g_amp[k] += bsh * (after - before) * w.rate;
        const float bafter = bsh * b.mamp[k] * w.rate, bbefore = -bafter;
        const float da = rhythm_slope(t + w.dt, b.minvp[k], b.mphase[k], b.moff[k], b.mduty[k], b.midu[k], b.micomp[k]);
        const float db = rhythm_slope(t, b.minvp[k], b.mphase[k], b.moff[k], b.mduty[k], b.midu[k], b.micomp[k]);
        a.g_phase[k] += bafter * da + bbefore * db;
        a.g_invp[k] += bafter * da * (t + w.dt) + bbefore * db * t;
    }
}

// ---------------------------------------------------------------- a trial

// The start state of a substep (positions, velocities, energy stores) as
// 4 MN + MM words, for checkpoints and the per-substep tape.
template <int MN, int MM>
struct StateIO {
    static constexpr int kWords = 4 * MN + MM;
};

// Runs a whole trial of `steps` steps forward from the loaded state, keeping
// a checkpoint every `seg` steps, then backwards segment by segment: each
// segment is recomputed from its checkpoint with the start state of every
// substep kept (`io.put`/`io.get`, slot = substep of the segment), and each
// substep is then recomputed once more with its tape and reversed. Leaves
// the gradient in `a` and returns the fitness (the centre of mass x at the
// end). `io` provides ckpt_put/ckpt_get(segment, body) and
// put/get(slot, body).
template <int MN, int MM, class IO>
EVO_HD float trial_gradient(const World& w, Body<MN, MM>& b, uint32_t steps, uint32_t substeps, uint32_t seg,
                            IO& io, Adj<MN, MM>& a) {
    for (uint32_t s = 0; s < steps; s++) {
        if (s % seg == 0u) { io.ckpt_put(s / seg, b); }
        demands(w, b, static_cast<float>(s) * w.dt);
        for (uint32_t sub = 0; sub < substeps; sub++) { substep<MN, MM, false>(w, b, nullptr); }
    }
    const float fitness = com_x(b);
    for (uint32_t i = 0; i < static_cast<uint32_t>(MN); i++) {
        a.px[i] = i < b.nodes ? b.nm[i] * b.inv_mass : 0.0f;
        a.py[i] = a.vx[i] = a.vy[i] = 0.0f;
        a.g_fric[i] = 0.0f;
    }
    for (uint32_t k = 0; k < static_cast<uint32_t>(MM); k++) {
        a.store[k] = a.demand[k] = 0.0f;
        a.g_amp[k] = a.g_invp[k] = a.g_phase[k] = a.g_hill[k] = a.g_cap[k] = a.g_invcap[k] = 0.0f;
    }
    const uint32_t nseg = (steps + seg - 1u) / seg;
    for (uint32_t g = nseg; g-- > 0u;) {
        const uint32_t s0 = g * seg, s1 = s0 + seg < steps ? s0 + seg : steps;
        io.ckpt_get(g, b);
        for (uint32_t s = s0; s < s1; s++) {
            demands(w, b, static_cast<float>(s) * w.dt);
            for (uint32_t sub = 0; sub < substeps; sub++) {
                io.put((s - s0) * substeps + sub, b);
                substep<MN, MM, false>(w, b, nullptr);
            }
        }
        for (uint32_t s = s1; s-- > s0;) {
            demands(w, b, static_cast<float>(s) * w.dt);
            for (uint32_t sub = substeps; sub-- > 0u;) {
                const uint32_t slot = (s - s0) * substeps + sub;
                Tape<MN> t;
                io.get(slot, b);
                substep<MN, MM, true>(w, b, &t);
                io.get(slot, b);
                substep_back(w, b, t, a);
            }
            demands_back(w, b, static_cast<float>(s) * w.dt, a);
        }
    }
    return fitness;
}

// The gradient in parameter order (6 per muscle, then 1 per node), `stride`
// values per creature, zeros past the creature's own parameters.
template <int MN, int MM>
EVO_HD void store_gradient(const Body<MN, MM>& b, const Adj<MN, MM>& a, float* out, uint32_t stride) {
    for (uint32_t q = 0; q < stride; q++) { out[q] = 0.0f; }
    for (uint32_t k = 0; k < b.muscles; k++) {
        out[6u * k + 0u] = a.g_amp[k];
        out[6u * k + 1u] = a.g_invp[k];
        out[6u * k + 2u] = a.g_phase[k];
        out[6u * k + 3u] = a.g_hill[k];
        out[6u * k + 4u] = a.g_cap[k];
        out[6u * k + 5u] = a.g_invcap[k];
    }
    for (uint32_t i = 0; i < b.nodes; i++) { out[6u * b.muscles + i] = a.g_fric[i]; }
}

}  // namespace adj
}  // namespace evo
