// Seeded random bodies written directly as packed records, for the benchmark
// workloads. A body is a tree: bone j joins its pivot node (any earlier node)
// to node j + 1, and bone 0, the neck, hangs from the head. The start pose
// follows src/physics2.rs:371-417 (`Model::start`, `kinematics`): absolute
// angles from relative joint angles, the center of mass over x = 0 and the
// lowest surface on the ground. The records are those of src/kernel.rs:336-411
// (`fill_creature`), including the foot rule, and the heads those of
// src/kernel.rs:554-565. Muscle constants follow src/physics2.rs:233-313 with
// a drawn strength in place of the subtree rule.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <thread>
#include <vector>

#include "host_common.hpp"

namespace {

using evo::Stream;

constexpr float kPi = 3.14159265358979f;

struct Counts {
    uint32_t nodes, muscles;
};

Counts draw_counts(const evo_synth_spec& s, uint32_t c) {
    Stream r(s.seed, 2ull * c);
    uint32_t n;
    if (s.node_p > 0.0f) {
        n = s.min_nodes;
        while (n < s.max_nodes && r.unit() >= s.node_p) { n++; }
    } else {
        n = s.min_nodes + r.index(s.max_nodes - s.min_nodes + 1u);
    }
    const uint32_t bones = n - 1u;
    uint32_t hi = s.max_muscles;
    if (s.max_constraints > 0u) { hi = std::min(hi, s.max_constraints > bones ? s.max_constraints - bones : 0u); }
    hi = std::min(hi, EVO_MAX_MUSCLES);
    const uint32_t lo = std::min(s.min_muscles, hi);
    const uint32_t m = lo + r.index(hi - lo + 1u);
    return {n, m};
}

bool spec_ok(const evo_synth_spec& s) {
    return s.min_nodes >= 3u && s.max_nodes <= EVO_MAX_NODES && s.min_nodes <= s.max_nodes &&
           s.min_muscles <= s.max_muscles && s.max_muscles <= EVO_MAX_MUSCLES && s.duty > 0.0f && s.duty < 1.0f &&
           s.mass > 0.0f && s.radius > 0.0f && s.strength_lo > 0.0f;
}

// Writes creature `gen` (its generated index) at packed position `c`.
void fill_one(const evo_synth_spec& s, uint32_t gen, uint32_t* rec, float* mus, uint32_t* heads, uint32_t c,
              uint32_t record_at, uint32_t muscle_at, const evo_world_config& world) {
    const Counts k = draw_counts(s, gen);
    const uint32_t n = k.nodes, bones = n - 1u, m = k.muscles;
    Stream r(s.seed, 2ull * gen + 1ull);
    uint32_t pivot[EVO_MAX_NODES];
    int parent[EVO_MAX_NODES];
    float length[EVO_MAX_NODES], rel[EVO_MAX_NODES], lo[EVO_MAX_NODES], hi[EVO_MAX_NODES];
    float friction[EVO_MAX_NODES];
    for (uint32_t i = 0; i < n; i++) { friction[i] = r.range(s.friction_lo, s.friction_hi); }
    for (uint32_t j = 0; j < bones; j++) {
        pivot[j] = j == 0 ? 0u : r.index(j + 1u);
        // src/physics2.rs:220 — the bone ending at the pivot, else the neck.
        parent[j] = pivot[j] > 0u ? static_cast<int>(pivot[j]) - 1 : (j > 0 ? 0 : -1);
        length[j] = r.range(s.bone_lo, s.bone_hi);
        // The neck points down from the head; other bones hang within a
        // quarter turn of their parent bone's direction.
        rel[j] = j == 0 ? r.range(-5.0f * kPi / 6.0f, -kPi / 6.0f) : r.range(-0.5f * kPi, 0.5f * kPi);
        const float below = r.range(s.joint_lo, s.joint_hi), above = r.range(s.joint_lo, s.joint_hi);
        lo[j] = rel[j] - below;
        hi[j] = rel[j] + above;
    }
    // Start pose (src/physics2.rs:402-417): absolute angles, then positions.
    float th[EVO_MAX_NODES], px[EVO_MAX_NODES], py[EVO_MAX_NODES];
    px[0] = 0.0f;
    py[0] = 0.0f;
    for (uint32_t j = 0; j < bones; j++) {
        th[j] = parent[j] >= 0 ? th[parent[j]] + rel[j] : rel[j];
        px[j + 1] = px[pivot[j]] + length[j] * std::cos(th[j]);
        py[j + 1] = py[pivot[j]] + length[j] * std::sin(th[j]);
    }
    const uint32_t hash = evo::quake_hash(static_cast<uint64_t>(s.id_base) + gen);
    const bool still = world.quake <= 0.0f || !world.ground;
    const float amp = world.terrain + (still ? 0.0f : world.quake * evo::quake_scale(hash));
    const float phase = still ? 0.0f : evo::quake_phase(hash);
    float total = 0.0f, cx = 0.0f;
    for (uint32_t i = 0; i < n; i++) {
        total += s.mass;
        cx += px[i] * s.mass;
    }
    cx /= total;
    float low = INFINITY;
    for (uint32_t i = 0; i < n; i++) {
        const float h = evo::ground_height(world, px[i] - cx, amp, phase);
        low = std::min(low, py[i] - s.radius - std::max(h, -1e6f));
    }
    if (!world.ground) { low = 0.0f; }
    for (uint32_t i = 0; i < n; i++) {
        px[i] -= cx;
        py[i] -= low;
    }
    // Foot rule of src/kernel.rs:340-352.
    uint32_t children[EVO_MAX_NODES] = {0};
    for (uint32_t j = 1; j < bones; j++) {
        if (parent[j] >= 0) { children[parent[j]]++; }
    }
    uint32_t* node_rec = rec + record_at;
    for (uint32_t i = 0; i < n; i++) {
        bool foot = false;
        if (i > 0) {
            const uint32_t j = i - 1;
            foot = children[j] == 0 && j > 0 && parent[j] > 0 && children[parent[j]] == 1;
        }
        uint32_t* q = node_rec + i * EVO_NODE_WORDS;
        q[0] = evo::as_uint(s.mass);
        q[1] = evo::as_uint(s.radius);
        q[2] = evo::as_uint(friction[i]);
        q[3] = evo::as_uint(px[i]);
        q[4] = evo::as_uint(py[i]);
        q[5] = foot ? 1u : 0u;
    }
    uint32_t* bone_rec = node_rec + n * EVO_NODE_WORDS;
    for (uint32_t j = 0; j < bones; j++) {
        uint32_t* q = bone_rec + j * EVO_BONE_WORDS;
        q[0] = pivot[j];
        q[1] = evo::as_uint(length[j]);
        q[2] = (parent[j] >= 0 && j > 0) ? static_cast<uint32_t>(parent[j]) : EVO_NONE;
        q[3] = evo::as_uint(lo[j]);
        q[4] = evo::as_uint(hi[j]);
        q[5] = 0u;
    }
    float first_period = 0.0f;
    for (uint32_t q = 0; q < m; q++) {
        const uint32_t ba = r.index(bones);
        uint32_t bb = r.index(bones - 1u);
        if (bb >= ba) { bb++; }
        const float aa = r.unit(), ab = r.unit();
        const uint32_t a0 = pivot[ba], a1 = ba + 1u, b0 = pivot[bb], b1 = bb + 1u;
        const float ax = px[a0] + (px[a1] - px[a0]) * aa, ay = py[a0] + (py[a1] - py[a0]) * aa;
        const float bx = px[b0] + (px[b1] - px[b0]) * ab, by = py[b0] + (py[b1] - py[b0]) * ab;
        const float span = std::max(std::hypot(bx - ax, by - ay), 0.02f);
        const float shortest = span * r.range(s.short_lo, s.short_hi);
        const float longest = span * r.range(s.long_lo, s.long_hi);
        const float period = r.range(s.period_lo, s.period_hi);
        const float phase_m = r.unit();
        const float strength = r.range(s.strength_lo, s.strength_hi);
        if (q == 0) { first_period = period; }
        const float duty = s.duty;
        // src/physics2.rs:233-262, src/kernel.rs:376-408
        const float amplitude =
            std::min(longest - shortest, 2.0f * 24.0f * period * std::min(duty, 1.0f - duty) / kPi);
        float* f = mus + muscle_at + q * EVO_MUSCLE_WORDS;
        f[0] = evo::as_float(a0 | (a1 << 8) | (b0 << 16) | (b1 << 24));
        f[1] = aa;
        f[2] = ab;
        f[3] = 1.0f / (8.0f * std::max(longest, 0.05f));
        f[4] = 200.0f * strength;
        f[5] = 1.0f / (120.0f * strength);
        f[6] = 0.0f;
        f[7] = std::max(longest, span);
        f[8] = amplitude;
        f[9] = 1.0f / period;
        f[10] = phase_m;
        f[11] = duty;
        f[12] = 1.0f / duty;
        f[13] = 1.0f / (1.0f - duty);
        f[14] = s.stiffness;
        f[15] = 0.0f;
        f[16] = evo::as_float(EVO_NONE);
        f[17] = 0.0f;
        f[18] = 0.0f;
        f[19] = 0.0f;
    }
    uint32_t* h = heads + 8ull * c;
    h[0] = n;
    h[1] = m;
    h[2] = hash;
    h[3] = record_at;
    h[4] = muscle_at;
    h[5] = evo::as_uint(total);
    h[6] = evo::as_uint(1.0f / total);
    h[7] = evo::f32_to_f16(first_period) & 0xffffu;
}

}  // namespace

extern "C" {

int evo_synth_preset(int preset, evo_synth_spec* s) {
    std::memset(s, 0, sizeof(*s));
    s->mass = 1.0f;
    s->radius = 0.1f;
    s->friction_lo = 0.2f;
    s->friction_hi = 1.0f;
    s->bone_lo = 0.3f;
    s->bone_hi = 1.2f;
    s->short_lo = 0.5f;
    s->short_hi = 1.0f;
    s->long_lo = 1.0f;
    s->long_hi = 1.5f;
    s->period_lo = 0.5f;
    s->period_hi = 2.0f;
    s->strength_lo = 0.1f;
    s->strength_hi = 1.0f;
    s->duty = 0.5f;
    s->stiffness = 40.0f;
    s->joint_lo = 0.5f;
    s->joint_hi = 2.0f;
    s->sorted = 1;
    s->id_base = 1;
    evo_world_default(&s->world);
    switch (preset) {
        case 0:
            s->seed = 1; s->count = 4096; s->min_nodes = 3; s->max_nodes = 6; s->min_muscles = 2; s->max_muscles = 8;
            break;
        case 1:
            s->seed = 2; s->count = 3000000; s->min_nodes = 3; s->max_nodes = 12; s->min_muscles = 1;
            s->max_muscles = 24; s->max_constraints = 24;
            break;
        case 2:
            s->seed = 3; s->count = 3000000; s->min_nodes = 3; s->max_nodes = 32; s->node_p = 0.25f;
            s->min_muscles = 1; s->max_muscles = 64; s->max_constraints = 64;
            break;
        case 3:
            s->seed = 4; s->count = 65536; s->min_nodes = 4; s->max_nodes = 10; s->min_muscles = 6; s->max_muscles = 16;
            break;
        case 4:
            // the bodies of preset 2 from their own seed, packed class-major
            s->seed = 5; s->count = 3000000; s->min_nodes = 3; s->max_nodes = 32; s->node_p = 0.25f;
            s->min_muscles = 1; s->max_muscles = 64; s->max_constraints = 64; s->sorted = 2;
            break;
        default:
            return EVO_ERR_INVALID;
    }
    return EVO_OK;
}

int evo_synth_sizes(const evo_synth_spec* s, uint64_t* record_words, uint64_t* muscle_floats) {
    if (!spec_ok(*s)) { return EVO_ERR_INVALID; }
    uint64_t rw = 0, mf = 0;
    for (uint32_t c = 0; c < s->count; c++) {
        const Counts k = draw_counts(*s, c);
        rw += static_cast<uint64_t>(k.nodes) * EVO_NODE_WORDS + static_cast<uint64_t>(k.nodes - 1u) * EVO_BONE_WORDS;
        mf += static_cast<uint64_t>(k.muscles) * EVO_MUSCLE_WORDS;
    }
    *record_words = rw;
    *muscle_floats = mf;
    return EVO_OK;
}

int evo_synth_fill(const evo_synth_spec* s, uint32_t* records, float* muscles, uint32_t* heads, uint32_t* slots,
                   const evo_world_config* worlds, const uint32_t* world_of) {
    if (!spec_ok(*s)) { return EVO_ERR_INVALID; }
    const uint32_t count = s->count;
    std::vector<Counts> counts(count);
    for (uint32_t c = 0; c < count; c++) { counts[c] = draw_counts(*s, c); }
    std::vector<uint32_t> order(count);
    std::iota(order.begin(), order.end(), 0u);
    if (s->sorted) {
        // src/kernel.rs:487-493 — muscles, then nodes, then index; with
        // sorted = 2 the body-size classes come first (at most 12 nodes and
        // 24 muscles; at most 12 nodes; at most 20 nodes; the rest), so that
        // each class is one contiguous range.
        const bool class_major = s->sorted == 2u;
        auto cls = [&](uint32_t c) {
            const Counts& k = counts[c];
            return k.nodes > 20u ? 3 : k.nodes > 12u ? 2 : k.muscles > 24u ? 1 : 0;
        };
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            if (class_major && cls(a) != cls(b)) { return cls(a) < cls(b); }
            if (counts[a].muscles != counts[b].muscles) { return counts[a].muscles < counts[b].muscles; }
            return counts[a].nodes < counts[b].nodes;
        });
    }
    std::vector<uint64_t> rec_at(count), mus_at(count);
    uint64_t rw = 0, mf = 0;
    for (uint32_t c = 0; c < count; c++) {
        const Counts k = counts[order[c]];
        rec_at[c] = rw;
        mus_at[c] = mf;
        rw += static_cast<uint64_t>(k.nodes) * EVO_NODE_WORDS + static_cast<uint64_t>(k.nodes - 1u) * EVO_BONE_WORDS;
        mf += static_cast<uint64_t>(k.muscles) * EVO_MUSCLE_WORDS;
    }
    if (rw > 0xffffffffull || mf > 0xffffffffull) { return EVO_ERR_INVALID; }
    unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            for (uint32_t c = t; c < count; c += threads) {
                const uint32_t gen = order[c];
                const evo_world_config& w = (worlds && world_of) ? worlds[world_of[gen]] : s->world;
                fill_one(*s, gen, records, muscles, heads, c, static_cast<uint32_t>(rec_at[c]),
                         static_cast<uint32_t>(mus_at[c]), w);
            }
        });
    }
    for (auto& th : pool) { th.join(); }
    if (slots) {
        for (uint32_t c = 0; c < count; c++) { slots[c] = order[c]; }
    }
    return EVO_OK;
}

}  // extern "C"
