// The worlds of the wild islands.
//
// `evo_wild_levels` and `evo_wild_world` restate src/environment.rs:479-536:
// 100 fixed mixes of one to three effects, drawn from the reference's
// counter-based stream (src/evolution.rs:455-564) with the fixed seed "wild",
// each effect at a level that is neither calm nor its harshest. The effect
// tables are src/environment.rs:120-160 and :216-383; `terrain` carries the
// bump height of the ground level (TERRAIN_AMPLITUDES, src/physics.rs:523).
//
// `evo_island_worlds` draws the per-island worlds of the batched benchmark:
// ground bumps {0, 3, 8, 15, 25} cm, gravity {1, 1.5, 2, 3} g, grip
// U[0.05, 1.2], water {0, 12, 35, 90} cm and a headwind of
// {0, -1, -3, -6} m/s^2 (the reference's wind table), from the same stream.
#include <algorithm>
#include <cstdint>
#include <set>
#include <vector>

#include "host_common.hpp"

namespace {

// src/evolution.rs:455-540: the splitmix64 finalizer, keyed draws, stream keys.
uint64_t finalize(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

uint64_t keyed(uint64_t key, uint32_t gene, uint32_t draw) {
    const uint64_t counter = (static_cast<uint64_t>(gene) << 32) | draw;
    return finalize(key + counter * 0x9e3779b97f4a7c15ull);
}

uint64_t stream_key(uint64_t seed, uint64_t generation, uint64_t round, uint64_t slot) {
    uint64_t key = finalize(seed + 0x632be59bd9b4e019ull);
    for (const uint64_t word : {generation, round, slot}) { key = finalize(key ^ (word * 0xd1342543de82ef95ull)); }
    return key;
}

// The cursor of `Rng` (src/evolution.rs:505-550).
struct Rng {
    uint64_t key;
    uint32_t counter = 0;
    Rng(uint64_t seed, uint32_t generation, uint64_t index) : key(stream_key(seed, generation, 0, index)) {}
    uint64_t next_u64() { return keyed(key, 0xffffffffu, counter++); }
    float unit() { return static_cast<float>(next_u64() >> 40) / 16777216.0f; }
    float range(float a, float b) { return a + unit() * (b - a); }
    uint32_t index(uint32_t n) { return static_cast<uint32_t>(next_u64() % n); }
};

// The effects after the autochange entry, in EFFECTS order: level count,
// calm level and the setting of each level.
constexpr float kTerrain[5] = {0.0f, 0.03f, 0.08f, 0.15f, 0.25f};
constexpr float kGravity[4] = {9.8f, 14.7f, 19.6f, 29.4f};
constexpr float kAir[4] = {1.0f, 0.995f, 0.985f, 0.96f};
constexpr float kGrip[5] = {3.0f, 1.5f, 1.0f, 0.6f, 0.3f};
constexpr float kHeat[4] = {1.0f, 0.7f, 0.5f, 0.35f};
constexpr float kDrought[4] = {1.0f, 0.6f, 0.3f, 0.1f};
constexpr float kSlope[5] = {0.0f, 0.03f, 0.08f, 0.15f, 0.25f};
constexpr float kWind[4] = {0.0f, -1.0f, -3.0f, -6.0f};
constexpr float kMud[4] = {0.0f, 0.02f, 0.05f, 0.10f};
constexpr float kBrambles[4] = {0.0f, 3.0f, 8.0f, 20.0f};
constexpr float kWater[4] = {0.0f, 0.12f, 0.35f, 0.9f};
constexpr float kPatches[4] = {0.0f, 0.5f, 0.8f, 0.95f};
constexpr float kGaps[4] = {0.0f, 0.35f, 0.8f, 1.5f};
constexpr float kHurdles[4] = {0.0f, 0.08f, 0.20f, 0.35f};
constexpr float kQuake[4] = {0.0f, 0.05f, 0.12f, 0.25f};

struct Effect {
    uint32_t levels, calm;
    const float* table;
    float evo_world_config::*field;
};

// Index 0 is the autochange level, which sets no physics.
const Effect kEffects[EVO_EFFECTS] = {
    {4, 0, nullptr, nullptr},
    {5, 0, kTerrain, &evo_world_config::terrain},
    {4, 0, kGravity, &evo_world_config::gravity},
    {4, 0, kAir, &evo_world_config::air_retention},
    {5, 1, kGrip, &evo_world_config::ground_friction},
    {4, 0, kHeat, &evo_world_config::muscle_energy},
    {4, 0, kDrought, &evo_world_config::muscle_recovery},
    {5, 0, kSlope, &evo_world_config::slope},
    {4, 0, kWind, &evo_world_config::wind},
    {4, 0, kMud, &evo_world_config::mud},
    {4, 0, kBrambles, &evo_world_config::brambles},
    {4, 0, kWater, &evo_world_config::water},
    {4, 0, kPatches, &evo_world_config::patches},
    {4, 0, kGaps, &evo_world_config::gaps},
    {4, 0, kHurdles, &evo_world_config::hurdles},
    {4, 0, kQuake, &evo_world_config::quake},
};

void set_level(evo_world_config& c, uint32_t effect, uint32_t level) {
    const Effect& e = kEffects[effect];
    if (!e.table) { return; }
    c.*e.field = e.table[level < e.levels ? level : e.levels - 1u];
}

}  // namespace

extern "C" {

int evo_wild_levels(uint32_t n, evo_wild_mix* out) {
    if (!out || n == 0u || n > EVO_WILD_ISLANDS) { return EVO_ERR_INVALID; }
    // src/environment.rs:490-521, drawn in full so the first n equal the
    // first n of the reference's 100
    Rng rng(0x77696c64ull, 0u, 0u);
    std::vector<uint32_t> choices;
    for (uint32_t e = 1; e < EVO_EFFECTS; e++) { choices.push_back(e); }
    std::set<std::vector<uint32_t>> seen;
    uint32_t made = 0;
    while (made < EVO_WILD_ISLANDS) {
        const uint32_t count = 1u + rng.index(3u);
        std::vector<std::pair<uint32_t, uint32_t>> mix;
        while (mix.size() < count) {
            const uint32_t e = choices[rng.index(static_cast<uint32_t>(choices.size()))];
            bool dup = false;
            for (const auto& m : mix) { dup = dup || m.first == e; }
            if (dup) { continue; }
            std::vector<uint32_t> levels;
            for (uint32_t l = 0; l + 1u < kEffects[e].levels; l++) {
                if (l != kEffects[e].calm) { levels.push_back(l); }
            }
            if (levels.empty()) { continue; }
            mix.push_back({e, levels[rng.index(static_cast<uint32_t>(levels.size()))]});
        }
        std::sort(mix.begin(), mix.end());
        std::vector<uint32_t> key;
        for (const auto& m : mix) {
            key.push_back(m.first);
            key.push_back(m.second);
        }
        if (!seen.insert(key).second) { continue; }
        if (made < n) {
            evo_wild_mix& w = out[made];
            w.count = static_cast<uint32_t>(mix.size());
            for (uint32_t i = 0; i < 3u; i++) {
                w.effect[i] = i < w.count ? mix[i].first : 0u;
                w.level[i] = i < w.count ? mix[i].second : 0u;
            }
        }
        made++;
    }
    return EVO_OK;
}

int evo_wild_world(const evo_world_config* base, const evo_wild_mix* mix, evo_world_config* out) {
    if (!base || !mix || !out || mix->count > 3u) { return EVO_ERR_INVALID; }
    // src/environment.rs:529-542: every effect calm, then the island's mix;
    // no screen and no rungs
    evo_world_config c = *base;
    for (uint32_t e = 1; e < EVO_EFFECTS; e++) { set_level(c, e, kEffects[e].calm); }
    for (uint32_t i = 0; i < mix->count; i++) {
        if (mix->effect[i] == 0u || mix->effect[i] >= EVO_EFFECTS) { return EVO_ERR_INVALID; }
        set_level(c, mix->effect[i], mix->level[i]);
    }
    evo_world_config calm;
    evo_world_default(&calm);
    c.screen_seconds = 0.0f;
    c.screen_bar = calm.screen_bar;
    c.screen_bar_young = calm.screen_bar_young;
    c.screen_bar_reshaped = calm.screen_bar_reshaped;
    c.r1 = calm.r1;
    c.r2 = calm.r2;
    *out = c;
    return EVO_OK;
}

int evo_island_worlds(uint64_t seed, uint32_t n, const evo_world_config* base, evo_world_config* out) {
    if (!base || !out) { return EVO_ERR_INVALID; }
    for (uint32_t i = 0; i < n; i++) {
        Rng rng(seed, 0u, i);
        evo_world_config c = *base;
        c.terrain = kTerrain[rng.index(5u)];
        c.gravity = kGravity[rng.index(4u)];
        c.ground_friction = rng.range(0.05f, 1.2f);
        c.water = kWater[rng.index(4u)];
        c.wind = kWind[rng.index(4u)];
        out[i] = c;
    }
    return EVO_OK;
}

uint64_t evo_splitmix_finalize(uint64_t z) { return finalize(z); }

}  // extern "C"
