/* ORACLE — test infrastructure only (see oracle_math.h).
 *
 * CPU restatement of what the archives read of a finished trial:
 *   trial_steps     src/scheduler.rs:859-870
 *   to_metrics      src/scheduler.rs:890-923 (GpuResult::feet,
 *                   src/creature_kernel.rs:97-99)
 *   descriptor      src/qd.rs:762-787
 *   niche_in        src/qd.rs:797-809, classes_of :125-131, bin :843-847,
 *                   class edges :56-84, MOVEMENT_BINS :16
 * in FP32 with Rust's float rules (clamp keeps NaN, max/min drop it, round is
 * half away from zero, float-to-integer casts saturate with NaN -> 0).
 * height_axis (src/qd.rs:837-842) is ln(max(h, 0.15)/0.15) / ln(0.6 * 2 / 0.15)
 * binned in sixths; its bin k is entered at h = 0.15 * 8^(k/6) = 0.15 * 2^(k/2),
 * so the bin is found by comparing h with those edges (rounded to float).
 * The aspect ratio reads the start positions of the node records (words 3
 * and 4), which are the genome's node positions moved as one.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../include/evo_types.h"

static float rs_clamp(float x, float lo, float hi) {
    if (x < lo) { return lo; }
    if (x > hi) { return hi; }
    return x;
}

static uint32_t rs_as_u8(float x) {
    if (isnan(x) || x <= 0.0f) { return 0u; }
    if (x >= 255.0f) { return 255u; }
    return (uint32_t)x;
}

static uint32_t rs_as_u32(float x) {
    if (isnan(x) || x <= 0.0f) { return 0u; }
    if (x >= 4294967296.0f) { return 0xffffffffu; }
    return (uint32_t)x;
}

static int32_t rs_as_i32(float x) {
    if (isnan(x)) { return 0; }
    if (x <= -2147483648.0f) { return INT32_MIN; }
    if (x >= 2147483648.0f) { return INT32_MAX; }
    return (int32_t)x;
}

static uint8_t qd_bin(float value, float low, float high, uint32_t count) {
    const float scaled = (rs_clamp(value, low, high) - low) / (high - low) * (float)count;
    uint32_t b = rs_as_u8(floorf(scaled));
    if (b > count - 1u) { b = count - 1u; }
    return (uint8_t)b;
}

static uint8_t height_bin(float mean_height) {
    uint8_t b = 0;
    for (int k = 1; k <= 5; k++) {
        const float edge = (float)((double)0.15f * pow(2.0, 0.5 * k));
        if (mean_height >= edge) { b = (uint8_t)k; }
    }
    return b;
}

/* Descriptor::niche_in for a descriptor's fields. */
void oracle_niche_in(uint32_t nodes, float ground_contact, float gait_frequency, float aspect_ratio,
                     float mean_height, float feet, uint32_t layout, uint8_t* niche) {
    static const float island_aspect[2] = {1.2f, 2.0f};
    static const uint32_t island_nodes[2] = {9u, 12u};
    static const float global_aspect[3] = {0.9f, 1.4f, 2.5f};
    static const uint32_t global_nodes[3] = {9u, 11u, 14u};
    const float* aspect_edges = layout == EVO_LAYOUT_GLOBAL ? global_aspect : island_aspect;
    const uint32_t* node_edges = layout == EVO_LAYOUT_GLOBAL ? global_nodes : island_nodes;
    const int edges = layout == EVO_LAYOUT_GLOBAL ? 3 : 2;
    uint8_t shape = 0, size = 0;
    for (int k = 0; k < edges; k++) {
        if (aspect_ratio >= aspect_edges[k]) { shape++; }
        if (nodes >= node_edges[k]) { size++; }
    }
    int32_t f = rs_as_i32(roundf(feet));
    if (f < 1) { f = 1; }
    if (f > 5) { f = 5; }
    niche[0] = qd_bin(ground_contact, 0.0f, 1.0f, 6u);
    niche[1] = qd_bin(gait_frequency, 0.0f, 6.0f, 8u);
    niche[2] = shape;
    niche[3] = height_bin(mean_height);
    niche[4] = (uint8_t)(f - 1);
    niche[5] = size;
}

void oracle_describe(const uint32_t* records, const uint32_t* heads, const evo_result* results, uint32_t base,
                     uint32_t count, uint32_t steps_full, uint32_t rate, uint32_t layout, evo_metrics* out) {
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t c = base + i;
        const uint32_t* h0 = heads + 8u * (uint64_t)c;
        const evo_result* r = results + c;
        const uint32_t nodes = h0[0] & 255u;
        evo_metrics* m = out + i;
        memset(m, 0, sizeof(*m));

        const uint32_t full = steps_full > 1u ? steps_full : 1u;
        const float ended = r->fall_time > 0.0f ? r->fall_time : r->screened;
        uint32_t steps = full;
        if (ended > 0.0f) {
            steps = rs_as_u32(roundf(ended * (float)rate));
            if (steps < 1u) { steps = 1u; }
            if (steps > full) { steps = full; }
        }
        uint32_t lift;
        memcpy(&lift, &r->lift_lo, 4);
        m->fitness = r->fitness;
        m->ground_contact = rs_clamp(r->ground_contact / (float)(steps * nodes), 0.0f, 1.0f);
        m->vertical_oscillation = isfinite(r->vertical_oscillation) ? fmaxf(r->vertical_oscillation, 0.0f) : 0.0f;
        m->gait_frequency = isfinite(r->gait_frequency) ? fmaxf(r->gait_frequency, 0.0f) : 0.0f;
        m->mean_height = fmaxf(r->height_sum / (float)steps, 0.0f);
        m->feet = (float)__builtin_popcount(lift);
        m->screen_x = r->screen_x;
        m->steps = steps;
        m->screened = r->screened > 0.0f;
        uint32_t code;
        memcpy(&code, &r->ground_hi, 4);
        m->fell = (code & 16u) != 0u;

        const uint32_t* rec = records + h0[3];
        float lo_x = INFINITY, hi_x = -INFINITY, lo_y = INFINITY, hi_y = -INFINITY;
        for (uint32_t n = 0; n < nodes; n++) {
            float x, y;
            memcpy(&x, rec + n * EVO_NODE_WORDS + 3u, 4);
            memcpy(&y, rec + n * EVO_NODE_WORDS + 4u, 4);
            lo_x = fminf(lo_x, x);
            hi_x = fmaxf(hi_x, x);
            lo_y = fminf(lo_y, y);
            hi_y = fmaxf(hi_y, y);
        }
        m->aspect_ratio = rs_clamp(fmaxf(hi_x - lo_x, 0.01f) / fmaxf(hi_y - lo_y, 0.01f), 0.0625f, 16.0f);
        oracle_niche_in(nodes, rs_clamp(m->ground_contact, 0.0f, 1.0f), rs_clamp(m->gait_frequency, 0.0f, 20.0f),
                        m->aspect_ratio, fmaxf(m->mean_height, 0.0f), fmaxf(m->feet, 0.0f), layout, m->niche);
    }
}
