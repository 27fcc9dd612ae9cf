// A recording decoded for viewing: engine::recorded (src/engine.rs:1062-1127)
// and engine::thin (src/engine.rs:127-152) over the frame layout of
// src/creature_kernel.rs:312-326 (positions of the node slots, then an
// (energy, force) pair per muscle, then a (normal, friction) pair per node
// slot after the muscles, then the broken-joint bits).
#include <cstdint>
#include <cstring>

#include "host_common.hpp"

extern "C" {

int64_t evo_decode_recording(const float* frames, uint64_t frame_floats, uint32_t stride, uint32_t count,
                             const uint32_t* order, uint32_t nodes, uint32_t muscles, uint32_t every,
                             float* positions, float* energy, float* pull, float* ground, float* friction,
                             uint64_t* broken) {
    if (!frames || !order || nodes == 0u || nodes > EVO_MAX_NODES || count == 0u) { return EVO_ERR_INVALID; }
    if (stride < 2u * EVO_MAX_NODES + muscles + 1u) { return EVO_ERR_INVALID; }
    if (static_cast<uint64_t>(count) * stride * 2u > frame_floats) { return EVO_ERR_INVALID; }
    for (uint32_t k = 0; k < nodes; k++) {
        if (order[k] >= nodes) { return EVO_ERR_INVALID; }
    }
    const uint32_t step = every > 1u ? every : 1u;
    int64_t kept = 0;
    for (uint32_t t = 0; t < count; t++) {
        if (t % step != 0u && t + 1u != count) { continue; }
        const float* f = frames + static_cast<uint64_t>(t) * stride * 2u;
        for (uint32_t k = 0; k < nodes; k++) {
            const uint32_t node = order[k];
            if (positions) {
                positions[(kept * nodes + node) * 2u] = f[2u * k];
                positions[(kept * nodes + node) * 2u + 1u] = f[2u * k + 1u];
            }
            const float* nf = f + 2u * (EVO_MAX_NODES + muscles + k);
            if (ground) { ground[kept * nodes + node] = nf[0]; }
            if (friction) { friction[kept * nodes + node] = nf[1]; }
        }
        for (uint32_t k = 0; k < muscles; k++) {
            const float* mf = f + 2u * (EVO_MAX_NODES + k);
            if (energy) { energy[kept * muscles + k] = mf[0]; }
            if (pull) { pull[kept * muscles + k] = mf[1]; }
        }
        if (broken) {
            const float* b = f + 2u * (stride - 1u);
            broken[kept] = static_cast<uint64_t>(evo::as_uint(b[0])) | (static_cast<uint64_t>(evo::as_uint(b[1])) << 32u);
        }
        kept++;
    }
    return kept;
}

}  // extern "C"
