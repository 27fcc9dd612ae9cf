// Adjoint kernel launchers: one set per (steps per second, substeps per
// step), each compiled in its own translation unit (adjoint_instances.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/evo_engine.h"

namespace evo {
namespace dev {
namespace adj {

// Body classes: the largest node and muscle counts each unrolls to. The
// large class is the adjoint's limit (EVO_ADJ_MAX_NODES, EVO_ADJ_MAX_MUSCLES).
constexpr int kSmallNodes = 10, kSmallMuscles = 16;
constexpr int kLargeNodes = EVO_ADJ_MAX_NODES, kLargeMuscles = EVO_ADJ_MAX_MUSCLES;

struct AdjointArgs {
    const uint32_t* records;
    const float* muscles;
    const uint4* heads;
    const evo_result* results;
    evo_params p;
    uint32_t flags, seg, segments;
    float* grads;
    float* fitness;
    float* workspace;
    uint64_t stride;  // lanes per workspace word: the launch's count rounded up to whole waves
};

using AdjointLauncher = hipError_t (*)(const AdjointArgs& x, hipStream_t stream);

// The launcher for a run's world and body sizes.
AdjointLauncher adjoint_pick_r60_s8(const evo_run& run);
AdjointLauncher adjoint_pick_r60_s16(const evo_run& run);
AdjointLauncher adjoint_pick_r120_s8(const evo_run& run);
AdjointLauncher adjoint_pick_r120_s16(const evo_run& run);

}  // namespace adj
}  // namespace dev
}  // namespace evo
