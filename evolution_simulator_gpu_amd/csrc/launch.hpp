// Kernel launchers: one set per (steps per second, substeps per step), each
// set compiled in its own translation unit (instances.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/evo_types.h"

namespace evo {
namespace dev {

using Launcher = hipError_t (*)(const uint32_t* records, const float* muscles, const uint4* heads,
                                evo_result* results, uint32_t* counter, const evo_params& p, uint32_t flags,
                                float2* frames, uint64_t frames_per_creature, hipStream_t stream, int grid_cap);

// The launcher for a batch with this run's world, recording switch and body
// sizes, or nullptr when none is compiled in.
Launcher pick_r60_s8(const evo_run& run);
Launcher pick_r60_s16(const evo_run& run);
Launcher pick_r120_s8(const evo_run& run);
Launcher pick_r120_s16(const evo_run& run);

// Body-size classes: the largest node and muscle counts each unrolls to.
// The last class runs its loops at run time and has no such limit below the
// simulator's maximum.
constexpr int kSmallN = 12, kSmallM = 8;
constexpr int kMidN = 12, kMidM = 24;

}  // namespace dev
}  // namespace evo
