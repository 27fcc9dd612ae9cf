// The device-resident entry point: picks a compiled kernel instance and
// launches it. The reference compiles its kernel at run time once per world,
// step rate and recording switch (src/cuda_engine.rs:889-947,
// src/kernel.rs:268-332); here every instance is compiled ahead of time for
// gfx950 (instances.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "../../include/evo_engine.h"
#include "launch.hpp"
#include <cstdlib>

namespace {

using evo::dev::Launcher;

Launcher pick(const evo_run& run) {
    if (run.rate == 60u && run.substeps == 8u) { return evo::dev::pick_r60_s8(run); }
    if (run.rate == 60u && run.substeps == 16u) { return evo::dev::pick_r60_s16(run); }
    if (run.rate == 120u && run.substeps == 8u) { return evo::dev::pick_r120_s8(run); }
    if (run.rate == 120u && run.substeps == 16u) { return evo::dev::pick_r120_s16(run); }
    return nullptr;
}

int grid_cap_from_env() {
    static int cap = -2;
    if (cap == -2) {
        const char* v = std::getenv("EVO_GRID_CAP");
        cap = v ? std::atoi(v) : 0;
    }
    return cap;
}

}  // namespace

extern "C" {

int evo_advance_device(const uint32_t* d_records, const float* d_muscles, const uint32_t* d_heads,
                       evo_result* d_results, uint32_t* d_counter, const evo_params* p, const evo_run* run,
                       float* d_frames, void* stream) {
    if (!p || !run || !d_records || !d_muscles || !d_heads || !d_results || !d_counter) { return EVO_ERR_INVALID; }
    if (run->record && !d_frames) { return EVO_ERR_INVALID; }
    if (p->steps == 0u || (run->flags & ~EVO_ALL_FLAGS) != 0u) { return EVO_ERR_INVALID; }
    if (run->max_nodes > EVO_MAX_NODES || run->max_muscles > EVO_MAX_MUSCLES) { return EVO_ERR_INVALID; }
    if (p->count == 0u) { return EVO_OK; }
    const Launcher l = pick(*run);
    if (!l) { return EVO_ERR_INVALID; }
    const uint64_t fpc = static_cast<uint64_t>(evo_settle_frames(run->rate)) + p->steps + 1u;
    const hipError_t e = l(d_records, d_muscles, reinterpret_cast<const uint4*>(d_heads), d_results, d_counter, *p,
                           run->flags, reinterpret_cast<float2*>(d_frames), fpc,
                           reinterpret_cast<hipStream_t>(stream), grid_cap_from_env());
    if (e == hipErrorOutOfMemory) { return EVO_ERR_OOM; }
    return e == hipSuccess ? EVO_OK : EVO_ERR_DEVICE;
}

const char* evo_build_info(void) {
    return "evo-mi355x gfx950: rates {60,120} x substeps {8,16} x worlds {flat, any} x sizes "
           "{12n/8m unrolled, 12n/24m unrolled, 32n/96m looped}; recording: any world, looped";
}

}  // extern "C"
