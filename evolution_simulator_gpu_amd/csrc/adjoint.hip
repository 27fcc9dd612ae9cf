// The adjoint launch: d(fitness)/d(parameters) of every creature of a batch,
// one creature per lane, from the functions of adjoint.hpp.
//
// Memory: every `seg` steps a checkpoint of the state (positions,
// velocities, energy stores) goes to the workspace; going backwards, a
// segment is recomputed from its checkpoint with the start state of each of
// its substeps written to the workspace, then reversed substep by substep.
// Workspace slots are laid out [slot][word][creature], so the lanes of a
// wave read and write consecutive words.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/evo_engine.h"
#include "adjoint.hpp"

namespace {

using evo::adj::Adj;
using evo::adj::Body;
using evo::adj::World;

constexpr int kN = EVO_ADJ_MAX_NODES;
constexpr int kM = EVO_ADJ_MAX_MUSCLES;
constexpr uint32_t kWords = 4u * kN + kM;

// The workspace of one launch: checkpoints, then the per-substep tape.
struct DeviceIO {
    float* ckpt;
    float* tape;
    uint32_t count, lane;
    __device__ void write(float* base, uint32_t slot, const Body<kN, kM>& b) const {
        float* p = base + static_cast<uint64_t>(slot) * kWords * count + lane;
        for (uint32_t i = 0; i < b.nodes; i++) {
            p[static_cast<uint64_t>(i) * count] = b.px[i];
            p[static_cast<uint64_t>(kN + i) * count] = b.py[i];
            p[static_cast<uint64_t>(2 * kN + i) * count] = b.vx[i];
            p[static_cast<uint64_t>(3 * kN + i) * count] = b.vy[i];
        }
        for (uint32_t k = 0; k < b.muscles; k++) { p[static_cast<uint64_t>(4 * kN + k) * count] = b.store[k]; }
    }
    __device__ void read(const float* base, uint32_t slot, Body<kN, kM>& b) const {
        const float* p = base + static_cast<uint64_t>(slot) * kWords * count + lane;
        for (uint32_t i = 0; i < b.nodes; i++) {
            b.px[i] = p[static_cast<uint64_t>(i) * count];
            b.py[i] = p[static_cast<uint64_t>(kN + i) * count];
            b.vx[i] = p[static_cast<uint64_t>(2 * kN + i) * count];
            b.vy[i] = p[static_cast<uint64_t>(3 * kN + i) * count];
        }
        for (uint32_t k = 0; k < b.muscles; k++) { b.store[k] = p[static_cast<uint64_t>(4 * kN + k) * count]; }
    }
    __device__ void ckpt_put(uint32_t g, const Body<kN, kM>& b) { write(ckpt, g, b); }
    __device__ void ckpt_get(uint32_t g, Body<kN, kM>& b) { read(ckpt, g, b); }
    __device__ void put(uint32_t s, const Body<kN, kM>& b) { write(tape, s, b); }
    __device__ void get(uint32_t s, Body<kN, kM>& b) { read(tape, s, b); }
};

__global__ void __launch_bounds__(64) adjoint_kernel(const uint32_t* __restrict__ records,
                                                     const float* __restrict__ muscles,
                                                     const uint32_t* __restrict__ heads,
                                                     const evo_result* __restrict__ results, const evo_params p,
                                                     uint32_t flags, uint32_t rate, uint32_t substeps, uint32_t seg,
                                                     float* __restrict__ grads, float* __restrict__ fitness,
                                                     float* __restrict__ ckpt, float* __restrict__ tape) {
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane >= p.count) { return; }
    const uint32_t c = p.base + lane;
    const World w = evo::adj::make_world(p, flags, rate, substeps);
    Body<kN, kM> b;
    Adj<kN, kM> a;
    float* out = grads + static_cast<uint64_t>(lane) * EVO_ADJ_STRIDE;
    const uint32_t code = __float_as_uint(results[c].ground_hi);
    // the trial's own length, never past the launch's (the workspace's size)
    const uint32_t steps = min(code >> 16u, p.steps);
    if (!evo::adj::load(b, w, c, records, muscles, heads) || (code & 32u) != 0u || steps == 0u) {
        // a failed trial scores a constant: its gradient is zero
        for (uint32_t q = 0; q < EVO_ADJ_STRIDE; q++) { out[q] = 0.0f; }
        fitness[lane] = results[c].fitness;
        return;
    }
    DeviceIO io{ckpt, tape, p.count, lane};
    fitness[lane] = evo::adj::trial_gradient(w, b, steps, substeps, seg, io, a);
    evo::adj::store_gradient(b, a, out, EVO_ADJ_STRIDE);
}

uint32_t max_segments(uint32_t steps, uint32_t seg) { return (steps + seg - 1u) / seg; }

}  // namespace

extern "C" {

uint64_t evo_adjoint_workspace_bytes(uint32_t count, uint32_t steps, uint32_t substeps, uint32_t seg) {
    if (seg == 0u) { return 0u; }
    const uint64_t slots = static_cast<uint64_t>(max_segments(steps, seg)) + static_cast<uint64_t>(seg) * substeps;
    return slots * kWords * count * sizeof(float);
}

int evo_adjoint_device(const uint32_t* d_records, const float* d_muscles, const uint32_t* d_heads,
                       const evo_result* d_results, const evo_params* p, const evo_run* run, uint32_t seg,
                       float* d_grads, float* d_fitness, void* d_workspace, uint64_t workspace_bytes, void* stream) {
    if (!d_records || !d_muscles || !d_heads || !d_results || !p || !run || !d_grads || !d_fitness || !d_workspace) {
        return EVO_ERR_INVALID;
    }
    if (seg == 0u || p->steps == 0u || run->record || (run->flags & ~evo::adj::kAdjointFlags) != 0u) {
        return EVO_ERR_INVALID;
    }
    if (run->max_nodes == 0u || run->max_nodes > static_cast<uint32_t>(kN) ||
        run->max_muscles > static_cast<uint32_t>(kM)) {
        return EVO_ERR_INVALID;
    }
    if (run->substeps != 8u && run->substeps != 16u) { return EVO_ERR_INVALID; }
    if (workspace_bytes < evo_adjoint_workspace_bytes(p->count, p->steps, run->substeps, seg)) {
        return EVO_ERR_INVALID;
    }
    if (p->count == 0u) { return EVO_OK; }
    float* ckpt = static_cast<float*>(d_workspace);
    float* tape = ckpt + static_cast<uint64_t>(max_segments(p->steps, seg)) * kWords * p->count;
    const uint32_t blocks = (p->count + 63u) / 64u;
    hipLaunchKernelGGL(adjoint_kernel, dim3(blocks), dim3(64), 0, reinterpret_cast<hipStream_t>(stream), d_records,
                       d_muscles, d_heads, d_results, *p, run->flags, run->rate, run->substeps, seg, d_grads,
                       d_fitness, ckpt, tape);
    const hipError_t e = hipGetLastError();
    if (e == hipErrorOutOfMemory) { return EVO_ERR_OOM; }
    return e == hipSuccess ? EVO_OK : EVO_ERR_DEVICE;
}

}  // extern "C"
