// The engine: submissions of packed units, completion by polling.
//
// Replaces the device side of the reference's CudaEngine
// (src/cuda_engine.rs:1819-2392) behind the `Device` contract of
// src/engine.rs:430-467: a fixed set of submission slots (standard ones, one
// for confirmation trials, one for replays), each with its own stream,
// events and device buffers that grow as units need and are kept for reuse;
// uploads through pinned staging buffers; results handed back in the
// caller's order through the unit's slot map; out-of-memory reported apart
// from other device errors so the caller can back off and retry. Kernels are
// compiled ahead of time, so there is no compile step and no kernel cache.
//
// One thread uses a handle at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/evo_engine.h"

namespace {

constexpr uint64_t kMinBuffer = 1u << 20;  // buffers start at 1 MiB (src/cuda_engine.rs:650-655)

struct DevBuf {
    void* p = nullptr;
    uint64_t size = 0;
};
struct HostBuf {
    void* p = nullptr;
    uint64_t size = 0;
};

enum Kind : uint32_t { kStandard = EVO_SLOT_STANDARD, kConfirm = EVO_SLOT_CONFIRM, kReplay = 2u };

struct Slot {
    Kind kind = kStandard;
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr, end = nullptr, done = nullptr;
    DevBuf records, muscles, heads, results, counter, frames;
    HostBuf h_records, h_muscles, h_heads, h_results, h_frames;
    bool busy = false;
    uint64_t ticket = 0;
    uint32_t count = 0;
    std::vector<uint32_t> order;  // caller index of each packed creature
    evo_result* out = nullptr;
    float* frames_out = nullptr;
    uint64_t frame_floats = 0, frames_per_creature = 0, stride = 0;
};

}  // namespace

struct evo_engine {
    evo_engine_config cfg{};
    std::vector<Slot> slots;
    uint64_t next_ticket = 1;
    uint64_t allocated = 0;
    bool failed = false;
    std::string error;
};

namespace {

int fail(evo_engine* e, hipError_t err, const char* what) {
    if (err == hipErrorOutOfMemory) {
        e->error = std::string(what) + ": out of device memory";
        return EVO_ERR_OOM;
    }
    e->failed = true;
    e->error = std::string(what) + ": " + hipGetErrorString(err);
    return EVO_ERR_DEVICE;
}

bool within_budget(const evo_engine* e, uint64_t more) {
    return e->cfg.device_budget == 0u || e->allocated + more <= e->cfg.device_budget;
}

// Grows a device buffer to at least `bytes` (by a quarter more, at least
// 1 MiB), keeping it for reuse otherwise.
int ensure(evo_engine* e, DevBuf& b, uint64_t bytes, const char* what) {
    if (b.size >= bytes) { return EVO_OK; }
    const uint64_t want = std::max<uint64_t>(kMinBuffer, bytes + bytes / 4u);
    if (!within_budget(e, want > b.size ? want - b.size : 0u)) {
        e->error = std::string(what) + ": over the engine's device memory budget";
        return EVO_ERR_OOM;
    }
    if (b.p) {
        hipFree(b.p);
        e->allocated -= b.size;
        b = DevBuf{};
    }
    void* p = nullptr;
    const hipError_t err = hipMalloc(&p, want);
    if (err != hipSuccess) {
        hipGetLastError();
        return fail(e, err, what);
    }
    b.p = p;
    b.size = want;
    e->allocated += want;
    return EVO_OK;
}

int ensure_host(evo_engine* e, HostBuf& b, uint64_t bytes, const char* what) {
    if (b.size >= bytes) { return EVO_OK; }
    const uint64_t want = std::max<uint64_t>(kMinBuffer, bytes + bytes / 4u);
    if (b.p) {
        hipHostFree(b.p);
        e->allocated -= b.size;
        b = HostBuf{};
    }
    void* p = nullptr;
    const hipError_t err = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (err != hipSuccess) {
        hipGetLastError();
        return fail(e, err, what);
    }
    b.p = p;
    b.size = want;
    e->allocated += want;
    return EVO_OK;
}

template <class B>
void drop(evo_engine* e, B& b, bool host) {
    if (!b.p) { return; }
    if (host) {
        hipHostFree(b.p);
    } else {
        hipFree(b.p);
    }
    e->allocated -= b.size;
    b = B{};
}

uint64_t release_slot(evo_engine* e, Slot& s) {
    const uint64_t before = e->allocated;
    drop(e, s.records, false);
    drop(e, s.muscles, false);
    drop(e, s.heads, false);
    drop(e, s.results, false);
    drop(e, s.counter, false);
    drop(e, s.frames, false);
    drop(e, s.h_records, true);
    drop(e, s.h_muscles, true);
    drop(e, s.h_heads, true);
    drop(e, s.h_results, true);
    drop(e, s.h_frames, true);
    return before - e->allocated;
}

bool standard_fits(const evo_engine* e, uint32_t index, uint32_t creatures) {
    return creatures <= e->cfg.big_unit || index < e->cfg.big_slots;
}

// The slot a unit of `creatures` creatures of this kind can take now, or -1.
int pick_slot(const evo_engine* e, Kind kind, uint32_t creatures) {
    const uint32_t standard = e->cfg.standard_slots;
    if (kind == kReplay) { return e->slots[standard + 1u].busy ? -1 : static_cast<int>(standard + 1u); }
    if (kind == kConfirm && !e->slots[standard].busy) { return static_cast<int>(standard); }
    // small units prefer the slots big units cannot use
    for (int pass = 0; pass < 2; pass++) {
        for (uint32_t i = 0; i < standard; i++) {
            const bool big_slot = i < e->cfg.big_slots;
            if ((pass == 0) == big_slot && creatures <= e->cfg.big_unit) { continue; }
            if (!e->slots[i].busy && standard_fits(e, i, creatures)) { return static_cast<int>(i); }
        }
    }
    return -1;
}

int submit_to(evo_engine* e, int index, const evo_pack* pack, const evo_world_config* world, bool record,
              evo_result* out, float* frames_out, uint64_t frame_floats, uint64_t* ticket) {
    Slot& s = e->slots[index];
    evo_params p;
    evo_run run;
    const uint32_t stride = record ? 2u * EVO_MAX_NODES + evo_max_muscles(pack->heads, 0u, pack->count) + 1u : 0u;
    if (evo_world_params(world, 0u, pack->count, stride, &p, &run) != EVO_OK) { return EVO_ERR_INVALID; }
    run.record = record ? 1u : 0u;
    run.max_nodes = 0u;
    run.max_muscles = 0u;
    for (uint32_t c = 0; c < pack->count; c++) {
        run.max_nodes = std::max(run.max_nodes, pack->heads[8u * c] & 255u);
        run.max_muscles = std::max(run.max_muscles, pack->heads[8u * c + 1u]);
    }
    bool tendons = false;
    for (uint32_t c = 0; c < pack->count && !tendons; c++) {
        const float* m = pack->muscles + pack->heads[8u * c + 4u];
        for (uint32_t k = 0; k < pack->heads[8u * c + 1u]; k++) { tendons = tendons || m[k * EVO_MUSCLE_WORDS + 6u] != 0.0f; }
    }
    run.no_tendons = tendons ? 0u : 1u;
    const uint64_t fpc = record ? static_cast<uint64_t>(evo_settle_frames(run.rate)) + p.steps + 1u : 0u;
    const uint64_t frames_needed = static_cast<uint64_t>(pack->count) * fpc * stride * 2u;
    if (record && (!frames_out || frame_floats < frames_needed)) { return EVO_ERR_INVALID; }

    const uint64_t rb = pack->record_words * 4u, mb = std::max<uint64_t>(pack->muscle_floats, 4u) * 4u;
    const uint64_t hb = static_cast<uint64_t>(pack->count) * 32u, ob = static_cast<uint64_t>(pack->count) * 76u;
    int rc = EVO_OK;
    if ((rc = ensure(e, s.records, rb, "records")) != EVO_OK) { return rc; }
    if ((rc = ensure(e, s.muscles, mb, "muscles")) != EVO_OK) { return rc; }
    if ((rc = ensure(e, s.heads, hb, "heads")) != EVO_OK) { return rc; }
    if ((rc = ensure(e, s.results, ob, "results")) != EVO_OK) { return rc; }
    if ((rc = ensure(e, s.counter, 256u, "counter")) != EVO_OK) { return rc; }
    if (record && (rc = ensure(e, s.frames, frames_needed * 4u, "frames")) != EVO_OK) { return rc; }
    if ((rc = ensure_host(e, s.h_records, rb, "staging")) != EVO_OK) { return rc; }
    if ((rc = ensure_host(e, s.h_muscles, mb, "staging")) != EVO_OK) { return rc; }
    if ((rc = ensure_host(e, s.h_heads, hb, "staging")) != EVO_OK) { return rc; }
    if ((rc = ensure_host(e, s.h_results, ob, "staging")) != EVO_OK) { return rc; }
    if (record && (rc = ensure_host(e, s.h_frames, frames_needed * 4u, "staging")) != EVO_OK) { return rc; }

    std::memcpy(s.h_records.p, pack->records, rb);
    std::memset(s.h_muscles.p, 0, mb);
    std::memcpy(s.h_muscles.p, pack->muscles, pack->muscle_floats * 4u);
    std::memcpy(s.h_heads.p, pack->heads, hb);
    hipError_t err = hipMemcpyAsync(s.records.p, s.h_records.p, rb, hipMemcpyHostToDevice, s.stream);
    if (err == hipSuccess) { err = hipMemcpyAsync(s.muscles.p, s.h_muscles.p, mb, hipMemcpyHostToDevice, s.stream); }
    if (err == hipSuccess) { err = hipMemcpyAsync(s.heads.p, s.h_heads.p, hb, hipMemcpyHostToDevice, s.stream); }
    if (err == hipSuccess && record) {
        // frame slots a creature does not use stay zero
        err = hipMemsetAsync(s.frames.p, 0, frames_needed * 4u, s.stream);
    }
    if (err == hipSuccess) { err = hipEventRecord(s.start, s.stream); }
    if (err != hipSuccess) {
        hipStreamSynchronize(s.stream);
        return fail(e, err, "upload");
    }
    rc = evo_advance_device(static_cast<const uint32_t*>(s.records.p), static_cast<const float*>(s.muscles.p),
                            static_cast<const uint32_t*>(s.heads.p), static_cast<evo_result*>(s.results.p),
                            static_cast<uint32_t*>(s.counter.p), &p, &run,
                            record ? static_cast<float*>(s.frames.p) : nullptr, s.stream);
    if (rc != EVO_OK) {
        hipStreamSynchronize(s.stream);
        if (rc == EVO_ERR_DEVICE) {
            e->failed = true;
            e->error = "launch failed";
        }
        return rc;
    }
    err = hipEventRecord(s.end, s.stream);
    if (err == hipSuccess) { err = hipMemcpyAsync(s.h_results.p, s.results.p, ob, hipMemcpyDeviceToHost, s.stream); }
    if (err == hipSuccess && record) {
        err = hipMemcpyAsync(s.h_frames.p, s.frames.p, frames_needed * 4u, hipMemcpyDeviceToHost, s.stream);
    }
    if (err == hipSuccess) { err = hipEventRecord(s.done, s.stream); }
    if (err != hipSuccess) {
        hipStreamSynchronize(s.stream);
        return fail(e, err, "download");
    }
    s.busy = true;
    s.ticket = e->next_ticket++;
    s.count = pack->count;
    s.order.assign(pack->count, 0u);
    for (uint32_t c = 0; c < pack->count; c++) { s.order[c] = pack->slots ? pack->slots[c] : c; }
    s.out = out;
    s.frames_out = frames_out;
    s.frame_floats = frames_needed;
    s.frames_per_creature = fpc;
    s.stride = stride;
    *ticket = s.ticket;
    return EVO_OK;
}

bool pack_ok(const evo_pack* pack) {
    if (!pack || !pack->records || !pack->muscles || !pack->heads || pack->count == 0u) { return false; }
    if (pack->slots) {
        std::vector<char> seen(pack->count, 0);
        for (uint32_t c = 0; c < pack->count; c++) {
            if (pack->slots[c] >= pack->count || seen[pack->slots[c]]) { return false; }
            seen[pack->slots[c]] = 1;
        }
    }
    uint32_t bad = 0;
    return evo_validate_batch(pack->records, pack->record_words, pack->muscles, pack->muscle_floats, pack->heads,
                              pack->count, &bad) == EVO_OK;
}

}  // namespace

extern "C" {

void evo_engine_default(evo_engine_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->device = 0;
    c->standard_slots = 24u;  // src/engine.rs:390-396
    c->big_unit = 8192u;      // src/cuda_engine.rs:505-510
    c->big_slots = 4u;
    c->device_budget = 0u;
}

int evo_open(const evo_engine_config* cfg, evo_engine** out) {
    if (!cfg || !out || cfg->standard_slots == 0u || cfg->standard_slots > 64u) { return EVO_ERR_INVALID; }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || cfg->device < 0 || cfg->device >= n) {
        hipGetLastError();
        return EVO_ERR_DEVICE;
    }
    if (hipSetDevice(cfg->device) != hipSuccess) { return EVO_ERR_DEVICE; }
    evo_engine* e = new evo_engine;
    e->cfg = *cfg;
    e->slots.resize(cfg->standard_slots + 2u);
    for (uint32_t i = 0; i < e->slots.size(); i++) {
        Slot& s = e->slots[i];
        s.kind = i < cfg->standard_slots ? kStandard : (i == cfg->standard_slots ? kConfirm : kReplay);
        if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&s.start) != hipSuccess || hipEventCreate(&s.end) != hipSuccess ||
            hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
            evo_close(e);
            return EVO_ERR_DEVICE;
        }
    }
    *out = e;
    return EVO_OK;
}

void evo_close(evo_engine* e) {
    if (!e) { return; }
    hipSetDevice(e->cfg.device);
    for (Slot& s : e->slots) {
        if (s.stream) { hipStreamSynchronize(s.stream); }
        release_slot(e, s);
        if (s.start) { hipEventDestroy(s.start); }
        if (s.end) { hipEventDestroy(s.end); }
        if (s.done) { hipEventDestroy(s.done); }
        if (s.stream) { hipStreamDestroy(s.stream); }
    }
    delete e;
}

uint32_t evo_free_slots(const evo_engine* e) {
    uint32_t n = 0;
    for (uint32_t i = 0; i < e->cfg.standard_slots; i++) { n += e->slots[i].busy ? 0u : 1u; }
    return n;
}

uint32_t evo_free_slots_for(const evo_engine* e, uint32_t creatures) {
    uint32_t n = 0;
    for (uint32_t i = 0; i < e->cfg.standard_slots; i++) {
        n += (!e->slots[i].busy && standard_fits(e, i, creatures)) ? 1u : 0u;
    }
    return n;
}

int evo_confirm_free(const evo_engine* e) {
    return !e->slots[e->cfg.standard_slots].busy || evo_free_slots(e) > 0u;
}

int evo_replay_free(const evo_engine* e) { return !e->slots[e->cfg.standard_slots + 1u].busy; }

int evo_submit(evo_engine* e, const evo_pack* pack, const evo_world_config* world, uint32_t kind,
               evo_result* results_out, uint64_t* ticket) {
    if (!e || !world || !results_out || !ticket || (kind != EVO_SLOT_STANDARD && kind != EVO_SLOT_CONFIRM)) {
        return EVO_ERR_INVALID;
    }
    if (e->failed) { return EVO_ERR_DEVICE; }
    if (!pack_ok(pack)) { return EVO_ERR_INVALID; }
    hipSetDevice(e->cfg.device);
    const int index = pick_slot(e, static_cast<Kind>(kind), pack->count);
    if (index < 0) { return EVO_ERR_BUSY; }
    return submit_to(e, index, pack, world, false, results_out, nullptr, 0u, ticket);
}

uint64_t evo_record_frame_floats(const evo_pack* pack, const evo_world_config* world) {
    if (!pack || !world || !pack->heads) { return 0u; }
    const uint32_t stride = 2u * EVO_MAX_NODES + evo_max_muscles(pack->heads, 0u, pack->count) + 1u;
    evo_params p;
    evo_run run;
    if (evo_world_params(world, 0u, pack->count, stride, &p, &run) != EVO_OK) { return 0u; }
    const uint64_t fpc = static_cast<uint64_t>(evo_settle_frames(run.rate)) + p.steps + 1u;
    return static_cast<uint64_t>(pack->count) * fpc * stride * 2u;
}

int evo_record(evo_engine* e, const evo_pack* pack, const evo_world_config* world, evo_result* results_out,
               float* frames_out, uint64_t frame_floats, uint64_t* ticket) {
    if (!e || !world || !results_out || !frames_out || !ticket) { return EVO_ERR_INVALID; }
    if (e->failed) { return EVO_ERR_DEVICE; }
    if (!pack_ok(pack)) { return EVO_ERR_INVALID; }
    hipSetDevice(e->cfg.device);
    const int index = pick_slot(e, kReplay, pack->count);
    if (index < 0) { return EVO_ERR_BUSY; }
    return submit_to(e, index, pack, world, true, results_out, frames_out, frame_floats, ticket);
}

int evo_poll(evo_engine* e, uint32_t timeout_us, evo_completion* out) {
    if (!e || !out) { return EVO_ERR_INVALID; }
    if (e->failed) { return EVO_ERR_DEVICE; }
    hipSetDevice(e->cfg.device);
    const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(timeout_us);
    for (;;) {
        int best = -1;
        bool any = false;
        for (uint32_t i = 0; i < e->slots.size(); i++) {
            Slot& s = e->slots[i];
            if (!s.busy) { continue; }
            any = true;
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) { continue; }
            if (q != hipSuccess) { return fail(e, q, "poll"); }
            if (best < 0 || s.ticket < e->slots[best].ticket) { best = static_cast<int>(i); }
        }
        if (best >= 0) {
            Slot& s = e->slots[best];
            const evo_result* staged = static_cast<const evo_result*>(s.h_results.p);
            for (uint32_t c = 0; c < s.count; c++) { s.out[s.order[c]] = staged[c]; }
            if (s.frames_out) {
                const float* f = static_cast<const float*>(s.h_frames.p);
                const uint64_t per = s.frames_per_creature * s.stride * 2u;
                for (uint32_t c = 0; c < s.count; c++) {
                    std::memcpy(s.frames_out + s.order[c] * per, f + c * per, per * 4u);
                }
            }
            float ms = 0.0f;
            hipEventElapsedTime(&ms, s.start, s.end);
            out->ticket = s.ticket;
            out->count = s.count;
            out->recorded = s.frames_out ? 1u : 0u;
            out->frame_floats = s.frames_out ? s.frame_floats : 0u;
            out->gpu_seconds = ms * 1e-3;
            s.busy = false;
            s.out = nullptr;
            s.frames_out = nullptr;
            return EVO_OK;
        }
        if (!any || std::chrono::steady_clock::now() >= until) { return EVO_ERR_NOT_READY; }
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

uint64_t evo_release_idle(evo_engine* e) {
    if (!e) { return 0u; }
    hipSetDevice(e->cfg.device);
    uint64_t freed = 0;
    for (Slot& s : e->slots) {
        if (!s.busy) { freed += release_slot(e, s); }
    }
    return freed;
}

uint64_t evo_allocated_bytes(const evo_engine* e) { return e ? e->allocated : 0u; }

const char* evo_last_error(const evo_engine* e) { return e ? e->error.c_str() : "no engine"; }

}  // extern "C"
