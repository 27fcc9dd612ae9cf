/* C ABI of the MI355X creature simulator.
 *
 * This library replaces the reference's CUDA path for one job: simulating
 * packed creatures to the end of their trials and returning one evo_result
 * each. It stands where src/cuda_engine.rs (driver + NVRTC + launch), the
 * kernel source assembly of src/kernel.rs:268-332 and the `Device` side of
 * src/engine.rs:430-501 stand. Kernels are compiled ahead of time for gfx950;
 * there is no run-time compilation.
 *
 * Conventions: every function returns EVO_OK (0) or a negative EVO_ERR_*
 * code; pointers marked `d_` are device memory, all others host memory.
 * Layouts are those of evo_types.h.
 */
#ifndef EVO_ENGINE_H
#define EVO_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#include "evo_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define EVO_OK 0
#define EVO_ERR_INVALID (-1) /* a malformed batch or argument */
#define EVO_ERR_OOM (-2)     /* device memory ran out; retry later (src/engine.rs:534-650) */
#define EVO_ERR_DEVICE (-3)  /* any other device error; reopen the engine */
#define EVO_ERR_BUSY (-4)    /* no free slot for a submission */
#define EVO_ERR_NOT_READY (-5)

/* ---------------------------------------------------------------- worlds */

/* The settings of src/config.rs:15-115 that a trial reads. `terrain` is the
 * bump height in meters (physics::terrain_amplitude of the level); the other
 * effects carry their table values (src/environment.rs:120-160). `rate` is
 * steps per second (60 standard, 120 fine) and `substeps` the substeps per
 * step (the reference fixes 16, src/kernel.rs:43). A screen applies when
 * screen_seconds > 0. */
typedef struct evo_world_config {
    float duration;
    float gravity;
    float air_retention;
    float ground_friction;
    int32_t ground;
    float terrain;
    float muscle_energy;
    float muscle_recovery;
    float slope;
    float wind;
    float mud;
    float brambles;
    float water;
    float patches;
    float gaps;
    float hurdles;
    float quake;
    uint32_t rate;
    uint32_t substeps;
    float screen_seconds;
    float screen_bar;
    float screen_bar_young;
    float screen_bar_reshaped;
    evo_rung r1;
    evo_rung r2;
} evo_world_config;

/* The calm default world (src/config.rs:118-162) at 60 Hz x 16 substeps. */
void evo_world_default(evo_world_config* c);
/* Effect bits of a world. Replaces src/kernel.rs:82-106 (`world_flags`). */
uint32_t evo_world_flags(const evo_world_config* c);
/* Launch parameters of creatures [base, base + count) in world `c`, with
 * `stride` float pairs per recorded frame. Replaces src/kernel.rs:212-251
 * (`params`); run->record is left 0. */
int evo_world_params(const evo_world_config* c, uint32_t base, uint32_t count, uint32_t stride, evo_params* p,
                     evo_run* run);
/* Frames of the start pose before a recording's first step
 * (src/physics.rs:82-84, `Fidelity::settle`). */
uint32_t evo_settle_frames(uint32_t rate);
/* src/physics.rs:621-631 (`quake_hash`). */
uint32_t evo_quake_hash(uint64_t id);
/* Ground height under x for a creature with bump amplitude and phase
 * (src/physics.rs:660-676, `ground`); -inf without ground. */
float evo_ground_height(const evo_world_config* c, float x, float amplitude, float phase);

/* ------------------------------------------------------ synthetic batches */

/* A seeded population of random bodies written directly as records, for the
 * benchmark workloads (no reference counterpart: the reference packs evolved
 * genomes, src/kernel.rs:471-593). Node counts are uniform on
 * [min_nodes, max_nodes], or min_nodes + Geometric(node_p) clipped to
 * max_nodes when node_p > 0. Muscle counts are uniform on
 * [min_muscles, max_muscles], cut so that bones + muscles <= max_constraints
 * when that is nonzero. Muscle lengths are factors of the anchors' start
 * distance. Every draw comes from a counter-based stream keyed by
 * (seed, creature), so any creature can be regenerated alone. */
typedef struct evo_synth_spec {
    uint64_t seed;
    uint32_t count;
    uint32_t min_nodes, max_nodes;
    float node_p;
    uint32_t min_muscles, max_muscles;
    uint32_t max_constraints;
    float mass, radius;
    float friction_lo, friction_hi;
    float bone_lo, bone_hi;
    float short_lo, short_hi;
    float long_lo, long_hi;
    float period_lo, period_hi;
    float strength_lo, strength_hi;
    float duty;
    float stiffness;
    float joint_lo, joint_hi; /* half ranges of the joints, radians */
    uint32_t sorted;          /* pack sorted by (muscles, nodes, index) */
    uint32_t id_base;         /* creature c has id id_base + c (quake hash) */
    /* Ground the start pose stands on (world of creature c: worlds[world_of[c]]
     * when world_of is given to evo_synth_fill, else `world`). */
    evo_world_config world;
} evo_synth_spec;

/* Fills `s` with the spec of a benchmark workload: 0 = 4,096 creatures of
 * 3-6 nodes and 2-8 muscles from seed 1; 1 = 3-12 nodes with bones + muscles
 * <= 24 from seed 2; 2 = 3-32 nodes (geometric, p = 0.25) with up to 64
 * constraints; 3 = 4-10 nodes with 6-16 muscles. `count` is then set by the
 * caller. */
int evo_synth_preset(int preset, evo_synth_spec* s);
/* Words of node and bone records and floats of muscle records the batch needs. */
int evo_synth_sizes(const evo_synth_spec* s, uint64_t* record_words, uint64_t* muscle_floats);
/* Writes the batch: records, muscles, heads (8 words per creature) and, when
 * not NULL, slots[c] = the generated index of packed creature c. `worlds` and
 * `world_of` (both may be NULL) give each creature the world its start pose
 * stands on. */
int evo_synth_fill(const evo_synth_spec* s, uint32_t* records, float* muscles, uint32_t* heads, uint32_t* slots,
                   const evo_world_config* worlds, const uint32_t* world_of);

/* ---------------------------------------------------- device-resident path */

/* Simulates creatures [p->base, p->base + p->count) of a batch that already
 * lives in device memory, on `stream` (a hipStream_t, or NULL for the null
 * stream), and writes d_results[cidx]. Replaces the launch of `advance`
 * (shaders/creature.cu:1266-1277, src/cuda_engine.rs:2171-2311).
 * d_counter is one device word of scratch that the call clears itself.
 * With run->record, d_frames receives every frame of every creature:
 * creature c's frames start at (c - base) * (settle + steps + 1) * p->stride
 * float pairs (frame layout src/creature_kernel.rs:312-326). The call is
 * asynchronous; it validates nothing on the device, so the batch must have
 * passed evo_validate_batch. */
int evo_advance_device(const uint32_t* d_records, const float* d_muscles, const uint32_t* d_heads,
                       evo_result* d_results, uint32_t* d_counter, const evo_params* p, const evo_run* run,
                       float* d_frames, void* stream);

/* Archive metrics and niche cells of creatures [base, base + count) from
 * their results, on the device: scheduler::to_metrics and trial_steps
 * (src/scheduler.rs:859-923), then qd::descriptor and Descriptor::niche_in
 * (src/qd.rs:762-809, :837-847), written to d_out[c - base]. `steps` is the
 * trial's full step count (Config::steps), `rate` its steps per second and
 * `layout` EVO_LAYOUT_ISLAND or EVO_LAYOUT_GLOBAL. The start pose's aspect
 * ratio is read from the node records' start positions. Asynchronous on
 * `stream`. */
int evo_describe_device(const uint32_t* d_records, const uint32_t* d_heads, const evo_result* d_results,
                        uint32_t base, uint32_t count, uint32_t steps, uint32_t rate, uint32_t layout,
                        evo_metrics* d_out, void* stream);

/* Host check of a packed batch: counts within the simulator's limits
 * (2-32 nodes, 0-96 muscles), every node index in range, bone pivots before
 * their child, parents before their bone, offsets inside the buffers.
 * Returns EVO_OK or EVO_ERR_INVALID with *bad = the first bad creature. */
int evo_validate_batch(const uint32_t* records, uint64_t record_words, const float* muscles, uint64_t muscle_floats,
                       const uint32_t* heads, uint32_t creatures, uint32_t* bad);

/* Largest muscle count of creatures [base, base + count) (for frame strides,
 * src/creature_kernel.rs:323-326). */
uint32_t evo_max_muscles(const uint32_t* heads, uint32_t base, uint32_t count);

/* Library build information: a static string naming the kernels compiled in. */
const char* evo_build_info(void);

#ifdef __cplusplus
}
#endif

#endif /* EVO_ENGINE_H */
