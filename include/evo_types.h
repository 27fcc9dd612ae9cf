/* Byte layouts that the creature simulator's boundary speaks.
 *
 * Every struct here keeps the layout of its counterpart in the reference, so a
 * batch that the reference's packer fills can be handed over unchanged:
 *   evo_rung    <- src/rungs.rs:83-93            (`Rung`, 32 B)
 *   evo_params  <- src/kernel.rs:145-208         (`Params`, 164 B)
 *                  shaders/creature.cu:75-103
 *   evo_result  <- src/creature_kernel.rs:15-91  (`GpuResult`, 76 B)
 *                  shaders/creature.cu:108-128
 * Record layouts (src/kernel.rs:24-37, :336-411, :554-565):
 *   node record   EVO_NODE_WORDS words: mass, radius, friction, start x,
 *                 start y, foot flag (bit 0)
 *   bone record   EVO_BONE_WORDS words: pivot node, length, parent bone or
 *                 EVO_NONE, joint low, joint high, spare
 *   muscle record EVO_MUSCLE_WORDS floats in five lines of four:
 *                 [nodes a0|a1<<8|b0<<16|b1<<24 (as bits), anchor_a, anchor_b, hill]
 *                 [cap, inv_capacity, tendon_k, slack]
 *                 [amplitude, inv_period, phase, duty]
 *                 [inv_duty, inv_complement, stiffness, reset]
 *                 [sensor node or EVO_NONE (as bits), spare x3]
 *   heads         two uint4 per creature:
 *                 h0 = [nodes, muscles, quake hash, word offset of its records]
 *                 h1 = [float offset of its muscle records, total mass bits,
 *                       inverse total mass bits, fp16 period | flags << 16]
 * A creature's records are its node records followed by its bone records;
 * bone j joins its pivot node to node j + 1, and node 0 is the head.
 */
#ifndef EVO_TYPES_H
#define EVO_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Largest body (node slots of a recorded frame) and most muscles
 * (src/kernel.rs:24-26, src/evolution.rs:271-273). */
#define EVO_MAX_NODES 32u
#define EVO_MAX_MUSCLES 96u
#define EVO_NODE_WORDS 6u
#define EVO_BONE_WORDS 6u
#define EVO_MUSCLE_WORDS 20u
/* No parent bone, no sensor node. */
#define EVO_NONE 0xffffffffu

/* World effect bits, in the order of src/kernel.rs:82-113 (`world_flags`). */
#define EVO_GROUND   (1u << 0)
#define EVO_TERRAIN  (1u << 1)
#define EVO_SLOPE    (1u << 2)
#define EVO_GAPS     (1u << 3)
#define EVO_HURDLES  (1u << 4)
#define EVO_QUAKE    (1u << 5)
#define EVO_MUD      (1u << 6)
#define EVO_WATER    (1u << 7)
#define EVO_ICE      (1u << 8)
#define EVO_WIND     (1u << 9)
#define EVO_AIR      (1u << 10)
#define EVO_BRAMBLES (1u << 11)
#define EVO_ALL_FLAGS 0xfffu

/* Creature flags in the high 16 bits of h1.w (src/rungs.rs:18-35). */
#define EVO_FLAG_AUDIT 1u
#define EVO_FLAG_EXEMPT 2u
#define EVO_FLAG_EXEMPT_R1 4u
#define EVO_FLAG_EXEMPT_R2 8u
#define EVO_FLAG_YOUNG 16u
#define EVO_FLAG_RESHAPED 32u

/* One early rung's rule: stop when the fused multiply-add chain of `w` with
 * the six features, from zero, is below `bias` and cadence band b is not set
 * in `off` (src/rungs.rs:176-221). */
typedef struct evo_rung {
    float w[6];
    float bias;
    uint32_t off;
} evo_rung;

/* Launch parameters: the creatures [base, base + count) of a batch and the
 * world they run in. Field meanings follow src/kernel.rs:145-251. The
 * simulator reads air_sub and inv_muscle_energy, not air and muscle_energy. */
typedef struct evo_params {
    uint32_t count;
    uint32_t base;
    uint32_t steps;
    uint32_t screen_step;
    uint32_t stride;
    float gravity;
    float air;
    float friction;
    float terrain;
    float muscle_energy;
    float muscle_recovery;
    float slope;
    float wind;
    float mud;
    float gaps;
    float hurdles;
    float quake;
    float screen_bar;
    float screen_bar_young;
    float screen_bar_reshaped;
    float water;
    float patches;
    float air_sub;
    float inv_muscle_energy;
    float brambles;
    evo_rung r1;
    evo_rung r2;
} evo_params;

/* One creature's trial result. Several fields carry bit sets or packed fp16
 * words as float bits once the trial has ended (src/creature_kernel.rs:15-218). */
typedef struct evo_result {
    float fitness;
    float ground_contact;
    float vertical_oscillation;
    float gait_frequency;
    float previous_center_y;
    float vertical_extremum;
    float vertical_trend;
    float gait_turns;
    float height_sum;
    float contact_lo;
    float contact_hi;
    float lift_lo;
    float lift_hi;
    float ground_lo;
    float ground_hi;
    float fall_time;
    float head_shake;
    float screen_x;
    float screened;
} evo_result;

/* How a batch runs: its world effect bits, steps per second (60 standard,
 * 120 fine), substeps per step (8 or 16), whether every frame of every
 * creature is recorded, the largest node and muscle counts in the batch
 * (0 when unknown), which pick the kernel's body-size class, and
 * no_tendons = 1 when the caller knows that every muscle's tendon stiffness
 * is 0 (which lets a kernel without tendon arithmetic run it). */
typedef struct evo_run {
    uint32_t flags;
    uint32_t rate;
    uint32_t substeps;
    uint32_t record;
    uint32_t max_nodes;
    uint32_t max_muscles;
    uint32_t no_tendons;
} evo_run;

/* Archive layouts of the niche bytes (src/qd.rs:56-84): ISLAND classes split
 * shapes at aspect 1.2 and 2.0 and sizes at 9 and 12 nodes; GLOBAL classes at
 * aspect 0.9, 1.4, 2.5 and at 9, 11, 14 nodes. */
#define EVO_LAYOUT_ISLAND 0u
#define EVO_LAYOUT_GLOBAL 1u

/* What the archives read of one trial: scheduler::to_metrics
 * (src/scheduler.rs:859-923) and qd::descriptor + Descriptor::niche_in
 * (src/qd.rs:762-809). The floats are the TrialMetrics values; `niche` holds
 * the cell bytes: contact, cadence, shape class, height, feet, size class. */
typedef struct evo_metrics {
    float fitness;
    float ground_contact;
    float vertical_oscillation;
    float gait_frequency;
    float mean_height;
    float feet;
    float screen_x;
    float aspect_ratio;
    uint32_t steps;    /* trial_steps: to the fall, the screen or the end */
    uint8_t niche[6];
    uint8_t screened;  /* 1 when the screen stopped the trial */
    uint8_t fell;      /* 1 when the trial ended in a fall (end code bit 4) */
} evo_metrics;

#ifdef __cplusplus
}
static_assert(sizeof(evo_rung) == 32, "evo_rung layout");
static_assert(sizeof(evo_params) == 164, "evo_params layout");
static_assert(sizeof(evo_result) == 76, "evo_result layout");
static_assert(sizeof(evo_metrics) == 44, "evo_metrics layout");
#else
_Static_assert(sizeof(evo_rung) == 32, "evo_rung layout");
_Static_assert(sizeof(evo_params) == 164, "evo_params layout");
_Static_assert(sizeof(evo_result) == 76, "evo_result layout");
_Static_assert(sizeof(evo_metrics) == 44, "evo_metrics layout");
#endif

#endif /* EVO_TYPES_H */
