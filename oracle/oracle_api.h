/* ORACLE — test infrastructure only (see oracle_math.h).
 *
 * The oracle's C entry points. `oracle_advance_<T>` runs creatures
 * [p->base, p->base + p->count) of a packed batch (the layouts of
 * include/evo_types.h) to the end of their trials on `threads` host threads
 * and writes results[cidx]. With run->record set it also writes every frame
 * of every creature: creature c's frames start at
 * (c - p->base) * frames_per_creature * p->stride float pairs, with
 * frames_per_creature = settle + steps + 1. `steps_out`, when not NULL,
 * receives the total steps simulated. Returns 0, or the negated index + 1 of
 * the first creature whose record the simulator cannot run.
 */
#ifndef ORACLE_API_H
#define ORACLE_API_H
#include <pthread.h>
#include <stdint.h>
#include "../include/evo_types.h"

#define ORACLE_DEFINE_RUNNER(T, FRAME)                                                                      \
    typedef struct {                                                                                          \
        const T##_World* w;                                                                                   \
        const uint32_t* records;                                                                              \
        const float* muscles;                                                                                 \
        const uint32_t* heads;                                                                                \
        const evo_params* p;                                                                                  \
        evo_result* results;                                                                                  \
        FRAME* frames;                                                                                        \
        uint64_t frames_per_creature;                                                                         \
        volatile uint32_t next;                                                                               \
        volatile long steps;                                                                                  \
        volatile long bad;                                                                                    \
    } T##_Job;                                                                                                \
    static void* T##_worker(void* arg) {                                                                      \
        T##_Job* job = (T##_Job*)arg;                                                                         \
        long local = 0;                                                                                       \
        for (;;) {                                                                                            \
            const uint32_t start = __atomic_fetch_add(&job->next, 16u, __ATOMIC_RELAXED);                     \
            if (start >= job->p->count) { break; }                                                            \
            const uint32_t end = start + 16u < job->p->count ? start + 16u : job->p->count;                   \
            for (uint32_t c = start; c < end; c++) {                                                          \
                const uint32_t cidx = job->p->base + c;                                                       \
                FRAME* frames = job->frames ? job->frames + 2u * job->frames_per_creature * job->p->stride * c \
                                            : NULL;                                                           \
                const long s = T##_run_creature(job->w, cidx, job->records, job->muscles, job->heads,         \
                                                job->results + cidx, frames);                                 \
                if (s < 0) {                                                                                  \
                    long want = (long)cidx + 1;                                                               \
                    long seen = job->bad;                                                                     \
                    while ((seen == 0 || want < seen) &&                                                      \
                           !__atomic_compare_exchange_n(&job->bad, &seen, want, 0, __ATOMIC_RELAXED,          \
                                                        __ATOMIC_RELAXED)) {}                                 \
                } else {                                                                                      \
                    local += s;                                                                               \
                }                                                                                             \
            }                                                                                                 \
        }                                                                                                     \
        __atomic_fetch_add(&job->steps, local, __ATOMIC_RELAXED);                                             \
        return NULL;                                                                                          \
    }                                                                                                         \
    int oracle_advance_##T(const uint32_t* records, const float* muscles, const uint32_t* heads,              \
                           const evo_params* p, const evo_run* run, evo_result* results, FRAME* frames,       \
                           int threads, long* steps_out) {                                                    \
        T##_World w;                                                                                          \
        T##_make_world(&w, p, run);                                                                           \
        T##_Job job;                                                                                          \
        job.w = &w;                                                                                           \
        job.records = records;                                                                                \
        job.muscles = muscles;                                                                                \
        job.heads = heads;                                                                                    \
        job.p = p;                                                                                            \
        job.results = results;                                                                                \
        job.frames = run->record ? frames : NULL;                                                             \
        job.frames_per_creature = (uint64_t)w.settle + p->steps + 1u;                                         \
        job.next = 0u;                                                                                        \
        job.steps = 0;                                                                                        \
        job.bad = 0;                                                                                          \
        if (threads < 1) { threads = 1; }                                                                     \
        if (threads > 256) { threads = 256; }                                                                 \
        pthread_t tids[256];                                                                                  \
        int started = 0;                                                                                      \
        for (int t = 1; t < threads; t++) {                                                                   \
            if (pthread_create(&tids[started], NULL, T##_worker, &job) == 0) { started++; }                   \
        }                                                                                                     \
        T##_worker(&job);                                                                                     \
        for (int t = 0; t < started; t++) { pthread_join(tids[t], NULL); }                                    \
        if (steps_out) { *steps_out = job.steps; }                                                            \
        return job.bad ? -(int)job.bad : 0;                                                                   \
    }

#endif
