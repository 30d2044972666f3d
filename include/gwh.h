/*
 * gwh.h -- C entry to the host-side engine (libgwhost.so): GOWalker's graph_engine
 * (engine/engine.hpp:54-638) driving libgw.so over an in-memory CSR graph, for callers
 * that are not C++ (the reference has no such entry; its apps call the engine directly,
 * apps/Node2vec.cu:173-176).
 */
#ifndef GWH_H
#define GWH_H
#include "gw.h"

#ifdef __cplusplus
extern "C" {
#endif

/* plain: GOwalker_scheduler_t (engine/schedule.hpp:58-365); learned: the MLP scheduler;
 * sowalker: SOwalker_scheduler_t's simulated annealing (engine/schedule.hpp:369-635) */
enum { GWH_SCHED_PLAIN = 0, GWH_SCHED_LEARNED = 1, GWH_SCHED_SOWALKER = 2 };
/* gwh_engine_opts.flags */
enum {
    GWH_NO_SWAP_PREFETCH = 1u << 0, /* learned: once the cache is full, do not keep a slot for
                                       staging the block the plain rule would load next (the
                                       default keeps one: each upload then overlaps the previous
                                       run, for a slot and some mispredicted copies) */
    GWH_MYOPIC_LABEL = 1u << 1   /* learned: a load's label is the next run's steps only */
};

typedef struct {
    uint64_t memory_bytes;     /* HBM budget (slots + walker state); 0 = every block resident */
    uint64_t block_bytes;      /* blocksize in bytes */
    uint64_t zero_threshold;   /* walkers below which the run ends with zero-copy reads */
    uint32_t scheduler;        /* GWH_SCHED_* */
    uint32_t verbose;
    gw_launch launch;
    int32_t device;
    uint32_t flags;            /* GWH_* flags */
    uint32_t prefetch_depth;   /* learned scheduler: staged block copies in flight (0 = 8) */
    uint32_t reserved;
} gwh_engine_opts;

typedef struct {
    uint64_t steps, trials, probes, mix_steps;
    uint64_t h2d_bytes, uploads, iterations, zero_copy_runs, zero_copy_steps;
    double wall_s, kernel_ms, filter_ms, h2d_ms;
    uint32_t nslots, nblocks;
    uint64_t learned_decisions, fallbacks;
    float final_loss;
    uint32_t pad;
    uint64_t staged;           /* blocks prefetched into free slots (their bytes are in h2d_bytes) */
    uint64_t zc_touches;       /* zero-copy reads of host-mapped blocks (about 64 B of PCIe each) */
    uint64_t tail_decisions;   /* learned: zero-copy chosen over an upload */
    double sched_ms;           /* host time spent scheduling and placing blocks */
    double tune_ms;            /* autotuner time (launch shape and search order) */
    gw_launch tuned;           /* what it picked (block 0: no tuning) */
} gwh_report;

/* Runs walkers to completion out of core: beg has `entries` u64 offsets (ids
 * 0..entries-2), csr the u32 targets; both in pinned memory (gw_alloc_pinned) for the
 * zero-copy tail.  hashes (nwalkers) and visits (entries-1) are optional outputs. */
gw_status gwh_run(const uint64_t* beg, const uint32_t* csr, uint64_t entries, const gwh_engine_opts* opts,
                  const gw_walk_params* params, uint64_t nwalkers, uint32_t walkers_per_vertex,
                  const uint32_t* sources, gwh_report* report, uint64_t* hashes, uint32_t* visits);
const char* gwh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
