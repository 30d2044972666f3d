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
 * sowalker: SOwalker_scheduler_t's simulated annealing (engine/schedule.hpp:369-635);
 * staging: the learned scheduler's staging, prefetch and zero-copy code with the network
 * left out (an ablation with no reference counterpart): fill order by the walkers a load
 * would release, swap loads and victims by the plain rule, the zero-copy tail at the
 * reference's walker threshold (apps/Node2vec.cu:100-104) */
enum { GWH_SCHED_PLAIN = 0, GWH_SCHED_LEARNED = 1, GWH_SCHED_SOWALKER = 2, GWH_SCHED_STAGING = 3 };

/* What the engine keeps from one walk of a workload to the next (no reference
 * counterpart): the learned scheduler's network parameters and its ring of block-access
 * samples (features at a decision, steps the block served afterwards), and the launch
 * tuner's step-kernel timings.  Hand the same model to successive gwh_run calls and each
 * walk starts from the network the earlier walks trained; with gwh_engine_opts.launch.block
 * == 0 each walk also runs the launch tuner's next untried candidate (workgroup size, waves
 * per CU, kernel variant; the membership search order stays the probe model's), and once
 * every candidate has run, the fastest (least step-kernel time per step).  Timings are kept
 * per workload (cap, block size, walkers, vertices).  NULL in gwh_engine_opts.model starts
 * every walk from an untrained network with the engine's default launch. */
typedef struct gwh_model gwh_model;
typedef struct {
    uint64_t samples;          /* samples recorded so far (all walks) */
    uint64_t walks;            /* walks that trained it */
    uint64_t adam_steps;       /* optimizer steps taken */
    float loss;                /* last training batch loss (MSE on standardized log1p steps) */
    uint32_t has_params;       /* 0 until a walk has stored parameters */
    uint64_t scored;           /* closed samples predicted before the network trained on them
                                  (once it has trained on 256 samples) */
    double net_abs_err;        /* its mean |log1p error| on them ... */
    double mean_abs_err;       /* ... and that of the running mean of earlier labels */
} gwh_model_info;
gw_status gwh_model_create(gwh_model** out);
gw_status gwh_model_destroy(gwh_model* m);
gw_status gwh_model_info_get(const gwh_model* m, gwh_model_info* out);
typedef struct {
    gw_launch launch;          /* a candidate launch (flags: the probe model's search order) */
    double ms_per_gstep;       /* its step-kernel time per 10^9 steps (mean over its walks) */
    uint32_t walks;            /* walks timed with it (0: not tried yet) */
    uint32_t pad;
} gwh_tune_entry;
/* The launch tuner's table for the workload of the model's last walk: up to cap entries
 * into out (may be NULL), their number in *n, the fastest timed one in *chosen. */
gw_status gwh_model_tuning(const gwh_model* m, gwh_tune_entry* out, uint32_t cap, uint32_t* n, uint32_t* chosen);
/* A graph's blocks in the transfer encoding (gw_pack_block), built once on the host
 * cores into pinned memory and reused by every gwh_run over the same beg / csr arrays and
 * block size (no reference counterpart).  Blocks whose encoding is not smaller than their
 * csr stay raw.  With gwh_engine_opts.packed set, block uploads send the packed bytes and
 * the GPU decodes them into the slots: walks are unchanged, fewer bytes cross PCIe. */
typedef struct gwh_packed gwh_packed;
gw_status gwh_pack_graph(const uint64_t* beg, const uint32_t* csr, uint64_t entries, uint64_t block_bytes,
                         gwh_packed** out);
/* raw_bytes: the blocks' beg and csr bytes; packed_bytes: what their uploads send instead */
gw_status gwh_packed_info(const gwh_packed* p, uint64_t* raw_bytes, uint64_t* packed_bytes, uint32_t* nblocks,
                          uint32_t* blocks_packed);
gw_status gwh_packed_destroy(gwh_packed* p);

/* gwh_engine_opts.flags */
enum {
    GWH_NO_SWAP_PREFETCH = 1u << 0, /* learned: once the cache is full, do not keep a slot for
                                       staging the block the plain rule would load next (the
                                       default keeps one: each upload then overlaps the previous
                                       run, for a slot and some mispredicted copies) */
    GWH_MYOPIC_LABEL = 1u << 1,  /* learned: a load's label is the next run's steps only */
    GWH_MODEL_SWAP = 1u << 2,    /* learned: once the cache is full, the network picks the load
                                    among the plain rule's near-best candidates (default: the
                                    plain rule's load and victim; the network ranks the fill
                                    and decides the zero-copy tail either way) */
    GWH_VALUE_VICTIM = 1u << 3,  /* learned, with GWH_MODEL_SWAP: a swap evicts the resident
                                    block whose eviction
                                    gives up the fewest steps (its walkers paired with the
                                    load plus the network's value of keeping it) instead of
                                    the least-paired one */
    GWH_NO_STICKY = 1u << 4,     /* learned / staging: a swap does not prefer the block already
                                    staged among near-best candidates */
    GWH_NO_TUNE = 1u << 5,       /* with a model: keep the engine's default launch (no tuning) */
    GWH_TUNE_FREEZE = 1u << 6,   /* with a model: run the fastest candidate timed so far, trying
                                    no new one (for walks that are measured) */
    GWH_TWO_LANES = 1u << 7      /* packed uploads on two copy lanes (gw_set_copy_lanes; default
                                    one) */
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
    gwh_model* model;          /* learned scheduler: state carried across walks (may be NULL) */
    const gwh_packed* packed;  /* block uploads in the transfer encoding (may be NULL: raw) */
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
    uint64_t tail_decisions;   /* learned / staging: zero-copy chosen over an upload (both kinds) */
    double sched_ms;           /* host time spent scheduling and placing blocks */
    double tune_ms;            /* autotuner time (launch shape and search order) */
    gw_launch tuned;           /* what it picked (block 0: no tuning) */
    uint64_t tail_model;       /* ... of which the network's rule chose */
    uint64_t tail_threshold;   /* ... of which the reference's walker threshold chose */
    uint64_t swap_model;       /* cache-full iterations whose load and victim the network chose */
    uint64_t swap_differs;     /* ... of which differ from the plain rule's choice */
    uint64_t fill_model;       /* free-slot iterations whose load order the network ranked */
    uint64_t model_samples;    /* samples the network had after the walk (all walks) */
    uint64_t sticky_swaps;     /* swaps that took the already staged near-best block */
    uint32_t tune_candidate;   /* the launch tuner's candidate this walk ran (~0u: no tuning) */
    uint32_t copy_lanes;       /* copy lanes of packed uploads (gw_set_copy_lanes) */
    uint64_t h2d_raw_bytes;    /* the uploads' size in the reference's block format (h2d_bytes:
                                  what crossed PCIe, smaller with packed uploads) */
    uint64_t inplace_blocks;   /* packed blocks decoded in their own slot (gw_upload_stats) ... */
    uint64_t ring_blocks;      /* ... through the staging ring ... */
    uint64_t inline_blocks;    /* ... and on the copy stream */
    uint64_t kernel_launches;  /* step-kernel launches: one per iteration, plus a further one
                                  each time an iteration's runnable walkers overflowed the
                                  worklist (gw_set_worklist_cap) */
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
