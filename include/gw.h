/*
 * gw.h -- C ABI of libgw.so, the MI355X (gfx950) walk engine behind GOWalker's host side.
 *
 * The reference has no library boundary: its host C++ (apps/, engine/, preprocess/)
 * calls the CUDA helpers of gpu/memory.hpp and the kernels of gpu/gpuwalk.hpp directly.
 * Each entry point below replaces one of those calls; the replaced reference interface
 * is cited on every declaration.
 *
 * Conventions
 *   - Every call returns gw_status (GW_OK == 0).  Nothing in the library exits or
 *     throws; gw_last_error() describes the last failure of a context.
 *   - The library owns all device memory and its three streams (compute, copy,
 *     readback -- the reference's update / graph / back streams, api/types.hpp:174-187).
 *   - A context is used by one host thread at a time.
 *   - Host buffers handed to gw_upload_block must stay valid until gw_wait_uploads()
 *     returns; pinned buffers (gw_alloc_pinned) make the copy asynchronous.
 *   - Integer widths follow api/types.hpp:12-20: vertex ids u32, edge offsets u64,
 *     block ids u32.
 */
#ifndef GW_H
#define GW_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t gw_status;
enum {
    GW_OK = 0,
    GW_ERR_ARG = -1,      /* invalid argument */
    GW_ERR_HIP = -2,      /* HIP runtime failure */
    GW_ERR_NOMEM = -3,    /* device or pinned allocation failed */
    GW_ERR_STATE = -4,    /* call out of order (e.g. run before init) */
    GW_ERR_BUDGET = -5,   /* HBM cache budget exceeded */
    GW_ERR_IO = -6,       /* file read/write failure */
    GW_ERR_FAULT = -7     /* failure injected through GW_FAULT (testing) */
};

enum { GW_NODE2VEC = 0, GW_SOPR = 1 };   /* AlgorithmType, api/types.hpp:27 (SOSR = SOPR step) */

typedef struct gw_ctx gw_ctx;

typedef struct {
    int32_t device;          /* HIP device ordinal (apps/Node2vec.cu:24-30 picks 0) */
    uint32_t reserved;
    uint64_t cache_bytes;    /* HBM budget for graph slots + walker state; 0 = unlimited
                                (the `memory` option, apps/Node2vec.cu:39) */
} gw_config;

/* Block descriptor, mirrors gpu_graph_block (api/types.hpp:159-166) without the
 * cache-index/status fields, which the library tracks itself. */
typedef struct {
    uint32_t blk;
    uint32_t start_vert;
    uint32_t nverts;
    uint32_t pad;
    uint64_t start_edge;
    uint64_t nedges;
} gw_block_desc;

/* One walker type.  Node2vec uses p, q; SOPR/SOSR uses alpha.  term is the geometric
 * termination probability per step (0 reproduces the reference's fixed length). */
typedef struct {
    uint32_t algo;           /* GW_NODE2VEC or GW_SOPR */
    uint32_t length;         /* hop cap, the `length` option */
    double p, q, alpha, term;
    uint32_t sopr_sampler;   /* GW_SOPR_PROPOSAL (default; duplicate-free adjacency lists) or
                                GW_SOPR_ENVELOPE (exact on any lists, slower on hubs) */
    uint32_t reserved;
} gw_walk_type;

enum { GW_SOPR_PROPOSAL = 0, GW_SOPR_ENVELOPE = 1 };

typedef struct {
    uint64_t seed;           /* Philox key */
    uint32_t ntypes;         /* 1, or 2 for a mixed batch (walker w has type w % 2) */
    uint32_t reserved;
    gw_walk_type type[2];
} gw_walk_params;

/* Output selection for gw_init_walkers */
enum {
    GW_OUT_VISITS = 1u << 0, /* per-vertex arrival counts (u32) */
    GW_OUT_HASH = 1u << 1,   /* per-walker 64-bit path hash */
    GW_OUT_PATHS = 1u << 2   /* full paths (length+1 u32 per walker; small runs only) */
};

/* Launch shape of the step kernel (all shapes give bit-identical walks).  block == 0
 * takes block, waves_per_cu, variant and the search order from gw_autotune (256 threads
 * of variant 0 before any tuning); only the caller's other flags are kept. */
typedef struct {
    uint32_t block;          /* threads per workgroup (64..1024, multiple of 64); 0 = tuned */
    uint32_t waves_per_cu;   /* persistent grid = CUs * waves_per_cu * 64 / block; 0 = as many
                                workgroups as the CUs hold */
    uint32_t variant;        /* 0 = per-step lane loop, 1 = flattened per-probe loop,
                                2 = lean per-load lane state machine (Node2vec and the SOPR
                                proposal sampler; other walker types run variant 0) */
    uint32_t flags;          /* GW_RUN_* */
} gw_launch;

enum {
    GW_RUN_BUCKETS = 1u << 0, /* after the run, the (prev_blk, cur_blk) bucket matrix of the
                                 walkers left, so gw_bucket_counts needs no extra pass: kept
                                 current from the run's worklist when it was current before
                                 (the filter takes the listed walkers out, a pass over the
                                 list puts the unfinished ones back), else recounted */
    GW_RUN_TRACE = 1u << 1,   /* record steps per current block (gw_trace_read) */
    GW_RUN_INTERP = 1u << 2,  /* membership tests by guarded interpolation search instead of
                                 bisection: same walks, fewer reads on evenly spread ids
                                 (randomly permuted graphs); step kernel variants 0 and 2 */
    GW_RUN_WINDOW = 1u << 3   /* membership tests read aligned 64-byte windows of 16 entries
                                 around each guess (bisected, or interpolated with
                                 GW_RUN_INTERP): same walks, one read where the last levels of
                                 a search took one each.  Probe counts equal the oracle's
                                 window orders when each block's edge words keep their offset
                                 modulo 64 bytes (HBM slots do; host-mapped blocks do when the
                                 host edge array starts on a 64-byte boundary); variants 0
                                 and 2 */
};

typedef struct {
    uint64_t steps;          /* successful transitions */
    uint64_t trials;         /* candidate draws */
    uint64_t probes;         /* adjacency reads for membership / intersection */
    uint64_t mix_steps;      /* SOPR steps taken by the mixture sampler */
    uint64_t walkers_run;    /* walkers whose (prev, cur) pair was resident */
    uint64_t walkers_ended;  /* walkers that finished during this run */
    uint64_t walkers_left;   /* unfinished walkers after this run */
    double kernel_ms;        /* step kernel time (events on the compute stream) */
    double filter_ms;        /* bucket pass time (GW_RUN_BUCKETS): the recount, or the pass
                                over the worklist */
    uint64_t zc_touches;     /* reads that went to host-mapped blocks (row, candidate, probe);
                                about 64 B of PCIe traffic each */
} gw_run_stats;

typedef struct {
    uint64_t steps, trials, probes, mix_steps;
    uint64_t h2d_bytes;      /* block upload bytes that crossed: (nverts+1)*8 + nedges*4 per raw
                                upload, the packed size per packed one */
    uint64_t h2d_uploads;
    double h2d_ms;           /* copy-stream busy time of block uploads */
    double kernel_ms;
    uint64_t hbm_bytes_held; /* bytes counted against cache_bytes */
    uint64_t zc_touches;     /* zero-copy reads of host-mapped blocks */
    uint64_t h2d_raw_bytes;  /* block uploads in the reference's format, (nverts+1)*8 + nedges*4
                                each, whether sent raw or packed (h2d_bytes counts what crossed) */
} gw_counters;

typedef struct {
    char name[64];
    uint32_t cus;
    uint32_t wave_size;
    uint64_t hbm_bytes;
    uint32_t lds_per_cu;
    uint32_t pad;
} gw_device_info;

/* ---- context: replaces graph_engine's GPU init (engine/engine.hpp:79-92) and the
 *      gpu_stream constructor (api/types.hpp:180-186) ---- */
gw_status gw_create(const gw_config* cfg, gw_ctx** out);
gw_status gw_destroy(gw_ctx* ctx);
const char* gw_last_error(const gw_ctx* ctx);
gw_status gw_device_info_get(gw_ctx* ctx, gw_device_info* out);

/* ---- pinned host memory: replaces cudaHostAlloc in cache_block::cudamalloc
 *      (engine/cache.hpp:106-120) and graph_buffer::alloc (api/graph_buffer.hpp:26-40) ---- */
gw_status gw_alloc_pinned(size_t bytes, void** out);
gw_status gw_free_pinned(void* p);

/* ---- block table: replaces initgpugraph (gpu/memory.hpp:28-49) ---- */
gw_status gw_set_block_table(gw_ctx* ctx, const gw_block_desc* blocks, uint32_t nblocks);

/* ---- HBM block cache: replaces initgpucache / initcacheblock (gpu/memory.hpp:51-107).
 *      Slots are sized for the largest block; fails with GW_ERR_BUDGET when
 *      nslots slots plus the walker state exceed cache_bytes. ---- */
gw_status gw_reserve_slots(gw_ctx* ctx, uint32_t nslots);
gw_status gw_slot_bytes(gw_ctx* ctx, uint64_t* bytes_per_slot);

/* Asynchronous upload of block `blk` into `slot` on the copy stream: replaces
 * copycachetogpu + pipeline_gpugraph (gpu/memory.hpp:124-139, :209-224).  beg holds
 * nverts+1 global offsets, csr holds nedges targets (the block's own arrays; pinned
 * memory, valid until the copy is done).  The host does not wait: the block is resident
 * for the next gw_run, whose kernels wait for the copy on the device. */
gw_status gw_upload_block(gw_ctx* ctx, uint32_t slot, uint32_t blk,
                          const uint64_t* beg, const uint32_t* csr);
/* Prefetch into a staging slot (the reference copies synchronously before each run,
 * engine/engine.hpp:415; this is the double buffer that overlaps the copy with the walk):
 * starts the copy of `blk` into `slot` at once; the block stays invisible to gw_run (which
 * does not wait for it) until gw_commit_block(slot).  The slot's previous block stops
 * being resident immediately.  gw_staged_block reports the staged block of a slot (-1:
 * none) and, if `ready` is not NULL, whether its copy has completed. */
gw_status gw_stage_block(gw_ctx* ctx, uint32_t slot, uint32_t blk,
                         const uint64_t* beg, const uint32_t* csr);
gw_status gw_commit_block(gw_ctx* ctx, uint32_t slot);
/* Packed uploads (no reference counterpart: the reference copies each block's beg and csr
 * as stored, 8 bytes per vertex and 4 per arc, gpu/memory.hpp:124-139).  `packed` is the
 * block's offsets and adjacency in the transfer encoding written by gw_pack_block (pinned
 * memory, valid until the copy is done; the block must have arcs).  The copy stream moves
 * the packed bytes into a staging buffer and two decode kernels write the block's beg and
 * csr into the slot, so the slot holds exactly what gw_upload_block / gw_stage_block would
 * put there; h2d_bytes counts the bytes that crossed, h2d_raw_bytes the block's size in the
 * reference's format.  The first time a block is sent from a buffer (its address, size and
 * a fingerprint of a few words), every header is checked against the block's vertex and arc
 * counts before anything is copied; later sends check the block header only, and the decode
 * kernels keep every read inside the staging buffer whatever the headers hold.  gw_set_pack_staging sizes the staging
 * buffer (counted against cache_bytes; 0 frees it): it holds a block's headers and packed
 * offsets (16 B per 256 arcs, 16 B per 256 vertices and the offsets' bits) whole and its arc
 * payload in pieces of the rest, so any size from those plus 1088 B up works; larger buffers
 * mean fewer, longer copies. */
gw_status gw_set_pack_staging(gw_ctx* ctx, uint64_t bytes);
/* Copy lanes (no reference counterpart): with 2, block copies alternate between two copy
 * streams, each decoding its packed blocks from its own half of the staging buffer, so one
 * lane's decode does not hold up the PCIe link while the other lane copies; a copy into a
 * slot whose last copy may still run stays on that copy's lane.  1 (the default) keeps one
 * copy stream.  Waits for every copy in flight.  gw_set_pack_staging gives each lane an
 * equal part of its size, rounded up to 256 B: set the lanes first. */
gw_status gw_set_copy_lanes(gw_ctx* ctx, uint32_t lanes);
gw_status gw_upload_block_packed(gw_ctx* ctx, uint32_t slot, uint32_t blk, const void* packed,
                                 uint64_t packed_bytes);
gw_status gw_stage_block_packed(gw_ctx* ctx, uint32_t slot, uint32_t blk, const void* packed,
                                uint64_t packed_bytes);
gw_status gw_staged_block(gw_ctx* ctx, uint32_t slot, int32_t* blk, uint32_t* ready);
/* Copies the beg (nverts + 1 entries) and csr (nedges) that slot holds for its resident or
 * staged block to host memory, after the slot's copy (and decode) is done; no reference
 * counterpart (an inspection call for tests of the upload paths). */
gw_status gw_read_slot(gw_ctx* ctx, uint32_t slot, uint64_t* beg, uint32_t* csr);
/* Marks a slot free and its block non-resident: replaces the INACTIVE update of
 * GOwalker_scheduler_t::copyblocks (engine/schedule.hpp:181-183). */
gw_status gw_evict(gw_ctx* ctx, uint32_t slot);
gw_status gw_wait_uploads(gw_ctx* ctx);
/* Makes block `blk` readable from pinned host memory (zero-copy), replacing the
 * mapped d_map table (engine/engine.hpp:85-92) used by zerocopy_update. */
gw_status gw_map_host_block(gw_ctx* ctx, uint32_t blk, const uint64_t* beg, const uint32_t* csr);
gw_status gw_unmap_host_blocks(gw_ctx* ctx);

/* ---- walk configuration and walkers: replaces gpu_createwalk and the generate
 *      kernels (engine/walk.hpp:32-95, :272-389) and cudarand (gpu/memory.hpp:109-122;
 *      Philox needs no state) ---- */
gw_status gw_set_walk(gw_ctx* ctx, const gw_walk_params* params);
/* Walker w starts at sources[w], or at w / walkers_per_vertex when sources is NULL. */
gw_status gw_init_walkers(gw_ctx* ctx, uint64_t nwalkers, uint32_t walkers_per_vertex,
                          const uint32_t* sources, uint32_t outputs);

/* Runs every unfinished walker whose (prev, cur) blocks are both resident until it
 * finishes or steps into a non-resident block: replaces the engine.hpp:418-448 loop
 * with gpu_run/SOPR_gpu_run, prefix, insertglobal and walkswap (gpu/gpuwalk.hpp:193-372).
 * With the lane kernel (variant 2), once a block resident in an earlier run since
 * gw_init_walkers is no longer resident, walkers whose current block is resident and
 * previous one is not run too, until a step needs the previous vertex's list: that step
 * stops unchanged, and the walker waits for the block (walks are the same either way). */
gw_status gw_run(gw_ctx* ctx, const gw_launch* launch, gw_run_stats* stats);

/* Launch-shape autotuner (no reference counterpart; the reference hard-codes 512 x 512,
 * apps/Node2vec.cu:47-48): times the step kernel on the first `sample` work items for each
 * candidate kernel variant, workgroup size and membership search order (GW_RUN_INTERP) that
 * `what` selects (GW_TUNE_*; 256 threads, variant 0, bisection is always a candidate),
 * restoring walker state after every trial, and makes the fastest the shape of later
 * gw_run calls that pass block == 0 (best->flags carries the search order).  Time it on
 * a sample that keeps the GPU as busy as the real runs: small samples favour shapes
 * that lose on full workloads. */
enum { GW_TUNE_BLOCK = 1u << 0, GW_TUNE_VARIANT = 1u << 1, GW_TUNE_SEARCH = 1u << 2, GW_TUNE_ALL = 7u };
gw_status gw_autotune(gw_ctx* ctx, uint64_t sample, uint32_t what, gw_launch* best, double* best_ms);

/* Bucket matrix W[prev_blk * nblocks + cur_blk] of unfinished walkers: replaces
 * copy_back's read of block_offset (engine/walk.hpp:488-498) and
 * walk_block::block_numwalks.  Free after gw_init_walkers or a gw_run with
 * GW_RUN_BUCKETS; otherwise one pass over the walkers. */
gw_status gw_bucket_counts(gw_ctx* ctx, uint64_t* W, uint64_t n);
/* Unfinished walkers: replaces walk_manager->nwalks() (engine/engine.hpp:519). */
gw_status gw_walkers_left(gw_ctx* ctx, uint64_t* left);
/* Unfinished walkers and the sum of their remaining hop caps (the work left; an upper
 * bound under geometric termination); the hop sum has no reference counterpart and feeds
 * the learned scheduler's zero-copy decision.  Free after a gw_run with GW_RUN_BUCKETS. */
gw_status gw_pending_work(gw_ctx* ctx, uint64_t* walkers, uint64_t* hops);

/* ---- learned block scheduler support (no reference counterpart: the reference's
 *      GOwalker_scheduler_t, engine/schedule.hpp:70-138, is the plain rule this
 *      scheduler falls back to) ---- */
/* Block-access trace of the last gw_run with GW_RUN_TRACE: walk steps taken while each
 * block held the walker's current vertex. */
gw_status gw_trace_read(gw_ctx* ctx, uint64_t* block_steps, uint64_t n);
/* 12 features per block (nblocks x 12, row major) from the current bucket matrix, the
 * resident set, the work left and a caller-kept age per block (iterations since last
 * use; may be NULL): log1p of the block's row, column and diagonal walker mass, of the
 * mass paired with resident blocks and of its smallest pairing, residency, log1p of the
 * walkers a load would release, age / 64, log1p(edges) / 24, the mean fraction of the
 * hop cap still ahead of unfinished walkers, log1p(mean degree) / 8, and the block's
 * share of all waiting walkers. */
gw_status gw_block_features(gw_ctx* ctx, const float* age, float* feats, uint64_t n);

/* MLP in -> h1 -> h2 -> 1 (ReLU), BF16 weights, FP32 accumulation, trained on the GPU
 * with Adam on FP32 master weights.  Parameters are exchanged as fp32 in the flat order
 * W1 (h1 x in), b1, W2 (h2 x h1), b2, w3 (h2), b3. */
typedef struct gw_mlp gw_mlp;
gw_status gw_mlp_create(gw_ctx* ctx, uint32_t in, uint32_t h1, uint32_t h2, uint64_t seed, gw_mlp** out);
gw_status gw_mlp_destroy(gw_ctx* ctx, gw_mlp* m);
gw_status gw_mlp_num_params(const gw_mlp* m, uint32_t* n);
gw_status gw_mlp_set_params(gw_ctx* ctx, gw_mlp* m, const float* params);
gw_status gw_mlp_get_params(gw_ctx* ctx, const gw_mlp* m, float* params);
/* x: n x in (host, row major); y: n outputs */
gw_status gw_mlp_forward(gw_ctx* ctx, gw_mlp* m, const float* x, uint32_t n, float* y);
/* mean-squared-error gradient of every parameter and the loss, for targets t */
gw_status gw_mlp_backward(gw_ctx* ctx, gw_mlp* m, const float* x, const float* t, uint32_t n,
                          float* grad, float* loss);
/* one Adam step (lr) on the batch; loss receives the batch loss before the step, or, with
 * loss NULL, the step is queued on the context's compute stream without a host wait (the
 * next call that waits for that stream completes it) */
gw_status gw_mlp_step(gw_ctx* ctx, gw_mlp* m, const float* x, const float* t, uint32_t n,
                      float lr, float* loss);
/* gw_block_features and the network's scores of every block in one call with one host
 * wait: the features go to feats (n >= nblocks rows), and the network (12 inputs) scores
 * them after standardizing each feature j as (f - x_mu[j]) / x_sigma[j] (the learned
 * scheduler's ranking, no reference counterpart; plug point engine/schedule.hpp:38-55). */
gw_status gw_block_scores(gw_ctx* ctx, gw_mlp* m, const float* age, const float* x_mu, const float* x_sigma,
                          float* feats, float* scores, uint64_t n);

/* ---- allocation-free walk (no counterpart in the reference).  For callers that cannot
 *      allocate device memory: code under HIP graph capture, tight serving loops, harnesses
 *      that forbid hipMalloc.  It is gw_init_walkers + gw_run over a graph the caller holds
 *      in device memory as one resident block: beg = nverts + 1 offsets into csr (beg[0] =
 *      0), csr = nedges sorted u32 targets.  Walker w = 0 .. nwalkers-1 starts at
 *      w / walkers_per_vertex; the walks equal gw_run's bit for bit.  hashes (nwalkers u64)
 *      and visits (nverts u32, cleared first) are optional device outputs.  The walker
 *      state, work queue and counters live in a fixed device array of the library,
 *      GW_NOALLOC_SCRATCH_BYTES per GPU, held from the first call on a GPU until the
 *      process ends; nwalkers above GW_NOALLOC_MAX_WALKERS is rejected (GW_ERR_ARG).  Work
 *      is queued on `stream` (a hipStream_t; NULL = the null stream).  stats == NULL: the
 *      call returns without waiting (make a first call outside a capture: it looks up the
 *      array's address on that GPU); otherwise it waits for the stream and fills steps,
 *      trials, probes, mix_steps and the walker counts (kernel_ms 0).  Nothing is
 *      allocated; calls that overlap on one GPU share its array and race.  Errors are read
 *      with gw_last_error(NULL). ---- */
#define GW_NOALLOC_MAX_WALKERS (1u << 20)
#define GW_NOALLOC_SCRATCH_BYTES (1024ull + 12ull * GW_NOALLOC_MAX_WALKERS)
gw_status gw_walk_resident_noalloc(const uint64_t* beg, const uint32_t* csr, uint32_t nverts, uint64_t nedges,
                                   const gw_walk_params* params, uint64_t nwalkers, uint32_t walkers_per_vertex,
                                   uint64_t* hashes, uint32_t* visits, void* stream, gw_run_stats* stats);

/* ---- outputs ---- */
gw_status gw_read_visits(gw_ctx* ctx, uint32_t* out, uint64_t n);
gw_status gw_read_hashes(gw_ctx* ctx, uint64_t* out, uint64_t n);
/* cur, prev, hop, status (1 length cap, 2 terminated, 3 dead end, 4 trial cap) */
gw_status gw_read_state(gw_ctx* ctx, uint32_t* out4, uint64_t n);
gw_status gw_read_paths(gw_ctx* ctx, uint32_t* out, uint64_t n);
gw_status gw_counters_get(gw_ctx* ctx, gw_counters* out);
/* Runs replayed from captured graphs and graphs captured since gw_create (no reference
 * counterpart: the reference launches every kernel of an iteration from the host,
 * engine/engine.hpp:406-448).  With GW_RUN_GRAPH=1 in the environment at gw_create, a
 * run's device work -- block table push, counter clears, filter pass, step kernel,
 * readbacks, bucket pass -- is captured as a HIP graph the first time its exact launch
 * sequence occurs and replayed when it recurs, and kernel_ms / filter_ms come from the
 * device clock (events in a graph cannot be timed); by default every kernel is launched
 * one by one and timed with HIP events. */
gw_status gw_graph_stats(gw_ctx* ctx, uint64_t* replays, uint64_t* captures);
/* Packed blocks by decode path since gw_create (no reference counterpart): decoded in their
 * own slot (payload copied to the slot's end, decoded in ordered passes), through the
 * staging ring in pieces, and on the copy stream (two copy lanes, GW_PACK_INLINE_DECODE=1,
 * or headers larger than a head area).  GW_PACK_INPLACE=0 in the environment at gw_create
 * sends every payload through the ring.  Waits until every upload asked for is enqueued. */
gw_status gw_upload_stats(gw_ctx* ctx, uint64_t* inplace, uint64_t* ring, uint64_t* copy_stream);
/* Worklist capacity in walker ids, taking effect at the next gw_init_walkers (no reference
 * counterpart: the reference's walker and bucket arrays hold every walker,
 * gpu/memory.hpp:15-20).
 * 0 = the default, nwalkers / 8 (at least min(nwalkers, 2^20)).  A gw_run whose runnable
 * walkers exceed it runs further filter + step rounds until none is left, with the same
 * walks; the bytes saved (4 per walker) are left to block slots. */
gw_status gw_set_worklist_cap(gw_ctx* ctx, uint64_t entries);
/* The worklist's capacity and the further rounds gw_run has taken since gw_create. */
gw_status gw_worklist_stats(gw_ctx* ctx, uint64_t* capacity, uint64_t* extra_rounds);

/* ---- synthetic graphs generated in HBM (the reference downloads real datasets,
 *      script/download_*.sh; its text->CSR converter cannot run as shipped,
 *      api/constants.hpp:5 + preprocess/graph_converter.hpp:227-231).  The result is the
 *      reference's CSR layout: u64 offsets indexed by vertex id (nverts+1), u32 sorted
 *      adjacency. ---- */
typedef struct gw_graph gw_graph;
enum { GW_GEN_UNDIRECTED = 1u << 0, GW_GEN_PERMUTE = 1u << 1, GW_GEN_DEDUP = 1u << 2 };
/* RMAT with 2^scale ids and nedges generated edges (self-loops dropped). */
gw_status gw_gen_rmat(gw_ctx* ctx, uint32_t scale, uint64_t nedges, double a, double b, double c,
                      uint64_t seed, uint32_t flags, gw_graph** out);
/* Directed Chung-Lu power law: out-degree exponent gamma, vertex 0 draws max_degree
 * out-edges, targets drawn by weight, lists sorted.  GW_GEN_DEDUP removes duplicate
 * targets and self-loops (drawing ~nedges/(1 - duplicate rate) so ~nedges remain). */
gw_status gw_gen_chung_lu(gw_ctx* ctx, uint64_t nverts, uint64_t nedges, double gamma,
                          uint64_t max_degree, uint64_t seed, uint32_t flags, gw_graph** out);
gw_status gw_graph_size(const gw_graph* g, uint64_t* nverts, uint64_t* nedges);
/* Copies the CSR to host memory (beg: nverts+1, csr: nedges). */
gw_status gw_graph_read(gw_ctx* ctx, const gw_graph* g, uint64_t* beg, uint32_t* csr);
gw_status gw_graph_free(gw_ctx* ctx, gw_graph* g);

/* ---- host-side block format helpers (no GPU needed) ---- */
/* The transfer encoding of one block (host only, no GPU): beg = the block's nverts + 1
 * global offsets, csr = its nedges arcs (a truncated giant vertex's list clipped at nedges).
 * Layout, in u32 words (arc and entry indices local to the block):
 *   header       {0x32505747, nvc = ceil((nverts + 1) / 256), vpw, 0};
 *   beg chunks   nvc x {beg[256k] low, high, payload word offset, widths w0..w3 (a byte
 *                each)}; then vpw payload words: per chunk, per group g of 64 entries, 2 w_g
 *                words, entry e of the group in bits [e w_g, (e + 1) w_g): beg[i] - beg[i-1]
 *                (0 for a chunk's first entry);
 *   arc chunks   ceil(nedges / 256) x {payload word offset, L | nz << 8, carry (the arc
 *                before the chunk when it starts inside a list, else 0), block-local vertex
 *                holding the chunk's first arc}; then per chunk of n arcs, with g_j the arc's
 *                difference (mod 2^32) to the arc before it in its list (to carry for a
 *                continued list's first arc in the chunk; its value at a list start) and S_j =
 *                g_0 + ... + g_j (64-bit): 8L words holding S_j mod 2^L in bits [jL, (j+1)L),
 *                then ceil((n + nz) / 32) words with bit (S_j >> L) + j set for each j (an
 *                Elias-Fano code; L = floor(log2(S_{n-1} / n)) or 0, nz = S_{n-1} >> L).
 * A block without arcs packs to 0 bytes (upload it raw).  With out == NULL only *bytes is
 * set. */
gw_status gw_pack_block(const uint64_t* beg, const uint32_t* csr, uint32_t nverts, uint64_t nedges,
                        void* out, uint64_t cap_bytes, uint64_t* bytes);

/* split_blocks (preprocess/split.hpp:15-87) over an in-memory .beg array. */
gw_status gw_split_blocks(const uint64_t* beg, uint64_t entries, uint64_t block_size,
                          uint32_t* vblocks, uint64_t* eblocks, uint64_t cap, uint64_t* ncuts);

#ifdef __cplusplus
}
#endif
#endif
