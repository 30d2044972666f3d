/*
 * gw_oracle.h -- CPU restatement of GOWalker's second-order random-walk step logic.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg load it.  The product (libgw.so and the
 * host engine) never links, includes or calls anything in oracle/.
 *
 * Contents and what pins each function (see the per-function comments in gw_oracle.c):
 *   orc_philox4x32_10   Philox-4x32-10 (Salmon et al., SC'11).  Pinned by the Random123
 *                       known-answer vectors and by rocRAND's / torch's open-source
 *                       implementations (tests/test_oracle_rng.py).
 *   orc_randnum_*       xoroshiro128+ RandNum of util/util.hpp:219-302, including the
 *                       literal type-punned inlineRandPair (util/util.hpp:291-301).
 *                       Pinned by golden vectors produced by the reference's own util.hpp
 *                       compiled into oracle/_ref (tests/golden/randnum_*.json).
 *   orc_walk / orc_run  Normative Philox step logic for Node2vec (gpu/gpuwalk.hpp:87-191
 *                       control flow with the CPU membership search of
 *                       engine/engine.hpp:224-243) and SOPR/SOSR (gpu/gpuwalk.hpp:207-307,
 *                       engine/engine.hpp:289-404).  Pinned by closed-form transition
 *                       probabilities on tiny graphs, brute-force enumeration of reduced
 *                       draws, and exact rational acceptance rates (tests/test_oracle_step.py).
 *   orc_ref_run         Transcription of sidewalkupdate / SOPR_sidewalkupdate
 *                       (engine/engine.hpp:158-404) run to completion, RandNum draws in
 *                       literal or fixed mode.  The CPU baseline; distributional anchor.
 *   orc_split_blocks    split_blocks (preprocess/split.hpp:15-87).  Pinned by hand-computed
 *                       cut points for every branch (tests/test_oracle_format.py).
 *   orc_max_choose      GOwalker_scheduler_t::max_choose_blocks (engine/schedule.hpp:70-138)
 *                       with lowest-id tie-break.  Pinned by hand-computed known answers.
 */
#ifndef GW_ORACLE_H
#define GW_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- RNG ---------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

typedef struct { uint64_t s0, s1; } orc_randnum;
void orc_randnum_init(orc_randnum* r, uint64_t seed);
uint64_t orc_randnum_lrand(orc_randnum* r);
/* inlineRandPair: literal=1 reproduces the reference's float type-pun byte for byte,
 * literal=0 is the 52-bit uniform the reference's comment describes. */
void orc_randnum_pair(orc_randnum* r, float max_val, uint64_t bound, int literal,
                      float* rand_val, uint64_t* rand_pos);

/* ---------------- graph / params ---------------- */
typedef struct {
    uint64_t nverts;        /* number of vertex ids; beg has nverts+1 entries */
    const uint64_t* beg;    /* global edge offsets */
    const uint32_t* csr;    /* sorted adjacency */
} orc_graph;

enum { ORC_N2V = 0, ORC_SOPR = 1 };

typedef struct {
    uint32_t algo;          /* ORC_N2V or ORC_SOPR */
    uint32_t length;        /* hop cap */
    uint64_t t_p, t_one, t_q, t_min;  /* node2vec integer thresholds, each in [0, 2^32] */
    uint32_t alpha_a, alpha_b;        /* SOPR fixed-point alpha: A = round(alpha*2^16), B = 2^16-A */
    uint64_t t_term;        /* geometric termination threshold floor(term*2^32); 0 = none */
    uint32_t sopr_sampler;  /* 0: proposal (duplicate-free lists), 1: envelope/mixture (any lists) */
    uint32_t search;        /* membership search order: 0 bisection (the reference's), 1 guarded
                               interpolation, 2 / 3 16-entry windows with bisected / interpolated
                               guesses; same answers, different probe counts */
} orc_typecfg;

typedef struct {
    uint64_t seed;
    orc_typecfg type[2];    /* per walker type; the walker's type picks one */
} orc_walkcfg;

typedef struct {
    uint32_t cur, prev, hop;
    uint32_t status;        /* 1 = length cap, 2 = terminated, 3 = dead end */
    uint64_t hash;
    uint64_t steps, trials, probes, mix_steps;
} orc_walk_result;

/* membership of x in the sorted list a[0..n) with the given search order (values are
 * below id_end; a[0] is entry `head` of the edge array, which places the window orders'
 * aligned reads); probes counts the list reads */
int orc_member(const uint32_t* a, uint64_t n, uint32_t x, uint32_t search, uint32_t id_end, uint64_t head,
               uint64_t* probes);

/* host-side parameter conversion (the product computes the same values independently) */
void orc_n2v_thresholds(double p, double q, orc_typecfg* t);
uint32_t orc_alpha_fx(double alpha);
uint64_t orc_term_threshold(double term);

/* one walker from its start state; path (optional) receives length+1 entries,
 * unused entries are set to 0xFFFFFFFF. */
void orc_walk(const orc_graph* g, const orc_walkcfg* cfg, uint32_t wid, uint32_t type,
              uint32_t source, uint32_t* path, orc_walk_result* res);

/* all walkers: walker w starts at sources[w] (or w / wpv when sources == NULL) with
 * type types[w] (or (w % ntypes_mod) when types == NULL and ntypes_mod > 0, else 0).
 * Any output pointer may be NULL.  visits[x] counts arrivals at x.  Uses OpenMP. */
typedef struct {
    uint64_t steps, trials, probes, mix_steps, walkers;
} orc_totals;
void orc_run(const orc_graph* g, const orc_walkcfg* cfg, uint64_t nwalkers,
             const uint32_t* sources, uint32_t wpv, const uint32_t* types, uint32_t ntypes_mod,
             uint32_t* paths, uint64_t* hashes, uint32_t* final_state /* cur, prev, hop, status per walker */,
             uint32_t* visits, orc_totals* totals, int nthreads);

/* sampler mode decision for SOPR (1 = mixture, 0 = rejection); exposed for tests */
int orc_sopr_use_mixture(uint64_t du, uint64_t dv, uint32_t a, uint32_t b);

/* ---------------- reference CPU path transcription ---------------- */
typedef struct {
    uint32_t algo;        /* ORC_N2V or ORC_SOPR */
    float p, q, alpha;    /* real_t in the reference */
    uint32_t length;
    double term;          /* build addition; 0 reproduces the reference */
    int literal;          /* 1: type-punned inlineRandPair; 0: fixed 52-bit uniform */
    uint64_t seed;        /* walker w uses RandNum(fmix64(seed) ^ fmix64(w + c)) */
} orc_refcfg;

/* Runs walkers [w0, w0+nw) to completion; returns real moves (the reference's bogus
 * dead-end move to vertex 0 is applied but not counted).  visits may be NULL.
 * max_seconds > 0 stops early (whole walkers only); *done receives walkers finished. */
uint64_t orc_ref_run(const orc_graph* g, const orc_refcfg* cfg, uint64_t w0, uint64_t nw,
                     const uint32_t* sources, uint32_t wpv, uint32_t* visits,
                     uint64_t* trials, double max_seconds, uint64_t* done, int nthreads);

/* ---------------- block format ---------------- */
/* split_blocks over a .beg array with `entries` u64 values.  Writes cut points into
 * vblocks / eblocks (capacity cap); returns the number of cut points written
 * (nblocks + 1), or 0 if cap is too small. */
size_t orc_split_blocks(const uint64_t* beg, uint64_t entries, uint64_t block_size,
                        uint32_t* vblocks, uint64_t* eblocks, size_t cap);

/* ---------------- plain scheduler ---------------- */
/* W: nblocks x nblocks walker counts (row = prev block, col = cur block).
 * buckets/nbuckets: resident set in slot order, updated in place.
 * Returns 1 if a block was added/replaced, 0 otherwise. */
int orc_max_choose(const uint64_t* W, uint32_t nblocks, uint32_t* buckets, uint32_t* nbuckets,
                   uint32_t ncblock, uint64_t zero_threshold);

#ifdef __cplusplus
}
#endif
#endif
