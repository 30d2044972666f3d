// gw_internal.h -- structures shared by the C-ABI layer (gw_api.cpp) and the gfx950
// kernels (gw_walk.hip, gw_gen.hip).  Not part of the public ABI.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

#include "gw.h"

namespace gw {

constexpr uint64_t kFnvOffset = 1469598103934665603ULL;
constexpr uint64_t kFnvPrime = 1099511628211ULL;
constexpr uint32_t kMaxTrials = 1u << 24;   // per step; the walk ends with status 4 beyond it
constexpr uint32_t kTagStep = 0, kTagTerm = 1;

// walker meta word: hop in bits 0..19, status in 20..23, type in 24
constexpr uint32_t kHopMask = 0xFFFFFu;
constexpr uint32_t kStatusShift = 20, kTypeShift = 24;
constexpr uint32_t kMaxLength = kHopMask;
enum : uint32_t { kActive = 0, kDoneLength = 1, kDoneTerm = 2, kDoneDeadEnd = 3, kDoneTrialCap = 4 };
// meta bit: the walker's step stopped where it needed its previous vertex's list, whose block
// was not resident (the lane kernel picks such walkers up again once that block is)
constexpr uint32_t kWaitPrev = 1u << 25;
// DevBlock::resident: 0 not readable, 1 in an HBM slot, 2 host-mapped (zero-copy reads)
enum : uint32_t { kNotResident = 0, kInSlot = 1, kMapped = 2 };

// One block as the kernels see it: the block's own beg (nverts+1 global offsets) and
// csr arrays, wherever they live (an HBM slot, a caller's device graph, or pinned host
// memory for the zero-copy tail).  resident == 0 means the pointers must not be read.
struct DevBlock {
    const uint64_t* beg;
    const uint32_t* csr;
    uint64_t start_edge;
    uint64_t nedges;      // possibly truncated (giant single-vertex block)
    uint32_t start_vert;
    uint32_t nverts;
    uint32_t resident;
    uint32_t pad;
};

// Per walker type, in the integer form the step logic uses.
struct DevType {
    uint64_t t_p, t_one, t_q, t_min;   // node2vec acceptance thresholds, each <= 2^32
    // the same thresholds as T - 1 (ra < T <=> ra <= T - 1; every T is at least 1), the
    // bounds of the band in which membership decides (min and max of t_one, t_q) and the
    // SOPR proposal's N(cur) threshold (alpha_b << 16) - 1, for 32-bit compares
    uint32_t m_p, m_one, m_q, m_min, m_lo, m_hi, m_prop;
    uint32_t pad0;                     // no implicit padding: run arguments are compared bytewise
    uint64_t t_term;                   // floor(term * 2^32), 0 = no geometric termination
    uint32_t algo;                     // 0 node2vec, 1 SOPR/SOSR
    uint32_t length;
    uint32_t alpha_a, alpha_b;         // A = round(alpha * 2^16), B = 2^16 - A
    uint32_t sampler;                  // SOPR: 0 proposal, 1 envelope/mixture
    uint32_t pad;
};

struct WalkArgs {
    const DevBlock* blocks;
    const uint32_t* vcut;     // nblocks+1 vertex cut points
    uint32_t nblocks;
    uint32_t ntypes;
    uint32_t seed_lo, seed_hi;
    DevType type0, type1;
    // walker state (SoA)
    uint32_t* cur;
    uint32_t* prev;
    uint32_t* meta;
    uint64_t* hash;           // optional
    uint32_t* paths;          // optional, path_stride entries per walker
    uint32_t path_stride;
    uint32_t* visits;         // optional: visits_parts arrays of visits_stride counters
    uint64_t visits_stride;
    uint32_t visits_parts;    // a workgroup counts into array blockIdx.x % visits_parts
    // work
    const uint32_t* worklist; // walker ids; nullptr = identity (finished walkers are skipped)
    uint32_t nwork;
    const uint32_t* nwork_dev;  // if set, the worklist length is read here (written by the filter pass)
    uint32_t zc_count;        // some block is host-mapped: count zero-copy reads (counters[7])
    uint32_t search_end;      // 0: membership by bisection; else interpolation, ids below it
    uint32_t search_win;      // probes read aligned 16-entry windows
    uint32_t relax;           // lane kernel: the worklist holds walkers whose previous block is not
                              // resident (FilterArgs::relax_prev): the instance that stops their
                              // steps where they need it
    uint32_t* queue;          // work-fetch counter, zero at launch
    uint32_t chunk;           // walkers a wavefront takes per queue atomic (multiple of 64)
    unsigned long long* counters;  // steps, trials, probes, mix, ended, left
    unsigned long long* block_steps;  // optional: steps per current block (access trace)
};

struct InitArgs {
    uint32_t ntypes;
    uint32_t len0, len1;
    uint64_t nwalkers;
    uint32_t wpv;
    const uint32_t* sources;  // device copy, optional
    uint32_t* cur;
    uint32_t* prev;
    uint32_t* meta;
    uint64_t* hash;
    uint32_t* paths;
    uint32_t path_stride;
};

struct FilterArgs {
    const DevBlock* blocks;
    uint32_t nblocks;
    uint64_t nwalkers;
    const uint32_t* meta;
    const uint32_t* cur;      // a walker's (prev, cur) blocks come from its vertices and the cut points
    const uint32_t* prev;
    const uint32_t* vcut;
    uint32_t* worklist;       // optional: ids of walkers whose (prev, cur) blocks are resident
    uint32_t* nwork;          // append counter of the worklist: every runnable walker, also past work_cap
    uint32_t work_cap;        // worklist entries (ids past it are counted, not written)
    uint32_t* bucket;         // optional nblocks*nblocks counters
    unsigned long long* left; // unfinished walkers
    unsigned long long* hops; // optional: sum of the remaining hop caps of unfinished walkers
    uint32_t len0, len1;      // hop caps of walker types 0 / 1
    uint32_t* bucket_part;    // optional: part_rows x nblocks^2 per-workgroup rows of the bucket matrix
    uint32_t part_rows;
    // 1: a walker whose current block is resident is runnable also when its previous one is
    // not, unless it stopped for it before (kWaitPrev); for the lane kernel, which stops a
    // step where the previous list is needed and lets the others run
    uint32_t relax_prev;
    // optional: the bucket matrix as it stands before a run; walkers the filter writes into
    // the worklist leave it (the pass over that worklist after the step kernel puts the
    // unfinished ones back), and taken[0] / taken[1] count them and their remaining hops
    uint32_t* bucket_sub;
    unsigned long long* taken;
    // optional: visit the walker ids gather[0 .. min(*gather_n, nwalkers)) instead of
    // 0 .. nwalkers-1 (a worklist); bucket counts then add to the matrix as it stands
    const uint32_t* gather;
    const uint32_t* gather_n;
};

// records a message for gw_last_error (gw_api.cpp)
gw_status gen_fail(gw_ctx* ctx, gw_status st, const char* msg);

// the integer form of one walker type (gw_api.cpp); nullptr on success, else the reason
const char* dev_type_from(const gw_walk_type& ty, DevType* out);

// launchers (gw_walk.hip)
hipError_t launch_init(const InitArgs& a, hipStream_t s);
hipError_t launch_filter(const FilterArgs& a, uint32_t grid, hipStream_t s);
hipError_t launch_walk(const WalkArgs& a, bool one_block, uint32_t variant, uint32_t grid,
                       uint32_t block, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
hipError_t launch_copy_u32(uint32_t* dst, const uint32_t* src, uint64_t n, hipStream_t s);
hipError_t launch_timestamp(unsigned long long* dst, hipStream_t s);   // device clock into dst
// the network's inputs per sample (gw_mlp.hip)
uint32_t mlp_num_inputs(const gw_mlp* m);
// forward pass on device-resident rows f (n x inputs), standardized per input with musig
// (the means, then the deviations; device memory); *y: the device buffer of the n outputs,
// valid until the network's next call.  Queued on the context's compute stream.
gw_status mlp_forward_device(gw_ctx* ctx, gw_mlp* m, const float* f, const float* musig, uint32_t n, const float** y);
// Packed blocks (gw_pack.hip: a header, beg chunks of bit-packed differences, arc chunks of
// Elias-Fano-coded gap sums).  Word offsets of the parts, from the block's first word.
constexpr uint32_t kPackChunkArcs = 256;
constexpr uint32_t kPackMagic = 0x32505747u;
struct PackLayout {
    uint64_t nvc, vpw, nch;   // beg chunks, beg payload words, arc chunks
    uint64_t vh_off, vp_off, ah_off, head_w;   // beg headers, beg payload, arc headers; words before the arc payload
};
// payload words of an arc chunk of n arcs with header h
inline uint64_t pack_chunk_words(const uint32_t* h, uint64_t n)
{
    return 8ull * (h[1] & 0xFFu) + ((n + (h[1] >> 8) + 31) >> 5);
}
// Fills lay from a packed block of `words` words of the block of nv vertices and ne arcs and
// checks its header; with `full`, every beg and arc chunk header too (offsets contiguous,
// widths and sizes in range, first vertices in the block and non-decreasing).  Returns an
// error message or nullptr.  Host code over host memory.
const char* pack_layout(const uint32_t* pk, uint64_t words, uint32_t nv, uint64_t ne, bool full, PackLayout* lay);
// FNV-1a over a packed block's size and a few of its words (the first 16, 16 at the middle,
// the last 16): tells blocks apart that reuse an address without reading them whole
inline uint64_t pack_fingerprint(const uint32_t* pk, uint64_t words)
{
    uint64_t h = 1469598103934665603ULL ^ words;
    auto mix = [&](uint64_t i0) {
        for (uint64_t i = i0; i < i0 + 16 && i < words; ++i) h = (h ^ pk[i]) * 1099511628211ULL;
    };
    mix(0);
    mix(words / 2);
    mix(words > 16 ? words - 16 : 0);
    return h;
}
// beg (nv + 1 global offsets) from a packed block whose first lay.head_w words are at staging
hipError_t launch_unpack_beg(const uint32_t* staging, const PackLayout& lay, uint32_t nv, uint64_t* beg, hipStream_t s);
// csr words of arc chunks [c_begin, c_end) into out; head holds the block's first lay.head_w
// words, payload `avail` of its arc payload words from word_base on; beg: the decoded
// offsets on the device
hipError_t launch_unpack_arcs(const uint32_t* head, const uint32_t* payload, const PackLayout& lay, const uint64_t* beg,
                              uint64_t start_edge, uint32_t nv, uint64_t ne, uint64_t c_begin, uint64_t c_end,
                              uint64_t word_base, uint64_t avail, uint32_t* out, hipStream_t s);
// v[i] = sum over p < parts of v[p * n + i]; the other arrays are cleared
hipError_t launch_fold_visits(uint32_t* v, uint64_t n, uint32_t parts, hipStream_t s);

// Small transfers (multiples of 4 bytes, up to a few MB) on the context's compute stream
// through its pinned ring and a copy kernel (no DMA engine).  h2d returns once the copy is
// enqueued; d2h returns with the data on the host.
hipError_t h2d_small(gw_ctx* ctx, void* dst, const void* src, uint64_t bytes);
hipError_t d2h_small(gw_ctx* ctx, void* dst, const void* src, uint64_t bytes);
uint32_t walk_blocks_per_cu(bool one_block, uint32_t nblocks, uint32_t variant, uint32_t block);   // resident workgroups per CU
hipError_t launch_features(const uint32_t* bucket, uint32_t nblocks, const DevBlock* blocks, const float* age, float ahead,
                           float* feats, hipStream_t s);
constexpr uint32_t kTraceLdsBlocks = 2048;   // per-workgroup LDS histogram size of the trace
constexpr uint32_t kLdsBlocks = 128;         // block tables up to this size are staged in LDS by the step kernel
constexpr uint32_t kLutSize = 2048;          // entries of the LDS vertex -> block lookup table
constexpr uint32_t kVisitParts = 8;          // visit-count arrays, one per XCD when HBM allows
constexpr uint32_t kNumFeatures = 12;
constexpr uint32_t kFilterLdsBuckets = 12288;  // bucket matrices up to 110 x 110 blocks count in LDS
constexpr uint32_t kFilterLdsResident = 32768; // resident sets of up to 32768 blocks are an LDS bit table

}  // namespace gw
