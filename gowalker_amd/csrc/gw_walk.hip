// gw_walk.hip -- gfx950 walker-step kernels for GOWalker's second-order walks.
//
// Replaces gpu/gpuwalk.hpp: gpu_update/gpu_run (:87-205, Node2vec), SOPR_gpu_update /
// SOPR_gpu_run (:207-319), prefix/insertglobal/walkswap (:321-372, here a worklist
// filter with wavefront ballot compaction), and the walker generators of
// engine/walk.hpp:32-95.
//
// Step semantics (identical, bit for bit, to the CPU restatement used as the oracle):
//   * Philox-4x32-10 keyed by the seed; counter (walker, hop, k, tag).  Trial k of a
//     step draws block k/2 with tag 0; even k uses (.x pos, .y accept), odd k uses
//     (.z, .w).  The geometric termination draw is block 0 with tag 1 (.x).
//   * pos = (u32 * deg) >> 32.
//   * Node2vec (gpuwalk.hpp:118-150): accept iff accept-word < T(class), classes
//     x == prev (1/p), x in N(prev) (1), other (1/q), T = floor(w/max * 2^32); the
//     smallest T is tested before the candidate is classified, and prev's list is
//     searched only when the accept word lies between T(1) and T(1/q), the one band in
//     which the two classes x != prev can fall into decide differently.
//   * SOPR (gpuwalk.hpp:241-266): P(x) ~ (1-a)/dv + a/du [x in N(prev), x != prev] with
//     a = A/2^16, sampled by envelope rejection or, when the envelope is loose, by an
//     exact two-component mixture after counting |N(cur) & N(prev)| (see sopr_step).
//   * Membership: the answer of the textbook binary search over prev's sorted adjacency
//     (engine/engine.hpp:224-243, not the defective device search of gpuwalk.hpp:26-48),
//     found by bisection, guarded interpolation or aligned 64-byte window reads (search
//     mode SM); the oracle restates every order, probe counts included.
#include "gw_device.h"

namespace gw {

typedef unsigned __int128 u128;

// per-lane counters (u32: a lane's share of one launch stays far below 2^32); summed in
// u64 across the wavefront at the end
struct Cnt {
    uint32_t steps, trials, probes, mix, ended, left, zc;
};

__device__ __forceinline__ uint32_t bitlen64(uint64_t n) { return n ? 64u - (uint32_t)__clzll(n) : 0u; }

// One window probe of the window order: the 16 list entries in the 64-byte piece (aligned
// to 64 bytes in memory, half a 128-byte line) that holds the guessed position.  Slots
// place each block so that an entry's address is 4 x its index in the whole graph's edge
// array modulo 64, so the pieces are those the oracle's window order reads.  scan_window
// looks at the entries [s, e) of the piece that belong to the remaining range: whether x
// is among them, and the first and last of them (x below the first or above the last
// keeps the search on that side; x between them is absent).
struct WinScan {
    bool found;
    uint32_t first, last;
};

// entry i in [0, 16) of the window q0..q3, by a tree of selects on the bits of i (no
// indexed registers, no private arrays)
__device__ __forceinline__ uint32_t pick16(const uint4& q0, const uint4& q1, const uint4& q2, const uint4& q3,
                                           uint32_t i)
{
    const bool b0 = i & 1u, b1 = i & 2u, b2 = i & 4u, b3 = i & 8u;
    const uint32_t a0 = b0 ? q0.y : q0.x, a1 = b0 ? q0.w : q0.z, a2 = b0 ? q1.y : q1.x, a3 = b0 ? q1.w : q1.z;
    const uint32_t a4 = b0 ? q2.y : q2.x, a5 = b0 ? q2.w : q2.z, a6 = b0 ? q3.y : q3.x, a7 = b0 ? q3.w : q3.z;
    const uint32_t c0 = b1 ? a1 : a0, c1 = b1 ? a3 : a2, c2 = b1 ? a5 : a4, c3 = b1 ? a7 : a6;
    const uint32_t d0 = b2 ? c1 : c0, d1 = b2 ? c3 : c2;
    return b3 ? d1 : d0;
}

// bits 4j..4j+3: which of q's four entries equal x
__device__ __forceinline__ uint32_t eq_bits(const uint4& q, uint32_t x, uint32_t j)
{
    return ((q.x == x ? 1u : 0u) | (q.y == x ? 2u : 0u) | (q.z == x ? 4u : 0u) | (q.w == x ? 8u : 0u)) << (4 * j);
}

__device__ __forceinline__ WinScan scan_window(const uint4& q0, const uint4& q1, const uint4& q2, const uint4& q3,
                                               uint32_t s, uint32_t e, uint32_t x)
{
    const uint32_t eq = eq_bits(q0, x, 0) | eq_bits(q1, x, 1) | eq_bits(q2, x, 2) | eq_bits(q3, x, 3);
    WinScan r;
    r.found = ((eq >> s) & ((1u << (e - s)) - 1u)) != 0;
    r.first = pick16(q0, q1, q2, q3, s);
    r.last = pick16(q0, q1, q2, q3, e - 1);
    return r;
}

__device__ __forceinline__ const uint32_t* window_of(const uint32_t* p)
{
    return (const uint32_t*)((uintptr_t)p & ~(uintptr_t)63);
}

// Window orders (the oracle's searches 2 and 3): each probe reads the window holding the
// guess, which bisects the remaining range (INTERP false) or is interpolated from its
// value bounds (INTERP true, bisecting after a probe that failed to halve the range); a
// miss either shows x absent (it falls between two entries read) or keeps the range on
// one side of the window, bounded by the value next to it.
template <bool INTERP>
__device__ __forceinline__ bool member_window(const uint32_t* __restrict__ a, uint64_t n, uint32_t x,
                                              uint32_t& probes, uint32_t id_end)
{
    uint64_t lo = 0, hi = n;   // remaining range [lo, hi), values in [vlo, vhi)
    uint32_t vlo = 0, vhi = id_end;
    bool bisect = false;
    if (INTERP && x >= vhi) return false;
    while (lo < hi) {
        const uint32_t len = (uint32_t)(hi - lo);
        uint64_t g;
        if (!INTERP || bisect || vhi - vlo < 2) {
            g = lo + ((len - 1) >> 1);
        } else {
            const uint32_t off = (uint32_t)((float)(x - vlo) * (float)len / (float)(vhi - vlo));
            g = lo + (off < len ? off : len - 1);
        }
        const uint32_t* wp = window_of(a + g);
        const uint4* q = (const uint4*)wp;
        const uint4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
        ++probes;
        const int64_t w0 = wp - a;   // list index of wp[0], in (g - 16, g]
        const int64_t s = w0 > (int64_t)lo ? w0 : (int64_t)lo;
        const int64_t e = w0 + 16 < (int64_t)hi ? w0 + 16 : (int64_t)hi;
        const WinScan r = scan_window(q0, q1, q2, q3, (uint32_t)(s - w0), (uint32_t)(e - w0), x);
        if (r.found) return true;
        if (x < r.first) {
            hi = (uint64_t)s;
            vhi = r.first;
        } else if (x > r.last) {
            lo = (uint64_t)e;
            vlo = r.last + 1;
        } else {
            return false;
        }
        bisect = (hi - lo) * 2 > len;
    }
    return false;
}

// engine/engine.hpp:224-243
// Membership of x in the sorted list a[0..n), probes counting the list reads.  Search
// mode SM: 0 bisects (the reference's search).  1 (every id below id_end) interpolates
// the next probe from the value bounds [vlo, vhi) of the remaining range -- a few probes
// when ids are spread evenly, as in randomly permuted graphs -- in exact fp32 arithmetic
// (the oracle computes the same positions), and bisects whenever a probe failed to halve
// the range and inside ranges of at most 16 entries (one 64-byte line), so no list takes
// more than about twice the bisection's probes.  2 and 3 read windows (member_window).
// All orders give the same answer.
template <uint32_t SM>
__device__ __forceinline__ bool member(const uint32_t* __restrict__ a, uint64_t n, uint32_t x, uint32_t& probes,
                                       uint32_t id_end)
{
    if (SM & 2) return member_window<(SM & 1) != 0>(a, n, x, probes, id_end);
    int64_t lo = 0, hi = (int64_t)n - 1;
    if (!(SM & 1)) {
        while (lo <= hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            const uint32_t m = a[mid];
            ++probes;
            if (m == x) return true;
            if (m < x) lo = mid + 1; else hi = mid - 1;
        }
        return false;
    }
    uint32_t vlo = 0, vhi = id_end;   // a[lo..hi] within [vlo, vhi)
    bool bisect = false;
    if (x >= vhi) return false;
    while (lo <= hi) {
        const uint32_t len = (uint32_t)(hi - lo + 1);
        int64_t mid;
        if (bisect || len <= 16 || vhi - vlo < 2) {
            mid = lo + ((hi - lo) >> 1);
        } else {
            uint32_t off = (uint32_t)((float)(x - vlo) * (float)len / (float)(vhi - vlo));
            mid = lo + (off < len ? off : len - 1);
        }
        const uint32_t m = a[mid];
        ++probes;
        if (m == x) return true;
        if (m < x) { lo = mid + 1; vlo = m + 1; } else { hi = mid - 1; vhi = m; }
        bisect = (uint64_t)(hi - lo + 1) * 2 > len;
    }
    return false;
}

__device__ __forceinline__ uint64_t lower_bound(const uint32_t* __restrict__ a, uint64_t lo, uint64_t hi,
                                                uint32_t x, uint32_t& probes)
{
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        ++probes;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t upper_bound(const uint32_t* __restrict__ a, uint64_t lo, uint64_t hi,
                                                uint32_t x, uint32_t& probes)
{
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        ++probes;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Distinct values y of N(u) (y != u) in ascending order, matched against their runs in
// N(v).  select == ~0: returns the number c of matching positions of N(v); otherwise
// returns the index in N(v) of the select-th matching position.
__device__ uint64_t intersect(const uint32_t* __restrict__ nu, uint64_t du,
                              const uint32_t* __restrict__ nv, uint64_t dv,
                              uint32_t u, uint64_t select, uint32_t& probes)
{
    uint64_t c = 0, base = 0;
    uint32_t last = 0;
    bool have_last = false;
    for (uint64_t i = 0; i < du && base < dv; ++i) {
        const uint32_t y = nu[i];
        ++probes;
        if ((have_last && y == last) || y == u) { last = y; have_last = true; continue; }
        last = y; have_last = true;
        const uint64_t lb = lower_bound(nv, base, dv, y, probes);
        if (lb < dv) {
            ++probes;
            if (nv[lb] == y) {
                const uint64_t ub = upper_bound(nv, lb, dv, y, probes);
                const uint64_t cnt = ub - lb;
                if (select != ~0ULL && select < c + cnt) return lb + (select - c);
                c += cnt;
                base = ub;
                continue;
            }
        }
        base = lb;
    }
    return c;
}

__device__ __forceinline__ bool sopr_use_mixture(uint64_t du, uint64_t dv, uint32_t a, uint32_t b)
{
    const uint64_t bdu = (uint64_t)b * du;
    const uint64_t num = bdu + (uint64_t)a * dv;
    const uint64_t e = (num + bdu - 1) / bdu;
    return du * (4ull * bitlen64(dv) + 2) < e * (1 + bitlen64(du));
}

template <uint32_t SM>
__device__ __forceinline__ bool n2v_step(const uint32_t* __restrict__ nv, uint64_t dv,
                                         const uint32_t* __restrict__ nu, uint64_t du, uint32_t prev,
                                         const DevType& tp, uint32_t w, uint32_t hop, uint32_t k0, uint32_t k1,
                                         uint32_t& x, Cnt& c, uint32_t id_end)
{
    uint4 d = make_uint4(0, 0, 0, 0);
    for (uint32_t k = 0; k < kMaxTrials; ++k) {
        if (!(k & 1)) d = philox(w, hop, k >> 1, kTagStep, k0, k1);
        const uint32_t rp = (k & 1) ? d.z : d.x;
        const uint32_t ra = (k & 1) ? d.w : d.y;   // accept iff ra < T, i.e. ra <= T - 1
        ++c.trials;
        x = nv[((uint64_t)rp * dv) >> 32];
        if (ra <= tp.m_min) return true;
        uint32_t t;
        if (x == prev) t = tp.m_p;
        else if (ra <= tp.m_lo) return true;      // accepted whether or not x is in N(prev)
        else if (ra > tp.m_hi) continue;          // rejected either way
        else t = member<SM>(nu, du, x, c.probes, id_end) ? tp.m_one : tp.m_q;
        if (ra <= t) return true;
    }
    return false;
}

template <uint32_t SM>
__device__ __forceinline__ bool sopr_step(const uint32_t* __restrict__ nv, uint64_t dv,
                                          const uint32_t* __restrict__ nu, uint64_t du, uint32_t prev,
                                          const DevType& tp, uint32_t w, uint32_t hop, uint32_t k0, uint32_t k1,
                                          uint32_t& x, Cnt& c, uint32_t id_end)
{
    if (tp.sampler == 0) {
        // proposal rejection: N(cur) with probability 1-a, else a prev-neighbour that is
        // also a cur-neighbour (duplicate-free lists)
        const uint64_t tb = (uint64_t)tp.alpha_b << 16;
        for (uint32_t k = 0; k < kMaxTrials; ++k) {
            const uint4 d = philox(w, hop, k, kTagStep, k0, k1);
            ++c.trials;
            if ((uint64_t)d.x < tb) {
                x = nv[((uint64_t)d.y * dv) >> 32];
                return true;
            }
            const uint32_t y = nu[((uint64_t)d.y * du) >> 32];
            if (y != prev && member<SM>(nv, dv, y, c.probes, id_end)) {
                x = y;
                return true;
            }
        }
        return false;
    }
    const uint64_t A = tp.alpha_a, B = tp.alpha_b;
    const u128 num32 = (u128)(B * du) << 32;
    if (sopr_use_mixture(du, dv, tp.alpha_a, tp.alpha_b)) {
        ++c.mix;
        ++c.trials;
        const uint4 d = philox(w, hop, 0, kTagStep, k0, k1);
        const uint64_t cnt = intersect(nu, du, nv, dv, prev, ~0ULL, c.probes);
        if (cnt == 0 || (u128)d.x * (B * du + A * cnt) < num32) {
            x = nv[((uint64_t)d.y * dv) >> 32];
        } else {
            const uint64_t k = ((uint64_t)d.z * cnt) >> 32;
            x = nv[intersect(nu, du, nv, dv, prev, k, c.probes)];
        }
        return true;
    }
    const uint64_t den = B * du + A * dv;
    uint4 d = make_uint4(0, 0, 0, 0);
    for (uint32_t k = 0; k < kMaxTrials; ++k) {
        if (!(k & 1)) d = philox(w, hop, k >> 1, kTagStep, k0, k1);
        const uint32_t rp = (k & 1) ? d.z : d.x;
        const uint64_t ra = (k & 1) ? d.w : d.y;
        ++c.trials;
        x = nv[((uint64_t)rp * dv) >> 32];
        if ((u128)ra * den < num32) return true;
        if (x != prev && member<SM>(nu, du, x, c.probes, id_end)) return true;
    }
    return false;
}

// vertex -> block: the block whose range [vcut[b], vcut[b+1]) holds v; empty blocks
// (duplicate cut points) never match, as in getblk (gpuwalk.hpp:69-76).
__device__ __forceinline__ uint32_t find_block(const uint32_t* __restrict__ vcut, uint32_t nblocks, uint32_t v)
{
    uint32_t lo = 0, hi = nblocks;   // largest b in [0, nblocks) with vcut[b] <= v
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (vcut[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

// The same answer from an LDS table: lut[i] = the largest b with vcut[b] <= i << shift,
// then forward to the largest b with vcut[b] <= v (one or two reads when blocks span more
// ids than a table entry).
__device__ __forceinline__ uint32_t find_block_lut(const uint16_t* lut, uint32_t shift, const uint32_t* vcut,
                                                   uint32_t nblocks, uint32_t v)
{
    uint32_t b = lut[v >> shift];
    while (b + 1 < nblocks && vcut[b + 1] <= v) ++b;
    return b;
}

// Copies a block table of at most kLdsBlocks entries and its cut points into LDS and
// builds the lookup table; returns its shift.  Every thread of the workgroup calls it.
__device__ __forceinline__ uint32_t stage_blocks(const WalkArgs& a, DevBlock* sblk, uint32_t* svcut, uint16_t* slut)
{
    for (uint32_t i = threadIdx.x; i < a.nblocks; i += blockDim.x) sblk[i] = a.blocks[i];
    for (uint32_t i = threadIdx.x; i <= a.nblocks; i += blockDim.x) svcut[i] = a.vcut[i];
    __syncthreads();
    const uint32_t top = svcut[a.nblocks] ? svcut[a.nblocks] - 1 : 0;
    uint32_t shift = 0;
    while ((top >> shift) >= kLutSize) ++shift;
    for (uint32_t i = threadIdx.x; i < kLutSize; i += blockDim.x)
        slut[i] = (uint16_t)find_block(svcut, a.nblocks, i << shift);
    return shift;
}

// The visit-count array this workgroup adds into.  Workgroups are dispatched round robin
// over the XCDs, so with one array per XCD the counter lines stay in one XCD's L2 instead
// of moving between them (C2: the step kernel 157 -> 93 ms); totals are unchanged, the
// arrays are summed before they are read.
__device__ __forceinline__ uint32_t* visit_array(const WalkArgs& a)
{
    return a.visits ? a.visits + (uint64_t)(blockIdx.x % a.visits_parts) * a.visits_stride : nullptr;
}

struct Row {
    const uint32_t* adj;
    uint64_t deg;
};

__device__ __forceinline__ Row load_row(const DevBlock& b, uint32_t v)
{
    const uint64_t* bp = b.beg + (v - b.start_vert);
    uint64_t t = bp[1] - b.start_edge;
    uint64_t h = bp[0] - b.start_edge;
    if (t > b.nedges) t = b.nedges;
    if (h > t) h = t;
    return Row{ b.csr + h, t - h };
}


// Worklist length and work-queue chunk: from the host, or, after a filter pass on the
// same stream, from device memory (no host round trip); every wavefront derives the same
// chunk (64..1024, a multiple of 64) so that each still gets several chunks.
__device__ __forceinline__ void work_size(const WalkArgs& a, uint32_t& nwork, uint32_t& chunk)
{
    if (!a.nwork_dev) {
        nwork = a.nwork;
        chunk = a.chunk;
        return;
    }
    nwork = *a.nwork_dev;
    if (nwork > a.nwork) nwork = a.nwork;   // the filter counts runnable walkers past the worklist's end
    const uint64_t waves = (uint64_t)gridDim.x * blockDim.x / 64;
    uint64_t ch = nwork / (waves * 8 + 1);
    ch = ch < 64 ? 64 : (ch > 1024 ? 1024 : ch);
    chunk = (uint32_t)(ch / 64 * 64);
}

// Wave-local work queue: a wavefront takes a chunk of walker indices with one atomic and
// hands them to its lanes as they free up, so the shared counter sees one atomic per
// chunk instead of one per walker.  The host sizes the chunk (64..1024) so that every
// wavefront still gets several chunks.  `next`/`end` stay wavefront-uniform.
struct WaveQueue {
    uint32_t next, end;
};

__device__ __forceinline__ uint32_t wave_take(WaveQueue& q, uint64_t mask, uint32_t* counter, uint32_t lane,
                                              uint32_t kChunk)
{
    const uint32_t cnt = (uint32_t)__popcll(mask);
    const uint32_t rank = lane_rank(mask);
    const uint32_t avail = q.end - q.next;
    if (avail >= cnt) {
        const uint32_t idx = q.next + rank;
        q.next += cnt;
        return idx;
    }
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t nb = 0;
    if ((int)lane == leader) nb = atomicAdd(counter, kChunk);
    nb = __shfl(nb, leader, 64);
    const uint32_t idx = rank < avail ? q.next + rank : nb + (rank - avail);
    q.next = nb + (cnt - avail);
    q.end = nb + kChunk;
    return idx;
}

// Persistent step kernel.  Each lane owns one walker at a time and takes one whole step
// per loop iteration; a lane whose walker ends or leaves the resident set takes the
// next one from the work queue (one atomic per wavefront), so a wavefront never waits
// for its longest walker -- only for its slowest step.  ONE: a single resident block,
// held in registers.  LB: a block table of at most kLdsBlocks entries, copied with its
// cut points into LDS, so that the descriptor reads and the vertex->block search a step
// across blocks makes stay on chip (the table does not change during a launch).  MIX:
// two walker types; with one, the type's parameters are wavefront-uniform (SGPRs).
template <bool ONE, uint32_t SM, bool LB, bool MIX>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu((SM & 2) ? 4 : 8))) walk_kernel(WalkArgs a)
{
    // block-access trace: steps per current block, histogrammed in LDS
    constexpr uint32_t kTrace = ONE ? 1u : (LB ? kLdsBlocks : kTraceLdsBlocks);
    __shared__ uint32_t trace[kTrace];
    __shared__ DevBlock sblk[LB ? kLdsBlocks : 1];
    __shared__ uint32_t svcut[LB ? kLdsBlocks + 1 : 1];
    __shared__ uint16_t slut[LB ? kLutSize : 1];
    const bool lds_trace = a.block_steps && a.nblocks <= kTrace;
    const uint32_t lut_shift = LB ? stage_blocks(a, sblk, svcut, slut) : 0u;
    if (a.block_steps)
        for (uint32_t i = threadIdx.x; i < kTrace; i += blockDim.x) trace[i] = 0;
    if (LB || a.block_steps) __syncthreads();
    const DevBlock* blocks = LB ? sblk : a.blocks;
    const uint32_t* vcut = LB ? svcut : a.vcut;
    uint32_t* const visits = visit_array(a);
    Cnt c = {0, 0, 0, 0, 0, 0, 0};
    const uint32_t lane = __lane_id();
    DevBlock one;
    if (ONE) one = a.blocks[0];
    bool have = false, exhausted = false;
    uint32_t w = 0, cur = 0, prev = 0, hop = 0, type = 0, bc = 0, bp = 0;
    uint64_t hash = 0;
    Row rv{nullptr, 0}, ru{nullptr, 0};
    const uint32_t k0 = a.seed_lo, k1 = a.seed_hi;
    WaveQueue q{0, 0};
    uint32_t nwork, chunk;
    work_size(a, nwork, chunk);

    for (;;) {
        const bool need = !have && !exhausted;
        const uint64_t mask = __ballot(need);
        if (mask) {
            const uint32_t idx = wave_take(q, mask, a.queue, lane, chunk);
            if (need) {
                const uint32_t m = idx < nwork ? a.meta[a.worklist ? a.worklist[idx] : idx] : 0u;
                if (idx >= nwork) {
                    exhausted = true;
                } else if (((m >> kStatusShift) & 0xFu) == kActive) {
                    w = a.worklist ? a.worklist[idx] : idx;
                    cur = a.cur[w];
                    prev = a.prev[w];
                    hop = m & kHopMask;
                    type = (m >> kTypeShift) & 1u;
                    if (!ONE) {   // the pair's blocks, from the cut points (no stored copy)
                        bc = LB ? find_block_lut(slut, lut_shift, vcut, a.nblocks, cur) : find_block(vcut, a.nblocks, cur);
                        bp = prev == cur ? bc
                             : (LB ? find_block_lut(slut, lut_shift, vcut, a.nblocks, prev) : find_block(vcut, a.nblocks, prev));
                    }
                    if (a.hash) hash = a.hash[w];
                    if (ONE) {
                        rv = load_row(one, cur);
                        ru = load_row(one, prev);
                    } else {
                        rv = load_row(blocks[bc], cur);
                        ru = load_row(blocks[bp], prev);
                    }
                    have = true;
                }
            }
        }
        if (!__any(have || !exhausted)) break;
        if (!have) continue;

        const DevType& tp = (MIX && type) ? a.type1 : a.type0;
        uint32_t status = kActive;
        bool release = false;
        if (tp.t_term) {
            const uint4 d = philox(w, hop, 0, kTagTerm, k0, k1);
            if ((uint64_t)d.x < tp.t_term) status = kDoneTerm;
        }
        if (status == kActive && rv.deg == 0) status = kDoneDeadEnd;
        if (status == kActive) {
            uint32_t x = 0;
            const uint32_t t0 = c.trials, p0 = c.probes;
            const bool ok = tp.algo == 0
                                ? n2v_step<SM>(rv.adj, rv.deg, ru.adj, ru.deg, prev, tp, w, hop, k0, k1, x, c,
                                                   a.search_end)
                                : sopr_step<SM>(rv.adj, rv.deg, ru.adj, ru.deg, prev, tp, w, hop, k0, k1, x, c,
                                                    a.search_end);
            if (a.zc_count) {
                // candidates come from cur's row, membership probes from prev's row
                const bool mv = (ONE ? one.resident : blocks[bc].resident) == kMapped;
                const bool mu = (ONE ? one.resident : blocks[bp].resident) == kMapped;
                c.zc += (mv ? 1u + (c.trials - t0) : 0u) + (mu ? c.probes - p0 : 0u);
            }
            if (!ok) {
                status = kDoneTrialCap;
            } else {
                if (a.block_steps) {
                    if (lds_trace) atomicAdd(trace + bc, 1u);
                    else atomicAdd(a.block_steps + bc, 1ull);
                }
                prev = cur;
                cur = x;
                ++hop;
                ++c.steps;
                hash = (hash ^ x) * kFnvPrime;
                if (a.visits) atomicAdd(visits + x, 1u);
                if (a.paths) a.paths[(uint64_t)w * a.path_stride + hop] = x;
                ru = rv;
                bp = bc;
                if (hop >= tp.length) {
                    status = kDoneLength;
                } else if (ONE) {
                    rv = load_row(one, cur);
                } else {
                    const DevBlock& cb = blocks[bc];
                    if (cur - cb.start_vert >= cb.nverts)
                        bc = LB ? find_block_lut(slut, lut_shift, vcut, a.nblocks, cur) : find_block(vcut, a.nblocks, cur);
                    const DevBlock& nb = blocks[bc];
                    if (nb.resident) rv = load_row(nb, cur);
                    else release = true;
                }
            }
        }
        if (status != kActive || release) {
            a.cur[w] = cur;
            a.prev[w] = prev;
            a.meta[w] = hop | (status << kStatusShift) | (type << kTypeShift);
            if (a.hash) a.hash[w] = hash;
            if (status != kActive) ++c.ended; else ++c.left;
            have = false;
        }
    }
    if (lds_trace) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < a.nblocks; i += blockDim.x)
            if (trace[i]) atomicAdd(a.block_steps + i, (unsigned long long)trace[i]);
    }
    const uint64_t s0 = wave_sum(c.steps), s1 = wave_sum(c.trials), s2 = wave_sum(c.probes);
    const uint64_t s3 = wave_sum(c.mix), s4 = wave_sum(c.ended), s5 = wave_sum(c.left);
    const uint64_t s7 = a.zc_count ? wave_sum(c.zc) : 0;
    if (lane == 0) {
        atomicAdd(a.counters + 0, (unsigned long long)s0);
        atomicAdd(a.counters + 1, (unsigned long long)s1);
        atomicAdd(a.counters + 2, (unsigned long long)s2);
        atomicAdd(a.counters + 3, (unsigned long long)s3);
        atomicAdd(a.counters + 4, (unsigned long long)s4);
        atomicAdd(a.counters + 5, (unsigned long long)s5);
        if (s7) atomicAdd(a.counters + 7, (unsigned long long)s7);
    }
}

// ------------------------------------------------------------------------------------
// Flattened step kernel (launch variant 1).  The same walk, restated as a per-lane state
// machine in which every loop iteration performs exactly one memory round trip for every
// busy lane: all lanes' next addresses are computed, one load instruction (two for an
// adjacency row) goes out for the whole wavefront, and each lane consumes its value and
// advances through compute-only states until it needs the next load.  A lane inside a
// long membership search or SOPR intersection therefore no longer holds the other 63
// lanes of its wavefront at the end of every step.  Trials, probes and draws are those of
// the textbook loops above, so results and counters are bit-identical to variant 0.
// Only used when every resident block lives in an HBM slot: a u32 is read through the
// 8-byte-aligned word that holds it, which stays inside a slot's padded buffer.
// ------------------------------------------------------------------------------------
enum : uint32_t {
    PH_IDLE = 0,
    // load states
    PH_ROW, PH_CAND, PH_PROBE, PH_MIX_U, PH_MIX_LB, PH_MIX_EQ, PH_MIX_UB, PH_PICK, PH_PROP_U, PH_PPROBE,
    // compute-only states
    PH_STEP, PH_TRIAL, PH_PTRIAL, PH_MIX_NEXT, PH_AFTER_LB, PH_AFTER_UB, PH_MOVE, PH_END
};

struct FlatLane {
    uint32_t phase, w, cur, prev, hop, type, bc, bp, status;
    bool release;
    uint64_t hash;
    const uint32_t* nv;
    const uint32_t* nu;
    uint64_t dv, du;
    // trial / search state
    uint32_t k, x, y, last;
    uint4 d;
    uint64_t ra;
    int64_t lo, hi, mid;
    uint64_t i, base, lb, cnt, sel, c, pos;
    bool have_last, select;
    // pending load
    const void* addr;
};

__device__ __forceinline__ uint32_t word32(uint64_t v, const void* exact)
{
    return ((uintptr_t)exact & 4) ? (uint32_t)(v >> 32) : (uint32_t)v;
}

template <bool ONE>
__device__ __forceinline__ void flat_advance(FlatLane& L, const WalkArgs& a, const DevBlock& one, Cnt& c,
                                             uint32_t* trace, bool lds_trace)
{
    const uint32_t k0 = a.seed_lo, k1 = a.seed_hi;
    for (;;) {
        const DevType& tp = L.type ? a.type1 : a.type0;
        switch (L.phase) {
        case PH_STEP: {
            if (tp.t_term) {
                const uint4 t = philox(L.w, L.hop, 0, kTagTerm, k0, k1);
                if ((uint64_t)t.x < tp.t_term) { L.status = kDoneTerm; L.phase = PH_END; break; }
            }
            if (L.dv == 0) { L.status = kDoneDeadEnd; L.phase = PH_END; break; }
            if (tp.algo == 1 && tp.sampler == 0) {
                L.k = 0;
                L.phase = PH_PTRIAL;
                break;
            }
            if (tp.algo == 1 && sopr_use_mixture(L.du, L.dv, tp.alpha_a, tp.alpha_b)) {
                ++c.mix;
                ++c.trials;
                L.d = philox(L.w, L.hop, 0, kTagStep, k0, k1);
                L.select = false;
                L.i = 0; L.base = 0; L.c = 0; L.have_last = false;
                L.phase = PH_MIX_NEXT;
            } else {
                L.k = 0;
                L.phase = PH_TRIAL;
            }
            break;
        }
        case PH_TRIAL: {
            if (L.k >= kMaxTrials) { L.status = kDoneTrialCap; L.phase = PH_END; break; }
            if (!(L.k & 1)) L.d = philox(L.w, L.hop, L.k >> 1, kTagStep, k0, k1);
            const uint32_t rp = (L.k & 1) ? L.d.z : L.d.x;
            L.ra = (L.k & 1) ? L.d.w : L.d.y;
            ++c.trials;
            L.addr = L.nv + (((uint64_t)rp * L.dv) >> 32);
            L.phase = PH_CAND;
            return;
        }
        case PH_PTRIAL: {
            if (L.k >= kMaxTrials) { L.status = kDoneTrialCap; L.phase = PH_END; break; }
            L.d = philox(L.w, L.hop, L.k, kTagStep, k0, k1);
            ++c.trials;
            if ((uint64_t)L.d.x < ((uint64_t)tp.alpha_b << 16)) {
                L.addr = L.nv + (((uint64_t)L.d.y * L.dv) >> 32);
                L.phase = PH_PICK;
            } else {
                L.addr = L.nu + (((uint64_t)L.d.y * L.du) >> 32);
                L.phase = PH_PROP_U;
            }
            return;
        }
        case PH_MIX_NEXT: {
            if (L.i < L.du && L.base < L.dv) {
                L.addr = L.nu + L.i;
                ++c.probes;
                L.phase = PH_MIX_U;
                return;
            }
            if (!L.select) {
                L.cnt = L.c;
                const uint64_t A = tp.alpha_a, B = tp.alpha_b;
                if (L.cnt == 0 || (u128)L.d.x * (B * L.du + A * L.cnt) < ((u128)(B * L.du) << 32)) {
                    L.addr = L.nv + (((uint64_t)L.d.y * L.dv) >> 32);
                    L.phase = PH_PICK;
                    return;
                }
                L.sel = ((uint64_t)L.d.z * L.cnt) >> 32;
                L.select = true;
                L.i = 0; L.base = 0; L.c = 0; L.have_last = false;
                break;
            }
            // unreachable: the selection pass always finds sel < cnt
            L.status = kDoneTrialCap;
            L.phase = PH_END;
            break;
        }
        case PH_AFTER_LB: {
            if (L.lb < L.dv) {
                L.addr = L.nv + L.lb;
                ++c.probes;
                L.phase = PH_MIX_EQ;
                return;
            }
            L.base = L.lb;
            ++L.i;
            L.phase = PH_MIX_NEXT;
            break;
        }
        case PH_AFTER_UB: {
            const uint64_t ub = (uint64_t)L.lo, cv = ub - L.lb;
            if (L.select && L.sel < L.c + cv) {
                L.addr = L.nv + L.lb + (L.sel - L.c);
                L.phase = PH_PICK;
                return;
            }
            L.c += cv;
            L.base = ub;
            ++L.i;
            L.phase = PH_MIX_NEXT;
            break;
        }
        case PH_MOVE: {
            if (a.block_steps) {
                if (lds_trace) atomicAdd(trace + L.bc, 1u);
                else atomicAdd(a.block_steps + L.bc, 1ull);
            }
            L.prev = L.cur;
            L.cur = L.x;
            ++L.hop;
            ++c.steps;
            L.hash = (L.hash ^ L.x) * kFnvPrime;
            if (a.visits) atomicAdd(visit_array(a) + L.x, 1u);
            if (a.paths) a.paths[(uint64_t)L.w * a.path_stride + L.hop] = L.x;
            L.nu = L.nv;
            L.du = L.dv;
            L.bp = L.bc;
            if (L.hop >= tp.length) { L.status = kDoneLength; L.phase = PH_END; break; }
            const DevBlock* nb = &one;
            if (!ONE) {
                const DevBlock& cb = a.blocks[L.bc];
                if (L.cur - cb.start_vert >= cb.nverts) L.bc = find_block(a.vcut, a.nblocks, L.cur);
                nb = &a.blocks[L.bc];
                if (!nb->resident) { L.release = true; L.phase = PH_END; break; }
            }
            L.addr = nb->beg + (L.cur - nb->start_vert);
            L.phase = PH_ROW;
            return;
        }
        case PH_END: {
            a.cur[L.w] = L.cur;
            a.prev[L.w] = L.prev;
            a.meta[L.w] = L.hop | (L.status << kStatusShift) | (L.type << kTypeShift);
            if (a.hash) a.hash[L.w] = L.hash;
            if (L.status != kActive) ++c.ended; else ++c.left;
            L.phase = PH_IDLE;
            return;
        }
        default:
            return;   // a load state or idle
        }
    }
}

template <bool ONE>
__device__ __forceinline__ void flat_consume(FlatLane& L, uint64_t v0, uint64_t v1, const WalkArgs& a,
                                             const DevBlock& one, Cnt& c)
{
    const DevType& tp = L.type ? a.type1 : a.type0;
    switch (L.phase) {
    case PH_ROW: {
        const DevBlock& b = ONE ? one : a.blocks[L.bc];
        uint64_t t = v1 - b.start_edge, h = v0 - b.start_edge;
        if (t > b.nedges) t = b.nedges;
        if (h > t) h = t;
        L.nv = b.csr + h;
        L.dv = t - h;
        L.phase = PH_STEP;
        break;
    }
    case PH_CAND: {
        L.x = word32(v0, L.addr);
        if (tp.algo == 0) {
            if (L.ra < tp.t_min) { L.phase = PH_MOVE; break; }
            if (L.x == L.prev) {
                if (L.ra < tp.t_p) { L.phase = PH_MOVE; break; }
                ++L.k;
                L.phase = PH_TRIAL;
                break;
            }
            // outside the band between T(1) and T(1/q) membership cannot change the answer
            if (L.ra < (tp.t_one < tp.t_q ? tp.t_one : tp.t_q)) { L.phase = PH_MOVE; break; }
            if (L.ra >= (tp.t_one < tp.t_q ? tp.t_q : tp.t_one)) { ++L.k; L.phase = PH_TRIAL; break; }
        } else {
            const uint64_t A = tp.alpha_a, B = tp.alpha_b;
            if ((u128)L.ra * (B * L.du + A * L.dv) < ((u128)(B * L.du) << 32)) { L.phase = PH_MOVE; break; }
            if (L.x == L.prev) { ++L.k; L.phase = PH_TRIAL; break; }
        }
        L.lo = 0;
        L.hi = (int64_t)L.du - 1;
        if (L.lo > L.hi) {
            // empty prev list: not a member
            const bool acc = tp.algo == 0 ? L.ra < tp.t_q : false;
            if (acc) L.phase = PH_MOVE; else { ++L.k; L.phase = PH_TRIAL; }
            break;
        }
        L.mid = L.lo + ((L.hi - L.lo) >> 1);
        L.addr = L.nu + L.mid;
        ++c.probes;
        L.phase = PH_PROBE;
        return;
    }
    case PH_PROBE: {
        const uint32_t m = word32(v0, L.addr);
        bool done = false, found = false;
        if (m == L.x) { done = true; found = true; }
        else {
            if (m < L.x) L.lo = L.mid + 1; else L.hi = L.mid - 1;
            if (L.lo > L.hi) done = true;
        }
        if (!done) {
            L.mid = L.lo + ((L.hi - L.lo) >> 1);
            L.addr = L.nu + L.mid;
            ++c.probes;
            return;
        }
        bool acc;
        if (tp.algo == 0) acc = L.ra < (found ? tp.t_one : tp.t_q);
        else acc = found;
        if (acc) L.phase = PH_MOVE; else { ++L.k; L.phase = PH_TRIAL; }
        break;
    }
    case PH_MIX_U: {
        const uint32_t y = word32(v0, L.addr);
        if ((L.have_last && y == L.last) || y == L.prev) {
            L.last = y; L.have_last = true; ++L.i;
            L.phase = PH_MIX_NEXT;
            break;
        }
        L.last = y; L.have_last = true; L.y = y;
        L.lo = (int64_t)L.base;
        L.hi = (int64_t)L.dv;
        if (L.lo < L.hi) {
            L.mid = L.lo + ((L.hi - L.lo) >> 1);
            L.addr = L.nv + L.mid;
            ++c.probes;
            L.phase = PH_MIX_LB;
            return;
        }
        L.lb = (uint64_t)L.lo;
        L.phase = PH_AFTER_LB;
        break;
    }
    case PH_MIX_LB: {
        const uint32_t m = word32(v0, L.addr);
        if (m < L.y) L.lo = L.mid + 1; else L.hi = L.mid;
        if (L.lo < L.hi) {
            L.mid = L.lo + ((L.hi - L.lo) >> 1);
            L.addr = L.nv + L.mid;
            ++c.probes;
            return;
        }
        L.lb = (uint64_t)L.lo;
        L.phase = PH_AFTER_LB;
        break;
    }
    case PH_MIX_EQ: {
        const uint32_t m = word32(v0, L.addr);
        if (m == L.y) {
            L.lo = (int64_t)L.lb;
            L.hi = (int64_t)L.dv;
            L.mid = L.lo + ((L.hi - L.lo) >> 1);
            L.addr = L.nv + L.mid;
            ++c.probes;
            L.phase = PH_MIX_UB;
            return;
        }
        L.base = L.lb;
        ++L.i;
        L.phase = PH_MIX_NEXT;
        break;
    }
    case PH_MIX_UB: {
        const uint32_t m = word32(v0, L.addr);
        if (m <= L.y) L.lo = L.mid + 1; else L.hi = L.mid;
        if (L.lo < L.hi) {
            L.mid = L.lo + ((L.hi - L.lo) >> 1);
            L.addr = L.nv + L.mid;
            ++c.probes;
            return;
        }
        L.phase = PH_AFTER_UB;
        break;
    }
    case PH_PICK: {
        L.x = word32(v0, L.addr);
        L.phase = PH_MOVE;
        break;
    }
    case PH_PROP_U: {
        L.y = word32(v0, L.addr);
        if (L.y == L.prev) { ++L.k; L.phase = PH_PTRIAL; break; }
        L.lo = 0;
        L.hi = (int64_t)L.dv - 1;
        L.mid = L.lo + ((L.hi - L.lo) >> 1);
        L.addr = L.nv + L.mid;
        ++c.probes;
        L.phase = PH_PPROBE;
        return;
    }
    case PH_PPROBE: {
        const uint32_t m = word32(v0, L.addr);
        if (m == L.y) { L.x = L.y; L.phase = PH_MOVE; break; }
        if (m < L.y) L.lo = L.mid + 1; else L.hi = L.mid - 1;
        if (L.lo <= L.hi) {
            L.mid = L.lo + ((L.hi - L.lo) >> 1);
            L.addr = L.nv + L.mid;
            ++c.probes;
            return;
        }
        ++L.k;
        L.phase = PH_PTRIAL;
        break;
    }
    default:
        break;
    }
}

template <bool ONE>
__global__ void __launch_bounds__(256) walk_flat_kernel(WalkArgs a)
{
    constexpr uint32_t kTrace = ONE ? 1u : kTraceLdsBlocks;
    __shared__ uint32_t trace[kTrace];
    const bool lds_trace = a.block_steps && a.nblocks <= kTrace;
    if (a.block_steps) {
        for (uint32_t i = threadIdx.x; i < kTrace; i += blockDim.x) trace[i] = 0;
        __syncthreads();
    }
    Cnt c = {0, 0, 0, 0, 0, 0, 0};
    const uint32_t lane = __lane_id();
    DevBlock one;
    if (ONE) one = a.blocks[0];
    FlatLane L;
    L.phase = PH_IDLE;
    L.release = false;
    bool exhausted = false;
    WaveQueue q{0, 0};
    uint32_t nwork, chunk;
    work_size(a, nwork, chunk);
    for (;;) {
        const bool need = L.phase == PH_IDLE && !exhausted;
        const uint64_t mask = __ballot(need);
        if (mask) {
            const uint32_t idx = wave_take(q, mask, a.queue, lane, chunk);
            if (need) {
                const uint32_t m = idx < nwork ? a.meta[a.worklist ? a.worklist[idx] : idx] : 0u;
                if (idx >= nwork) {
                    exhausted = true;
                } else if (((m >> kStatusShift) & 0xFu) == kActive) {
                    L.w = a.worklist ? a.worklist[idx] : idx;
                    L.cur = a.cur[L.w];
                    L.prev = a.prev[L.w];
                    L.hop = m & kHopMask;
                    L.type = (m >> kTypeShift) & 1u;
                    L.bc = ONE ? 0u : find_block(a.vcut, a.nblocks, L.cur);
                    L.bp = ONE || L.prev == L.cur ? L.bc : find_block(a.vcut, a.nblocks, L.prev);
                    L.hash = a.hash ? a.hash[L.w] : 0;
                    const Row rv = ONE ? load_row(one, L.cur) : load_row(a.blocks[L.bc], L.cur);
                    const Row ru = ONE ? load_row(one, L.prev) : load_row(a.blocks[L.bp], L.prev);
                    L.nv = rv.adj; L.dv = rv.deg; L.nu = ru.adj; L.du = ru.deg;
                    L.status = kActive;
                    L.release = false;
                    L.phase = PH_STEP;
                    flat_advance<ONE>(L, a, one, c, trace, lds_trace);
                }
            }
        }
        const bool busy = L.phase != PH_IDLE;
        if (!__any(busy || !exhausted)) break;
        if (!busy) continue;
        // one round trip: an 8-byte aligned word (adjacency rows take both beg words)
        const uint64_t* p0 = (const uint64_t*)((uintptr_t)L.addr & ~(uintptr_t)7);
        const uint64_t v0 = *p0;
        uint64_t v1 = 0;
        if (L.phase == PH_ROW) v1 = *((const uint64_t*)L.addr + 1);
        flat_consume<ONE>(L, v0, v1, a, one, c);
        flat_advance<ONE>(L, a, one, c, trace, lds_trace);
    }
    if (lds_trace) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < a.nblocks; i += blockDim.x)
            if (trace[i]) atomicAdd(a.block_steps + i, (unsigned long long)trace[i]);
    }
    const uint64_t s0 = wave_sum(c.steps), s1 = wave_sum(c.trials), s2 = wave_sum(c.probes);
    const uint64_t s3 = wave_sum(c.mix), s4 = wave_sum(c.ended), s5 = wave_sum(c.left);
    if (lane == 0) {
        atomicAdd(a.counters + 0, (unsigned long long)s0);
        atomicAdd(a.counters + 1, (unsigned long long)s1);
        atomicAdd(a.counters + 2, (unsigned long long)s2);
        atomicAdd(a.counters + 3, (unsigned long long)s3);
        atomicAdd(a.counters + 4, (unsigned long long)s4);
        atomicAdd(a.counters + 5, (unsigned long long)s5);
    }
}

// ------------------------------------------------------------------------------------
// Lane kernel (launch variant 2).  Node2vec and the SOPR proposal sampler restated as a
// per-lane state machine: every loop iteration issues one round of loads for every busy
// lane -- picking up a walker (work item, walker state, both adjacency rows) included --
// then each lane consumes its value and runs compute-only states up to its next load.
// A lane deep in a membership search therefore no longer keeps the other lanes of its
// wavefront waiting, and a wavefront load carries a request for nearly every lane (the
// per-step loop of walk_kernel averages about 8 of 64 on RMAT graphs).  The state is
// 32-bit (degrees, search bounds, thresholds stored as T - 1), so the kernel keeps 8
// waves per SIMD.  Draws, trials, probes and moves are those of n2v_step and sopr_step
// (proposal sampler), so walks and counters are bit-identical to variant 0.  The host
// uses it only when every walker type is Node2vec or the SOPR proposal sampler and no
// block holds 2^32 or more edges.
// ------------------------------------------------------------------------------------
enum : uint32_t {
    LS_IDLE = 0,
    // load states (one round trip each)
    LS_FETCH,   // worklist[idx] -> walker id
    LS_STATE,   // meta, cur, prev, block pair (and hash) of the walker
    LS_ROWS,    // beg words of cur's and prev's rows
    LS_ROW,     // beg words of the new cur's row
    LS_CAND,    // Node2vec candidate N(cur)[pos]
    LS_PICK,    // SOPR: N(cur)[pos] is the move
    LS_PROPU,   // SOPR: proposal y = N(prev)[pos]
    LS_PROBE,   // membership probe
    // compute states
    LS_GUESS, LS_MOVE, LS_STEP, LS_TRIAL, LS_END
};

// Lane::fl bits; kLaneNoPrev: the previous vertex's block is not resident (its list unread)
constexpr uint32_t kLaneBisect = 1u, kLaneMv = 2u, kLaneMu = 4u, kLaneNoPrev = 8u;

struct Lane {
    uint32_t ph, w, cur, prev;
    uint32_t meta;                  // the stored word: hop | status << 20 | type << 24
    uint32_t bcp;                   // cur block | prev block << 16
    uint32_t k, x, ra, dz, dw;      // trial, candidate or search target, accept word, odd trial's draws
    uint32_t lo, hx, vlo, vhi, fl;  // search range [lo, hx), its value bounds, flags
    uint32_t dv, du;
    uint32_t blo, bhi;              // Node2vec: prev's positions in N(cur) lie in [blo, bhi)
    uint64_t hash;
    const uint32_t* nv;
    const uint32_t* nu;             // between LS_STATE and LS_ROWS: prev's beg words
    const void* addr;               // the pending load
};

__device__ __forceinline__ void lane_row(const DevBlock& b, uint64_t v0, uint64_t v1, const uint32_t*& adj,
                                         uint32_t& deg)
{
    uint64_t t = v1 - b.start_edge, h = v0 - b.start_edge;
    if (t > b.nedges) t = b.nedges;
    if (h > t) h = t;
    adj = b.csr + h;
    deg = (uint32_t)(t - h);
}

// next probe of the search for L.x in [L.lo, L.hx): member's positions
template <uint32_t SM>
__device__ __forceinline__ uint32_t lane_mid(const Lane& L)
{
    const uint32_t len = L.hx - L.lo;
    if (!(SM & 1) || (L.fl & kLaneBisect) || (!(SM & 2) && len <= 16) || L.vhi - L.vlo < 2)
        return L.lo + ((len - 1) >> 1);
    const uint32_t off = (uint32_t)((float)(L.x - L.vlo) * (float)len / (float)(L.vhi - L.vlo));
    return L.lo + (off < len ? off : len - 1);
}

// address of the next probe: the entry, or the window holding it
template <uint32_t SM>
__device__ __forceinline__ const void* lane_probe_addr(const Lane& L, const uint32_t* list)
{
    const uint32_t* p = list + lane_mid<SM>(L);
    return (SM & 2) ? (const void*)window_of(p) : (const void*)p;
}

// starts the membership search of L.x in a list of n entries (the first probe's address
// is computed in LS_GUESS); false = decided (absent) without a probe
template <uint32_t SM>
__device__ __forceinline__ bool lane_search(Lane& L, uint32_t n, uint32_t id_end)
{
    if (n == 0 || ((SM & 1) && L.x >= id_end)) return false;
    L.lo = 0;
    L.hx = n;
    L.vlo = 0;
    L.vhi = id_end;
    L.fl &= ~kLaneBisect;
    return true;
}

// RX: the worklist may hold walkers whose previous block is not resident (WalkArgs::relax):
// their steps stop where they need the previous vertex's list
template <bool ONE, uint32_t SM, bool LB, bool MIX, bool RX>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu((SM & 2) ? 4 : 8))) walk_lane_kernel(WalkArgs a)
{
    constexpr uint32_t kTrace = ONE ? 1u : (LB ? kLdsBlocks : kTraceLdsBlocks);
    __shared__ uint32_t trace[kTrace];
    __shared__ DevBlock sblk[LB ? kLdsBlocks : 1];
    __shared__ uint32_t svcut[LB ? kLdsBlocks + 1 : 1];
    __shared__ uint16_t slut[LB ? kLutSize : 1];
    const bool lds_trace = a.block_steps && a.nblocks <= kTrace;
    const uint32_t lut_shift = LB ? stage_blocks(a, sblk, svcut, slut) : 0u;
    if (a.block_steps)
        for (uint32_t i = threadIdx.x; i < kTrace; i += blockDim.x) trace[i] = 0;
    if (LB || a.block_steps) __syncthreads();
    const DevBlock* blocks = LB ? sblk : a.blocks;
    const uint32_t* vcut = LB ? svcut : a.vcut;
    uint32_t* const visits = visit_array(a);
    Cnt c = {0, 0, 0, 0, 0, 0, 0};
    const uint32_t lane = __lane_id();
    DevBlock one;
    if (ONE) one = a.blocks[0];
    uint32_t k0 = a.seed_lo, k1 = a.seed_hi;
    // The walker arrays, the worklist and the Philox key are used by some lane in almost
    // every iteration.  In the interpolated-window instance (the out-of-core default, run
    // at 4 waves per SIMD) they are held in VGPRs (the empty asm statement makes them
    // opaque), which leaves the SGPRs to the type parameters and exec masks: otherwise
    // those spill into VGPR lanes and come back through v_readlane in the loop (141 of the
    // loop's 959 VALU instructions; C3 step kernel -2%).  The copies cost ~35 VGPRs, so
    // the other instances keep the arguments in SGPRs: built for 64 VGPRs they would
    // spill, and the bisected-window instance keeps 7 waves per SIMD for resident graphs
    // (C2: 85 ms at 69 VGPRs, 93 ms at 101).
    uint32_t* w_cur = a.cur;
    uint32_t* w_prev = a.prev;
    uint32_t* w_meta = a.meta;
    uint64_t* w_hash = a.hash;
    const uint32_t* w_list = a.worklist;
    if constexpr (SM == 3)
        asm volatile("" : "+v"(w_cur), "+v"(w_prev), "+v"(w_meta), "+v"(w_hash), "+v"(w_list),
                     "+v"(k0), "+v"(k1));
    Lane L{};
    L.ph = LS_IDLE;
    bool exhausted = false;
    WaveQueue q{0, 0};
    uint32_t nwork, chunk;
    work_size(a, nwork, chunk);

    for (;;) {
        const bool need = L.ph == LS_IDLE && !exhausted;
        const uint64_t mask = __ballot(need);
        if (mask) {
            const uint32_t idx = wave_take(q, mask, a.queue, lane, chunk);
            if (need) {
                if (idx >= nwork) {
                    exhausted = true;
                } else if (w_list) {
                    L.addr = w_list + idx;
                    L.ph = LS_FETCH;
                } else {
                    L.w = idx;
                    L.ph = LS_STATE;
                }
            }
        }
        const bool busy = L.ph != LS_IDLE;
        if (!__any(busy || !exhausted)) break;
        if (!busy) continue;

        // ---- one round of loads
        uint64_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
        uint32_t v = 0;
        uint4 q0 = make_uint4(0, 0, 0, 0), q1 = q0, q2 = q0, q3 = q0;
        if ((SM & 2) && L.ph == LS_PROBE) {
            const uint4* q = (const uint4*)L.addr;   // one window: 64 aligned bytes
            q0 = q[0];
            q1 = q[1];
            q2 = q[2];
            q3 = q[3];
        } else if (L.ph == LS_STATE) {
            L.meta = w_meta[L.w];
            L.cur = w_cur[L.w];
            L.prev = w_prev[L.w];
            L.hash = w_hash ? w_hash[L.w] : 0;
        } else if (L.ph == LS_ROWS) {
            const uint64_t* pc = (const uint64_t*)L.addr;
            const uint64_t* pp = (const uint64_t*)L.nu;
            r0 = pc[0];
            r1 = pc[1];
            if (!RX || !(L.fl & kLaneNoPrev)) {
                r2 = pp[0];
                r3 = pp[1];
            }
        } else if (L.ph == LS_ROW) {
            const uint64_t* pc = (const uint64_t*)L.addr;
            r0 = pc[0];
            r1 = pc[1];
        } else {
            v = *(const uint32_t*)L.addr;   // worklist entry, candidate, proposal or probe
        }

        // ---- consume
        const DevType& tp = (MIX && ((L.meta >> kTypeShift) & 1u)) ? a.type1 : a.type0;
        switch (L.ph) {
        case LS_FETCH:
            L.w = v;
            L.ph = LS_STATE;
            break;
        case LS_STATE: {
            if (((L.meta >> kStatusShift) & 0xFu) != kActive) {
                L.ph = LS_IDLE;   // finished: left untouched
                break;
            }
            if (!ONE) {   // the pair's blocks, from the cut points (no stored copy)
                const uint32_t bc = LB ? find_block_lut(slut, lut_shift, vcut, a.nblocks, L.cur)
                                       : find_block(vcut, a.nblocks, L.cur);
                const uint32_t bp = L.prev == L.cur ? bc
                                    : (LB ? find_block_lut(slut, lut_shift, vcut, a.nblocks, L.prev)
                                          : find_block(vcut, a.nblocks, L.prev));
                L.bcp = bc | (bp << 16);
            } else {
                L.bcp = 0;
            }
            const DevBlock& B = ONE ? one : blocks[L.bcp & 0xFFFFu];
            const DevBlock& P = ONE ? one : blocks[L.bcp >> 16];
            L.addr = B.beg + (L.cur - B.start_vert);
            L.nu = (const uint32_t*)(P.beg + (L.prev - P.start_vert));
            // a walker picked up with its previous block not resident (the filter's relaxed
            // pickup) reads only its current row; its step stops where it needs the other
            L.fl = (B.resident == kMapped ? kLaneMv : 0u) | (P.resident == kMapped ? kLaneMu : 0u) |
                   (RX && !P.resident ? kLaneNoPrev : 0u);
            if (RX) L.meta &= ~kWaitPrev;
            L.ph = LS_ROWS;
            break;
        }
        case LS_ROWS:
            lane_row(ONE ? one : blocks[L.bcp & 0xFFFFu], r0, r1, L.nv, L.dv);
            if (RX && (L.fl & kLaneNoPrev)) {
                L.nu = nullptr;
                L.du = 0;
            } else {
                lane_row(ONE ? one : blocks[L.bcp >> 16], r2, r3, L.nu, L.du);
            }
            L.ph = LS_STEP;
            break;
        case LS_ROW:
            lane_row(ONE ? one : blocks[L.bcp & 0xFFFFu], r0, r1, L.nv, L.dv);
            L.ph = LS_STEP;
            break;
        case LS_CAND:   // Node2vec: accept iff ra < T(class); prev's list decides only inside the band
            L.x = v;
            if (SM & 2) {   // the read bounds prev's positions (sorted list)
                const uint32_t idx = (uint32_t)((const uint32_t*)L.addr - L.nv);
                if (v < L.prev) L.blo = idx + 1 > L.blo ? idx + 1 : L.blo;
                else if (v > L.prev) L.bhi = idx < L.bhi ? idx : L.bhi;
            }
            if (L.ra <= tp.m_min) {
                L.ph = LS_MOVE;
            } else if (L.x == L.prev) {
                L.ph = L.ra <= tp.m_p ? LS_MOVE : LS_TRIAL;
            } else if (L.ra <= tp.m_lo) {
                L.ph = LS_MOVE;
            } else if (L.ra > tp.m_hi) {
                L.ph = LS_TRIAL;
            } else if (RX && (L.fl & kLaneNoPrev)) {
                // the class needs prev's list, which is not resident: the step stops here,
                // its trials uncounted, and runs again from its first trial (same draws) once
                // prev's block is resident
                c.trials -= L.k + 1;
                L.meta |= kWaitPrev;
                L.ph = LS_END;
            } else if (lane_search<SM>(L, L.du, a.search_end)) {
                ++c.probes;
                if (a.zc_count && (L.fl & kLaneMu)) ++c.zc;
                L.ph = LS_GUESS;
            } else {
                L.ph = L.ra <= tp.m_q ? LS_MOVE : LS_TRIAL;
            }
            if (L.ph == LS_TRIAL) ++L.k;
            break;
        case LS_PICK:
            L.x = v;
            L.ph = LS_MOVE;
            break;
        case LS_PROPU:   // SOPR: a prev-neighbour y != prev is accepted iff y is in N(cur)
            L.x = v;
            if (L.x != L.prev && lane_search<SM>(L, L.dv, a.search_end)) {
                ++c.probes;
                if (a.zc_count && (L.fl & kLaneMv)) ++c.zc;
                L.ph = LS_GUESS;
            } else {
                ++L.k;
                L.ph = LS_TRIAL;
            }
            break;
        case LS_PROBE: {
            const bool sopr = tp.algo != 0;
            const uint32_t* list = sopr ? L.nv : L.nu;
            bool done = true, found;
            const uint32_t len = L.hx - L.lo;
            if (SM & 2) {
                // window [w0, w0 + 16) of the list; the entries [s, e) of it are in range
                const int64_t w0 = (const uint32_t*)L.addr - list;
                const int64_t s0 = w0 > (int64_t)L.lo ? w0 : (int64_t)L.lo;
                const int64_t e0 = w0 + 16 < (int64_t)L.hx ? w0 + 16 : (int64_t)L.hx;
                const WinScan r = scan_window(q0, q1, q2, q3, (uint32_t)(s0 - w0), (uint32_t)(e0 - w0), L.x);
                found = r.found;
                if (!found) {
                    if (L.x < r.first) {
                        L.hx = (uint32_t)s0;
                        L.vhi = r.first;
                        done = false;
                    } else if (L.x > r.last) {
                        L.lo = (uint32_t)e0;
                        L.vlo = r.last + 1;
                        done = false;
                    }
                }
            } else {
                const uint32_t mid = (uint32_t)((const uint32_t*)L.addr - list);
                found = v == L.x;
                if (!found) {
                    if (v < L.x) {
                        L.lo = mid + 1;
                        L.vlo = v + 1;
                    } else {
                        L.hx = mid;
                        L.vhi = v;
                    }
                    done = false;
                }
            }
            if (!done) {
                if (L.lo < L.hx) {
                    if (SM & 1) L.fl = ((uint64_t)(L.hx - L.lo) * 2 > len) ? (L.fl | kLaneBisect) : (L.fl & ~kLaneBisect);
                    L.ph = LS_GUESS;
                    ++c.probes;
                    if (a.zc_count && (L.fl & (sopr ? kLaneMv : kLaneMu))) ++c.zc;
                } else {
                    done = true;
                }
            }
            if (done) {
                const bool acc = sopr ? found : L.ra <= (found ? tp.m_one : tp.m_q);
                if (acc) {
                    L.ph = LS_MOVE;
                } else {
                    ++L.k;
                    L.ph = LS_TRIAL;
                }
            }
            break;
        }
        default:
            break;
        }

        // ---- compute states, in the order they can follow each other
        if (L.ph == LS_GUESS) {   // next probe of a membership search (one site for every phase)
            L.addr = lane_probe_addr<SM>(L, tp.algo != 0 ? L.nv : L.nu);
            L.ph = LS_PROBE;
        }
        if (L.ph == LS_MOVE) {
            const uint32_t bc = L.bcp & 0xFFFFu;
            if (a.block_steps) {
                if (lds_trace) atomicAdd(trace + bc, 1u);
                else atomicAdd(a.block_steps + bc, 1ull);
            }
            L.prev = L.cur;
            L.cur = L.x;
            ++L.meta;   // hop + 1 (hop < length <= 2^20 - 1)
            ++c.steps;
            L.hash = (L.hash ^ L.x) * kFnvPrime;
            if (a.visits) atomicAdd(visits + L.x, 1u);
            const uint32_t hop = L.meta & kHopMask;
            if (a.paths) a.paths[(uint64_t)L.w * a.path_stride + hop] = L.x;
            L.nu = L.nv;
            L.du = L.dv;
            L.fl = (L.fl & ~(kLaneMu | (RX ? kLaneNoPrev : 0u))) | ((L.fl & kLaneMv) ? kLaneMu : 0u);
            if (hop >= tp.length) {
                L.meta |= kDoneLength << kStatusShift;
                L.bcp = bc | (bc << 16);
                L.ph = LS_END;
            } else if (ONE) {
                L.bcp = bc | (bc << 16);
                L.addr = one.beg + (L.cur - one.start_vert);
                if (a.zc_count && (L.fl & kLaneMv)) ++c.zc;
                L.ph = LS_ROW;
            } else {
                uint32_t nbc = bc;
                const DevBlock& cb = blocks[bc];
                if (L.cur - cb.start_vert >= cb.nverts)
                    nbc = LB ? find_block_lut(slut, lut_shift, vcut, a.nblocks, L.cur) : find_block(vcut, a.nblocks, L.cur);
                const DevBlock& nb = blocks[nbc];
                L.bcp = nbc | (bc << 16);
                if (nb.resident) {
                    L.fl = (L.fl & ~kLaneMv) | (nb.resident == kMapped ? kLaneMv : 0u);
                    L.addr = nb.beg + (L.cur - nb.start_vert);
                    if (a.zc_count && (L.fl & kLaneMv)) ++c.zc;
                    L.ph = LS_ROW;
                } else {
                    L.ph = LS_END;   // steps into a block that is not resident: released, still active
                }
            }
        }
        if (L.ph == LS_STEP) {
            const uint32_t hop = L.meta & kHopMask;
            if (tp.t_term && (uint64_t)philox(L.w, hop, 0, kTagTerm, k0, k1).x < tp.t_term) {
                L.meta |= kDoneTerm << kStatusShift;
                L.ph = LS_END;
            } else if (L.dv == 0) {
                L.meta |= kDoneDeadEnd << kStatusShift;
                L.ph = LS_END;
            } else {
                L.k = 0;
                if (SM & 2) {
                    L.blo = 0;
                    L.bhi = L.dv;
                }
                L.ph = LS_TRIAL;
            }
        }
        // A Node2vec trial whose accept word lies above both classes x != prev decides on
        // x == prev alone; with sorted lists, a candidate position outside [blo, bhi) is not
        // prev (the candidates read in this step bound prev's positions), so such a trial is
        // rejected without a read and, in the window instances, the next trial is drawn in
        // the same iteration (once per iteration: C3's step kernel took 268 ms with one such
        // pass, 269 ms with two, 274 ms with four; the rule applies to about 0.8 of its 3.7
        // candidate reads per step).  Trials are counted as before.
        constexpr uint32_t kSkipPasses = (SM & 2) ? 1u : 0u;
        for (uint32_t pass = 0; pass <= kSkipPasses && L.ph == LS_TRIAL; ++pass) {
            if (L.k >= kMaxTrials) {
                L.meta |= kDoneTrialCap << kStatusShift;
                L.ph = LS_END;
            } else {
                const uint32_t hop = L.meta & kHopMask;
                ++c.trials;
                if (tp.algo == 0) {
                    uint32_t rp;
                    if (!(L.k & 1)) {
                        const uint4 d = philox(L.w, hop, L.k >> 1, kTagStep, k0, k1);
                        rp = d.x;
                        L.ra = d.y;
                        L.dz = d.z;
                        L.dw = d.w;
                    } else {
                        rp = L.dz;
                        L.ra = L.dw;
                    }
                    const uint32_t idx = (uint32_t)(((uint64_t)rp * L.dv) >> 32);
                    if (pass < kSkipPasses && L.ra > tp.m_hi && (L.ra > tp.m_p || idx < L.blo || idx >= L.bhi)) {
                        ++L.k;
                        continue;
                    }
                    L.addr = L.nv + idx;
                    if (a.zc_count && (L.fl & kLaneMv)) ++c.zc;
                    L.ph = LS_CAND;
                } else {
                    const uint4 d = philox(L.w, hop, L.k, kTagStep, k0, k1);
                    if (d.x <= tp.m_prop) {
                        L.addr = L.nv + (((uint64_t)d.y * L.dv) >> 32);
                        if (a.zc_count && (L.fl & kLaneMv)) ++c.zc;
                        L.ph = LS_PICK;
                    } else if (RX && (L.fl & kLaneNoPrev)) {
                        // the proposal draws from prev's list, which is not resident: the
                        // step stops here (as for a Node2vec class test above)
                        c.trials -= L.k + 1;
                        L.meta |= kWaitPrev;
                        L.ph = LS_END;
                    } else {
                        L.addr = L.nu + (((uint64_t)d.y * L.du) >> 32);
                        if (a.zc_count && (L.fl & kLaneMu)) ++c.zc;
                        L.ph = LS_PROPU;
                    }
                }
            }
        }
        if (L.ph == LS_END) {
            w_cur[L.w] = L.cur;
            w_prev[L.w] = L.prev;
            w_meta[L.w] = L.meta;
            if (w_hash) w_hash[L.w] = L.hash;
            if (((L.meta >> kStatusShift) & 0xFu) != kActive) ++c.ended; else ++c.left;
            L.ph = LS_IDLE;
        }
    }
    if (lds_trace) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < a.nblocks; i += blockDim.x)
            if (trace[i]) atomicAdd(a.block_steps + i, (unsigned long long)trace[i]);
    }
    const uint64_t s0 = wave_sum(c.steps), s1 = wave_sum(c.trials), s2 = wave_sum(c.probes);
    const uint64_t s4 = wave_sum(c.ended), s5 = wave_sum(c.left);
    const uint64_t s7 = a.zc_count ? wave_sum(c.zc) : 0;
    if (lane == 0) {
        atomicAdd(a.counters + 0, (unsigned long long)s0);
        atomicAdd(a.counters + 1, (unsigned long long)s1);
        atomicAdd(a.counters + 2, (unsigned long long)s2);
        atomicAdd(a.counters + 4, (unsigned long long)s4);
        atomicAdd(a.counters + 5, (unsigned long long)s5);
        if (s7) atomicAdd(a.counters + 7, (unsigned long long)s7);
    }
}

// Walker generation (engine/walk.hpp:32-49): walker w at sources[w] or w / wpv, with
// source = previous = current and hop 0; type w % 2 in a mixed batch.
__global__ void __launch_bounds__(256) init_kernel(InitArgs a)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < a.nwalkers; w += stride) {
        const uint32_t src = a.sources ? a.sources[w] : (uint32_t)(w / a.wpv);
        const uint32_t type = a.ntypes == 2 ? (uint32_t)(w & 1) : 0u;
        const uint32_t len = type ? a.len1 : a.len0;
        a.cur[w] = src;
        a.prev[w] = src;
        a.meta[w] = (len == 0 ? (kDoneLength << kStatusShift) : 0u) | (type << kTypeShift);
        if (a.hash) a.hash[w] = (kFnvOffset ^ src) * kFnvPrime;
        if (a.paths) {
            uint32_t* p = a.paths + w * a.path_stride;
            p[0] = src;
            for (uint32_t i = 1; i < a.path_stride; ++i) p[i] = 0xFFFFFFFFu;
        }
    }
}

// The bucket matrix is the count matrix that prefix() scans in gpuwalk.hpp:321-338; the
// worklist replaces insertglobal's scatter (gpuwalk.hpp:340-372).
constexpr uint32_t kFilterThreads = 1024, kFilterItems = 16;
constexpr uint32_t kFilterLdsCuts = 2048;   // cut points staged in LDS (with a lookup table) up to this many blocks

// Bucket / worklist pass over all walkers.  A workgroup takes a tile of
// kFilterThreads * kFilterItems walkers (coalesced: item k of thread t is walker
// tile + k * kFilterThreads + t), so each lane has 16 status words, then the vertex pairs
// of the unfinished ones, in flight before it waits.  A walker's (prev, cur) blocks come
// from its vertices through the cut points' lookup table in LDS, and the resident set is
// a bit table in LDS.  Unfinished walkers are
// counted into the (prev_blk, cur_blk) histogram (in LDS when the matrix fits) and, when a
// worklist is requested, the runnable ones are appended with one atomic per tile (ids past
// work_cap are counted but not written: the run takes another round for them).  An LDS
// histogram leaves the workgroup as one row of plain stores into bucket_part, summed by
// bucket_reduce_kernel: a device-scope atomic per bin and workgroup (1.5M of them on a
// 54-block table) made the pass 2x slower and held up host-to-device copies meanwhile.
// Two forms keep the matrix current across a run without a pass over all walkers: the
// worklist filter with bucket_sub counts the walkers it writes into the worklist (their
// pairs leave the matrix: bucket_reduce_kernel subtracts the rows), and the pass with
// gather visits only that worklist after the step kernel (the pairs of the unfinished ones
// come back: the rows are added).  GATHER and SUB are those forms (a.gather, a.bucket_sub),
// as template arguments so that the plain pass keeps its registers (80 VGPRs; 103 with the
// pairs that SUB holds until the worklist is written, 116 with every form in one kernel).
template <bool GATHER, bool SUB>
__global__ void __launch_bounds__(1024) filter_kernel(FilterArgs a)
{
    extern __shared__ uint32_t hist[];
    __shared__ uint32_t wave_cnt[kFilterThreads / 64];
    __shared__ uint32_t tile_base;
    __shared__ uint32_t res_bits[kFilterLdsResident / 32];
    __shared__ uint32_t svcut[kFilterLdsCuts + 1];
    __shared__ uint16_t slut[kLutSize];
    const uint32_t nbb = a.nblocks * a.nblocks;
    const bool lds_hist = (a.bucket || SUB) && nbb <= kFilterLdsBuckets;
    const bool lds_res = a.worklist && a.nblocks <= kFilterLdsResident;
    uint64_t nw = a.nwalkers;
    if (GATHER && *a.gather_n < nw) nw = *a.gather_n;
    const bool lds_cuts = a.nblocks <= kFilterLdsCuts;
    uint32_t lut_shift = 0;
    if (lds_cuts) {
        for (uint32_t i = threadIdx.x; i <= a.nblocks; i += blockDim.x) svcut[i] = a.vcut[i];
        __syncthreads();
        const uint32_t top = svcut[a.nblocks] ? svcut[a.nblocks] - 1 : 0;
        while ((top >> lut_shift) >= kLutSize) ++lut_shift;
        for (uint32_t i = threadIdx.x; i < kLutSize; i += blockDim.x)
            slut[i] = (uint16_t)find_block(svcut, a.nblocks, i << lut_shift);
    }
    if (lds_hist)
        for (uint32_t i = threadIdx.x; i < nbb; i += blockDim.x) hist[i] = 0;
    if (lds_res) {
        for (uint32_t i = threadIdx.x; i < (a.nblocks + 31) / 32; i += blockDim.x) res_bits[i] = 0;
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < a.nblocks; b += blockDim.x)
            if (a.blocks[b].resident) atomicOr(res_bits + (b >> 5), 1u << (b & 31));
    }
    __syncthreads();
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64, nwaves = blockDim.x / 64;
    const uint64_t tile = (uint64_t)blockDim.x * kFilterItems;
    uint64_t left = 0, hops = 0, tk = 0, tk_hops = 0;
    for (uint64_t t0 = (uint64_t)blockIdx.x * tile; t0 < nw; t0 += (uint64_t)gridDim.x * tile) {
        uint32_t m[kFilterItems], vc[kFilterItems], vp[kFilterItems], id[kFilterItems];
#pragma unroll
        for (uint32_t k = 0; k < kFilterItems; ++k) {
            const uint64_t j = t0 + (uint64_t)k * blockDim.x + threadIdx.x;
            id[k] = j < nw ? (GATHER ? a.gather[j] : (uint32_t)j) : 0u;
        }
#pragma unroll
        for (uint32_t k = 0; k < kFilterItems; ++k) {
            const uint64_t j = t0 + (uint64_t)k * blockDim.x + threadIdx.x;
            m[k] = j < nw ? a.meta[id[k]] : (kDoneLength << kStatusShift);
        }
#pragma unroll
        for (uint32_t k = 0; k < kFilterItems; ++k) {
            const bool act = ((m[k] >> kStatusShift) & 0xFu) == kActive;
            vc[k] = act ? a.cur[id[k]] : 0u;
            vp[k] = act ? a.prev[id[k]] : 0u;
        }
        uint64_t masks[kFilterItems];
        uint32_t pair[kFilterItems];
        uint32_t cnt = 0;
#pragma unroll
        for (uint32_t k = 0; k < kFilterItems; ++k) {
            bool take = false;
            if (((m[k] >> kStatusShift) & 0xFu) == kActive) {
                ++left;
                const uint32_t len = ((m[k] >> kTypeShift) & 1u) ? a.len1 : a.len0;
                const uint32_t hop = m[k] & kHopMask;
                hops += len > hop ? len - hop : 0u;
                const uint32_t bc = lds_cuts ? find_block_lut(slut, lut_shift, svcut, a.nblocks, vc[k])
                                             : find_block(a.vcut, a.nblocks, vc[k]);
                const uint32_t bp = vp[k] == vc[k] ? bc
                                    : (lds_cuts ? find_block_lut(slut, lut_shift, svcut, a.nblocks, vp[k])
                                                : find_block(a.vcut, a.nblocks, vp[k]));
                if (a.worklist) {
                    const bool rc = lds_res ? ((res_bits[bc >> 5] >> (bc & 31)) & 1u) != 0 : a.blocks[bc].resident != 0;
                    const bool rp = lds_res ? ((res_bits[bp >> 5] >> (bp & 31)) & 1u) != 0 : a.blocks[bp].resident != 0;
                    take = rc && (rp || (a.relax_prev && !(m[k] & kWaitPrev)));
                }
                if (SUB) pair[k] = bp * a.nblocks + bc;
                if (a.bucket) {
                    if (lds_hist) atomicAdd(hist + bp * a.nblocks + bc, 1u);
                    else atomicAdd(a.bucket + (uint64_t)bp * a.nblocks + bc, 1u);
                }
            }
            masks[k] = __ballot(take);
            cnt += (uint32_t)__popcll(masks[k]);
        }
        if (a.worklist) {   // uniform: one atomic per tile, waves write at their offsets
            if (lane == 0) wave_cnt[wave] = cnt;
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t tot = 0;
                for (uint32_t i = 0; i < nwaves; ++i) { const uint32_t c = wave_cnt[i]; wave_cnt[i] = tot; tot += c; }
                tile_base = tot ? atomicAdd(a.nwork, tot) : 0u;
            }
            __syncthreads();
            uint32_t off = tile_base + wave_cnt[wave];
#pragma unroll
            for (uint32_t k = 0; k < kFilterItems; ++k) {
                if ((masks[k] >> lane) & 1ull) {
                    const uint32_t at = off + lane_rank(masks[k]);
                    if (at < a.work_cap) {   // (no gather with a worklist: the id is the index)
                        a.worklist[at] = (uint32_t)(t0 + (uint64_t)k * blockDim.x + threadIdx.x);
                        if (SUB) {
                            if (lds_hist) atomicAdd(hist + pair[k], 1u);
                            else atomicSub(a.bucket_sub + pair[k], 1u);
                            const uint32_t len = ((m[k] >> kTypeShift) & 1u) ? a.len1 : a.len0;
                            const uint32_t hop = m[k] & kHopMask;
                            ++tk;
                            tk_hops += len > hop ? len - hop : 0u;
                        }
                    }
                }
                off += (uint32_t)__popcll(masks[k]);
            }
            __syncthreads();   // wave_cnt / tile_base are reused by the next tile
        }
    }
    left = wave_sum(left);
    if (lane == 0 && left) atomicAdd(a.left, (unsigned long long)left);
    if (a.hops) {
        hops = wave_sum(hops);
        if (lane == 0 && hops) atomicAdd(a.hops, (unsigned long long)hops);
    }
    if (SUB) {
        tk = wave_sum(tk);
        tk_hops = wave_sum(tk_hops);
        if (lane == 0 && tk) {
            atomicAdd(a.taken, (unsigned long long)tk);
            atomicAdd(a.taken + 1, (unsigned long long)tk_hops);
        }
    }
    if (lds_hist) {
        __syncthreads();
        if (a.bucket_part) {
            uint32_t* row = a.bucket_part + (uint64_t)blockIdx.x * nbb;
            for (uint32_t i = threadIdx.x; i < nbb; i += blockDim.x) row[i] = hist[i];
        } else if (SUB) {
            for (uint32_t i = threadIdx.x; i < nbb; i += blockDim.x)
                if (hist[i]) atomicSub(a.bucket_sub + i, hist[i]);
        } else {
            for (uint32_t i = threadIdx.x; i < nbb; i += blockDim.x)
                if (hist[i]) atomicAdd(a.bucket + i, hist[i]);
        }
    }
}

// bucket[i] = the sum of the filter workgroups' rows of bucket_part (integer sums: the
// same matrix in any order); mode 1 adds the sum to bucket[i], mode 2 subtracts it.
__global__ void __launch_bounds__(256) bucket_reduce_kernel(const uint32_t* __restrict__ part, uint32_t nrows,
                                                            uint32_t nbb, uint32_t* __restrict__ bucket, uint32_t mode)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbb) return;
    uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    uint32_t r = 0;
    for (; r + 4 <= nrows; r += 4) {
        s0 += part[(uint64_t)r * nbb + i];
        s1 += part[(uint64_t)(r + 1) * nbb + i];
        s2 += part[(uint64_t)(r + 2) * nbb + i];
        s3 += part[(uint64_t)(r + 3) * nbb + i];
    }
    for (; r < nrows; ++r) s0 += part[(uint64_t)r * nbb + i];
    const uint32_t sum = s0 + s1 + s2 + s3;
    bucket[i] = mode == 0 ? sum : (mode == 1 ? bucket[i] + sum : bucket[i] - sum);
}

// Per-block features of the learned scheduler, one thread per block b, from the bucket
// matrix W (prev-major) of unfinished walkers and the resident set:
//  0 log1p(row sum W[b,:])            1 log1p(column sum W[:,b])      2 log1p(W[b,b])
//  3 log1p(sum over resident j != b of W[b,j] + W[j,b])               4 resident flag
//  5 log1p(min over resident j != b of W[b,j] + W[j,b])  (0 if none)
//  6 log1p(plain add score: feature 3 + W[b,b])       7 age (iterations since last use) / 64
//  8 log1p(nedges) / 24    9 log1p(nverts) / 24    10 log1p(mean degree) / 8
// 11 fraction of all unfinished walkers touching b
__global__ void __launch_bounds__(256) features_kernel(const uint32_t* __restrict__ W, uint32_t nb,
                                                        const DevBlock* __restrict__ blocks,
                                                        const float* __restrict__ age, float ahead,
                                                        float* __restrict__ f)
{
    // one workgroup: the total of W is the sum of the row sums, reduced in LDS (integer
    // counts, exact in double in any order)
    __shared__ double part[256];
    double t = 0;
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x)
        for (uint32_t j = 0; j < nb; ++j) t += W[(uint64_t)b * nb + j];
    part[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    const double total = part[0];
    for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
        double row = 0, col = 0, res = 0;
        double mn = -1;
        for (uint32_t j = 0; j < nb; ++j) {
            const double r = W[(uint64_t)b * nb + j], c = W[(uint64_t)j * nb + b];
            row += r;
            col += c;
            if (j != b && blocks[j].resident == kInSlot) {
                res += r + c;
                if (mn < 0 || r + c < mn) mn = r + c;
            }
        }
        const double diag = W[(uint64_t)b * nb + b];
        const DevBlock& d = blocks[b];
        float* o = f + (uint64_t)b * kNumFeatures;
        o[0] = (float)log1p(row);
        o[1] = (float)log1p(col);
        o[2] = (float)log1p(diag);
        o[3] = (float)log1p(res);
        o[4] = d.resident == kInSlot ? 1.f : 0.f;
        o[5] = (float)log1p(mn < 0 ? 0.0 : mn);
        o[6] = (float)log1p(res + diag);
        o[7] = age ? age[b] / 64.f : 0.f;
        o[8] = (float)(log1p((double)d.nedges) / 24.0);
        o[9] = ahead;   // run phase: mean fraction of the hop cap still ahead of unfinished walkers
        o[10] = (float)(log1p(d.nverts ? (double)d.nedges / d.nverts : 0.0) / 8.0);
        o[11] = total > 0 ? (float)((row + col - diag) / total) : 0.f;
    }
}

hipError_t launch_features(const uint32_t* bucket, uint32_t nblocks, const DevBlock* blocks, const float* age, float ahead,
                           float* feats, hipStream_t s)
{
    hipLaunchKernelGGL(features_kernel, dim3(1), dim3(256), 0, s, bucket, nblocks, blocks, age, ahead, feats);
    return hipGetLastError();
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

hipError_t launch_init(const InitArgs& a, hipStream_t s)
{
    uint64_t blocks = (a.nwalkers + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(init_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_filter(const FilterArgs& a, uint32_t grid, hipStream_t s)
{
    if (a.nwalkers == 0) return hipSuccess;
    const uint64_t need = (a.nwalkers + kFilterThreads * kFilterItems - 1) / (kFilterThreads * kFilterItems);
    if (grid == 0 || grid > need) grid = (uint32_t)(need < 65536 ? need : 65536);
    const uint32_t nbb = a.nblocks * a.nblocks;
    const bool lds_hist = (a.bucket || a.bucket_sub) && nbb <= kFilterLdsBuckets;
    FilterArgs f = a;
    if (!lds_hist || a.part_rows == 0) f.bucket_part = nullptr;
    if (f.bucket_part && grid > a.part_rows) grid = a.part_rows;
    const size_t lds = lds_hist ? nbb * 4 : 0;
    if (f.gather)
        hipLaunchKernelGGL((filter_kernel<true, false>), dim3(grid), dim3(kFilterThreads), lds, s, f);
    else if (f.bucket_sub)
        hipLaunchKernelGGL((filter_kernel<false, true>), dim3(grid), dim3(kFilterThreads), lds, s, f);
    else
        hipLaunchKernelGGL((filter_kernel<false, false>), dim3(grid), dim3(kFilterThreads), lds, s, f);
    if (f.bucket_part)
        hipLaunchKernelGGL(bucket_reduce_kernel, dim3((nbb + 255) / 256), dim3(256), 0, s, f.bucket_part, grid, nbb,
                           f.bucket_sub ? f.bucket_sub : f.bucket, f.bucket_sub ? 2u : (f.gather ? 1u : 0u));
    return hipGetLastError();
}

// Step kernel instances: variant x single block x search order x LDS table x walker types.
template <bool ONE, uint32_t SM, bool LB, bool MIX>
static void launch_shape(const WalkArgs& a, uint32_t variant, uint32_t grid, uint32_t block, hipStream_t s)
{
    if (variant == 2) {
        if constexpr (!ONE) {
            if (a.relax) {
                hipLaunchKernelGGL((walk_lane_kernel<ONE, SM, LB, MIX, true>), dim3(grid), dim3(block), 0, s, a);
                return;
            }
        }
        hipLaunchKernelGGL((walk_lane_kernel<ONE, SM, LB, MIX, false>), dim3(grid), dim3(block), 0, s, a);
    } else {
        hipLaunchKernelGGL((walk_kernel<ONE, SM, LB, MIX>), dim3(grid), dim3(block), 0, s, a);
    }
}

template <bool ONE, uint32_t SM, bool LB>
static void launch_types(const WalkArgs& a, uint32_t variant, uint32_t grid, uint32_t block, hipStream_t s)
{
    if (a.ntypes == 2) launch_shape<ONE, SM, LB, true>(a, variant, grid, block, s);
    else launch_shape<ONE, SM, LB, false>(a, variant, grid, block, s);
}

template <uint32_t SM>
static void launch_search(const WalkArgs& a, bool one_block, uint32_t variant, uint32_t grid, uint32_t block,
                          hipStream_t s)
{
    if (one_block) launch_types<true, SM, false>(a, variant, grid, block, s);
    else if (a.nblocks <= kLdsBlocks) launch_types<false, SM, true>(a, variant, grid, block, s);
    else launch_types<false, SM, false>(a, variant, grid, block, s);
}

hipError_t launch_walk(const WalkArgs& a, bool one_block, uint32_t variant, uint32_t grid,
                       uint32_t block, hipStream_t s)
{
    if (a.nwork == 0) return hipSuccess;
    if (variant == 1) {
        if (one_block) hipLaunchKernelGGL(walk_flat_kernel<true>, dim3(grid), dim3(block), 0, s, a);
        else hipLaunchKernelGGL(walk_flat_kernel<false>, dim3(grid), dim3(block), 0, s, a);
    } else if (a.search_end != 0 && a.search_win) {
        launch_search<3>(a, one_block, variant, grid, block, s);
    } else if (a.search_win) {
        launch_search<2>(a, one_block, variant, grid, block, s);
    } else if (a.search_end != 0) {
        launch_search<1>(a, one_block, variant, grid, block, s);
    } else {
        launch_search<0>(a, one_block, variant, grid, block, s);
    }
    return hipGetLastError();
}

uint32_t walk_blocks_per_cu(bool one_block, uint32_t nblocks, uint32_t variant, uint32_t block)
{
    int n = 0;
    hipError_t e;
    if (variant == 1)
        e = one_block ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, walk_flat_kernel<true>, (int)block, 0)
                      : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, walk_flat_kernel<false>, (int)block, 0);
    else if (variant == 2)
        e = one_block ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, walk_lane_kernel<true, false, false, true, false>, (int)block, 0)
            : nblocks <= kLdsBlocks
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, walk_lane_kernel<false, false, true, true, false>, (int)block, 0)
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, walk_lane_kernel<false, false, false, true, false>, (int)block, 0);
    else if (one_block)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, walk_kernel<true, false, false, true>, (int)block, 0);
    else if (nblocks <= kLdsBlocks)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, walk_kernel<false, false, true, true>, (int)block, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, walk_kernel<false, false, false, true>, (int)block, 0);
    return (e == hipSuccess && n > 0) ? (uint32_t)n : 1u;
}

// Sums the per-XCD visit arrays into the first and clears the others.
__global__ void fold_visits_kernel(uint32_t* __restrict__ v, uint64_t n, uint32_t parts)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t t = v[i];
        for (uint32_t p = 1; p < parts; ++p) {
            t += v[p * n + i];
            v[p * n + i] = 0;
        }
        v[i] = t;
    }
}

hipError_t launch_fold_visits(uint32_t* v, uint64_t n, uint32_t parts, hipStream_t s)
{
    if (parts <= 1 || n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fold_visits_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, v, n, parts);
    return hipGetLastError();
}

// Word copy by a kernel: moves small host<->device transfers through pinned, mapped host
// memory instead of the DMA engines, so they never queue behind large block uploads.
__global__ void copy_u32_kernel(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy_u32(uint32_t* dst, const uint32_t* src, uint64_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(copy_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

// The device's constant-rate clock into dst: kernel timing inside captured graphs, whose
// event records cannot be timed.
__global__ void timestamp_kernel(unsigned long long* dst)
{
    if (threadIdx.x == 0) *dst = wall_clock64();
}

hipError_t launch_timestamp(unsigned long long* dst, hipStream_t s)
{
    hipLaunchKernelGGL(timestamp_kernel, dim3(1), dim3(64), 0, s, dst);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

}  // namespace gw
