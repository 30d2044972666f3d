// schedule.hpp -- block schedulers (engine/schedule.hpp:38-365).
//
// A scheduler owns `buckets`, the set of blocks that should be resident, in slot-position
// order (scheduler::buckets, schedule.hpp:44).  Each call to schedule() sees the bucket
// matrix W[prev_blk * nblocks + cur_blk] of unfinished walkers and changes at most one
// block; it returns false when nothing is worth loading (the reference then finishes the
// remaining walkers by zero-copy, engine/engine.hpp:593-634).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <random>
#include <vector>

#include "gw.h"

namespace gwh {

struct SchedInput {
    const uint64_t* W;          // nblocks x nblocks
    uint32_t nblocks;
    uint32_t ncblock;           // cache slots
    uint64_t zero_threshold;
    const gw_block_desc* blocks;
    uint64_t iteration;
    const uint8_t* staged = nullptr;   // per block: 0 not staged, 1 copy in flight, 2 copy done
};

// What one iteration achieved; fed back to schedulers that learn.
struct IterationOutcome {
    uint64_t iteration;
    uint64_t steps;             // walk steps taken in the iteration
    uint64_t trials, probes;    // candidate draws and adjacency probes of those steps
    uint64_t h2d_bytes;         // block bytes uploaded for it
    double kernel_ms;
    std::vector<uint32_t> loaded;   // blocks uploaded for it
    uint64_t walkers_run = 0;   // walkers whose (prev, cur) pair was resident
    uint64_t zc_touches = 0;    // zero-copy reads (the tail run)
    bool zero_copy = false;     // the run read non-resident blocks from host memory
};

class Scheduler {
public:
    virtual ~Scheduler() {}
    // Sets `buckets`, the blocks resident for the next run; false = nothing is worth
    // loading (the engine then finishes the walk by zero-copy reads).
    virtual bool schedule(const SchedInput& in) = 0;
    // The blocks that should be staged while the next run goes on (copies overlapped with
    // the walk; a later schedule() commits them by putting them in `buckets`): staged
    // blocks it leaves out may be overwritten, new ones are copied into slots that hold no
    // wanted block.  At most `spare` = slots - |buckets| entries.
    virtual void prefetch(const SchedInput& in, uint32_t spare, std::vector<uint32_t>& out)
    {
        (void)in; (void)spare; (void)out;
    }
    virtual void feedback(const SchedInput& in, const IterationOutcome& out) { (void)in; (void)out; }
    // The walk is over (every walker finished).
    virtual void finish() {}
    virtual const char* name() const = 0;
    virtual uint32_t run_flags() const { return 0; }   // extra GW_RUN_* flags the scheduler needs
    std::vector<uint32_t> buckets;
};

// GOwalker_scheduler_t::max_choose_blocks (engine/schedule.hpp:70-138).  For each block c
// outside the resident set C:
//   C not full: score(c) = sum_{j in C} (W[c,j] + W[j,c]) + W[c,c]; load the best c;
//   C full:     m = argmin_j (W[c,j] + W[j,c]), score(c) = sum_j (...) - min; the best c
//               replaces m (the diagonal is not counted).
// Accept only if score > 0 and score > zero_threshold.  Ties: the reference's std::sort
// leaves them unspecified; the lowest block id wins here, so H2D bytes are reproducible.
class PlainScheduler : public Scheduler {
public:
    const char* name() const override { return "plain"; }

    // score and replacement position for candidate c (no side effects)
    static uint64_t score(const SchedInput& in, const std::vector<uint32_t>& C, uint32_t c, uint32_t* pos)
    {
        const uint64_t n = in.nblocks;
        uint64_t s = 0;
        *pos = 0;
        if (C.size() >= in.ncblock) {
            uint64_t mn = 0xffffffffULL, all = 0;
            for (uint32_t j = 0; j < C.size(); ++j) {
                uint64_t x = in.W[c * n + C[j]] + in.W[(uint64_t)C[j] * n + c];
                all += x;
                if (x < mn) { mn = x; *pos = j; }
            }
            s = all - mn;
        } else {
            for (uint32_t j : C) s += in.W[c * n + j] + in.W[(uint64_t)j * n + c];
            s += in.W[c * n + c];
        }
        return s;
    }

    bool schedule(const SchedInput& in) override
    {
        std::vector<char> cached(in.nblocks, 0);
        for (uint32_t b : buckets) cached[b] = 1;
        bool found = false;
        uint64_t best = 0;
        uint32_t best_c = 0, best_pos = 0;
        for (uint32_t c = 0; c < in.nblocks; ++c) {
            if (cached[c]) continue;
            uint32_t pos;
            uint64_t s = score(in, buckets, c, &pos);
            if (!found || s > best) { found = true; best = s; best_c = c; best_pos = pos; }
        }
        if (!found || best == 0 || best <= in.zero_threshold) return false;
        if (buckets.size() >= in.ncblock) buckets[best_pos] = best_c;
        else buckets.push_back(best_c);
        return true;
    }
};

// The SOWalker scheduler (SOwalker_scheduler_t::choose_blocks, engine/schedule.hpp:460-580):
// every iteration picks a whole resident set of `ncblock` blocks.  Start: the ncblock - 1
// blocks holding the most walkers as their current block (column sums of W, :480-488),
// plus the block that pairs the most walkers with them (:490-503).  Then simulated
// annealing (:518-578): swap a random member for a random non-member; the score of a set
// is the walkers whose (prev, cur) blocks both lie in it, per block of the set that is not
// resident yet; better sets are kept, worse ones with probability exp(dy / T) (when
// dy > 0 is not met but the new score is positive), T from 100 to 0.001 by x0.998 per
// swap.  It never finishes by zero-copy (its schedule() always returns true, :626-634).
// Differences, all deterministic: the annealing draws come from a seeded generator (the
// reference seeds rand() with the time, :530), the swap budget C(nblocks, ncblock) is
// computed without the int overflow of :444-456, ranking ties keep the lower block id,
// and a set in which no walker can step is completed with the blocks of the fullest
// bucket (the reference would choose it again forever); returning false (no set can make
// progress) hands the walk to the engine's zero-copy tail.
class SOWalkerScheduler : public Scheduler {
public:
    explicit SOWalkerScheduler(uint64_t seed = 42) : rng_(seed) {}
    const char* name() const override { return "sowalker"; }

    // Blocks ranked for the starting set: the first min(ncblock, nblocks) entries are it
    // (the ncblock - 1 with the most walkers as their current block, then the block
    // pairing the most walkers with those), the rest is the pool the annealing swaps from.
    static std::vector<uint32_t> initial_order(const SchedInput& in)
    {
        const uint32_t nb = in.nblocks;
        const uint64_t n = nb;
        const uint32_t k = std::min(in.ncblock, nb);
        std::vector<uint64_t> part(nb, 0);
        for (uint32_t p = 0; p < nb; ++p)
            for (uint32_t c = 0; c < nb; ++c) part[c] += in.W[p * n + c];
        std::vector<uint32_t> idx(nb);
        for (uint32_t b = 0; b < nb; ++b) idx[b] = b;
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return part[a] > part[b]; });
        if (k == 0) return idx;
        uint64_t most = 0;
        uint32_t best = k - 1;
        for (uint32_t pi = k - 1; pi < nb; ++pi) {
            uint64_t w = 0;
            for (uint32_t ci = 0; ci + 1 < k; ++ci) w += in.W[idx[pi] * n + idx[ci]] + in.W[(uint64_t)idx[ci] * n + idx[pi]];
            if (w > most) { most = w; best = pi; }
        }
        std::swap(idx[k - 1], idx[best]);
        return idx;
    }

    bool schedule(const SchedInput& in) override
    {
        const uint32_t nb = in.nblocks;
        const uint64_t n = nb;
        const uint32_t k = std::min(in.ncblock, nb);
        if (k == 0) return false;
        uint64_t total = 0;
        for (uint64_t i = 0; i < n * n; ++i) total += in.W[i];
        if (total == 0) return false;
        std::vector<uint32_t> idx = initial_order(in);
        std::vector<char> cached(nb, 0);
        for (uint32_t b : buckets) cached[b] = 1;
        std::vector<uint32_t> cand(idx.begin(), idx.begin() + k);
        uint64_t S = 0;
        uint32_t comm = 0;
        for (uint32_t p : cand) {
            comm += cached[p];
            for (uint32_t c : cand) S += in.W[p * n + c];
        }
        const double inf = std::numeric_limits<double>::infinity();
        double y = comm < k ? (double)S / (double)(k - comm) : inf;   // the reference divides by zero here
        if (k < nb) {
            // walkers paired with x inside the current set (both directions, diagonal once)
            auto pair_sum = [&](uint32_t x) {
                uint64_t t = 0;
                for (uint32_t j : cand)
                    if (j != x) t += in.W[(uint64_t)x * n + j] + in.W[(uint64_t)j * n + x];
                return t;
            };
            double T = 100.0;
            const uint64_t budget = swap_budget(nb, k);
            for (uint64_t it = 0; it < budget && T > 0.001; ++it) {
                const uint32_t pos = k + (uint32_t)(rng_() % (nb - k)), tp = (uint32_t)(rng_() % k);
                const uint32_t a = cand[tp], b = idx[pos];
                // S of the set with a swapped for b
                const uint64_t without_a = S - pair_sum(a) - in.W[(uint64_t)a * n + a];
                cand[tp] = b;
                const uint64_t S2 = without_a + pair_sum(b) + in.W[(uint64_t)b * n + b];
                const uint32_t comm2 = comm - cached[a] + cached[b];
                const double y2 = comm2 < k ? (double)S2 / (double)(k - comm2) : 0.0;
                bool keep = y2 > y;
                if (!keep) {
                    const double u = (double)(rng_() >> 11) * (1.0 / 9007199254740992.0);
                    keep = y2 > 0 && u < std::exp((y2 - y) / T);
                }
                if (keep) {
                    idx[pos] = a;   // a goes back to the pool in b's place
                    S = S2;
                    comm = comm2;
                    y = y2;
                } else {
                    cand[tp] = a;
                }
                T *= 0.998;
            }
        }
        if (S == 0) {
            // no walker can step in this set: take the blocks of the fullest bucket
            uint64_t bw = 0;
            uint32_t bp = 0, bc = 0;
            for (uint32_t p = 0; p < nb; ++p)
                for (uint32_t c = 0; c < nb; ++c)
                    if (in.W[p * n + c] > bw && (p == c || k >= 2)) { bw = in.W[p * n + c]; bp = p; bc = c; }
            if (bw == 0) return false;
            std::vector<uint32_t> s = {bp};
            if (bc != bp) s.push_back(bc);
            for (uint32_t b : idx) {
                if (s.size() >= k) break;
                if (b != bp && b != bc) s.push_back(b);
            }
            cand = s;
        }
        buckets = cand;
        return true;
    }

    // min(C(n, k), 2^40): the reference's loop bound, without its int overflow
    static uint64_t swap_budget(uint32_t n, uint32_t k)
    {
        const uint64_t cap = 1ull << 40;
        uint64_t r = 1;
        for (uint32_t i = 0; i < k; ++i) {
            r = r * (n - i) / (i + 1);
            if (r >= cap) return cap;
        }
        return r;
    }

private:
    std::mt19937_64 rng_;
};

// Slot assignment of GOwalker_scheduler_t::copyblocks (engine/schedule.hpp:144-197): a
// block already in a slot stays there; otherwise it takes the first free slot, else the
// first slot whose block is no longer wanted.  Slots may also hold a staged (prefetched,
// not yet resident) block; such a slot counts as free once its block is not wanted.
class SlotMap {
public:
    void reset(uint32_t nslots, uint32_t nblocks)
    {
        slot_blk.assign(nslots, -1);
        slot_stg.assign(nslots, -1);
        blk_slot.assign(nblocks, -1);
        blk_stg.assign(nblocks, -1);
    }
    // returns the slot to upload into, -1 if the block is already resident, -2 if every
    // slot holds a wanted block
    int place(uint32_t blk, const std::vector<uint32_t>& wanted)
    {
        if (blk_slot[blk] >= 0) return -1;
        const int s = pick(wanted);
        if (s < 0) return -2;
        assign((uint32_t)s, blk);
        return s;
    }
    // slot to stage `blk` into (free or unwanted), or -1
    int stage(uint32_t blk, const std::vector<uint32_t>& wanted)
    {
        if (blk_slot[blk] >= 0 || blk_stg[blk] >= 0) return -1;
        const int s = pick(wanted);
        if (s < 0) return -1;
        clear(s);
        slot_stg[s] = (int)blk;
        blk_stg[blk] = s;
        return s;
    }
    // the staged copy of blk becomes resident in its slot
    void commit(uint32_t blk)
    {
        const int s = blk_stg[blk];
        slot_stg[s] = -1;
        blk_stg[blk] = -1;
        slot_blk[s] = (int)blk;
        blk_slot[blk] = s;
    }
    uint32_t free_slots(const std::vector<uint32_t>& wanted) const
    {
        uint32_t n = 0;
        for (size_t s = 0; s < slot_blk.size(); ++s) n += usable((int)s, wanted) ? 1u : 0u;
        return n;
    }
    std::vector<int> slot_blk, slot_stg, blk_slot, blk_stg;

private:
    static bool in(const std::vector<uint32_t>& v, int b) { return b >= 0 && std::find(v.begin(), v.end(), (uint32_t)b) != v.end(); }
    bool usable(int s, const std::vector<uint32_t>& wanted) const
    {
        return !in(wanted, slot_blk[s]) && !in(wanted, slot_stg[s]);
    }
    int pick(const std::vector<uint32_t>& wanted) const
    {
        // first empty slot, then the first slot with an unwanted staged block, then the
        // first slot whose resident block is unwanted
        for (size_t s = 0; s < slot_blk.size(); ++s)
            if (slot_blk[s] < 0 && slot_stg[s] < 0) return (int)s;
        for (size_t s = 0; s < slot_blk.size(); ++s)
            if (slot_blk[s] < 0 && slot_stg[s] >= 0 && !in(wanted, slot_stg[s])) return (int)s;
        for (size_t s = 0; s < slot_blk.size(); ++s)
            if (slot_blk[s] >= 0 && !in(wanted, slot_blk[s])) return (int)s;
        return -1;
    }
    void clear(int s)
    {
        if (slot_blk[s] >= 0) blk_slot[slot_blk[s]] = -1;
        if (slot_stg[s] >= 0) blk_stg[slot_stg[s]] = -1;
        slot_blk[s] = -1;
        slot_stg[s] = -1;
    }
    void assign(uint32_t s, uint32_t blk)
    {
        clear((int)s);
        slot_blk[s] = (int)blk;
        blk_slot[blk] = (int)s;
    }
};

}  // namespace gwh
