// schedule.hpp -- block schedulers (engine/schedule.hpp:38-365).
//
// A scheduler owns `buckets`, the set of blocks that should be resident, in slot-position
// order (scheduler::buckets, schedule.hpp:44).  Each call to schedule() sees the bucket
// matrix W[prev_blk * nblocks + cur_blk] of unfinished walkers and changes at most one
// block; it returns false when nothing is worth loading (the reference then finishes the
// remaining walkers by zero-copy, engine/engine.hpp:593-634).
#pragma once
#include <algorithm>
#include <cstdint>
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
