// tuner.hpp -- launch tuner of the out-of-core engine (no reference counterpart: the
// reference hard-codes 512 x 512 threads, apps/Node2vec.cu:47-48).  Timing a candidate
// launch on a sample needs a resident cache, which an out-of-core walk only has near its
// end, and short samples favour shapes that lose on whole walks; so the tuner times whole
// walks instead: the first walk of a workload (same graph, walker count, cap and block
// size) runs the engine's default and is not timed (the first walk of a process or a
// workload runs slower: cold caches, first launches), successive walks each run the next
// untried candidate, and once every candidate has run, every later walk runs the one with
// the least step-kernel time per step.  Walks are
// bit-identical under every launch, so tuning changes time only.  The search order of
// the membership tests stays the engine's probe model's choice; a candidate fixes the
// workgroup size, the resident waves per CU (walkers in flight) and the kernel variant.
#pragma once
#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

#include "gw.h"

namespace gwh {

class LaunchTuner {
public:
    struct Entry {
        gw_launch launch;
        double ms_per_gstep = 0;   // step-kernel ms per 10^9 steps (mean over its walks)
        uint32_t walks = 0;
    };
    // workload key: (memory budget, block bytes, walkers, vertices)
    using Key = std::tuple<uint64_t, uint64_t, uint64_t, uint64_t>;

    // candidates in the order they are tried: the engine's default first, then the shapes
    // that came within 2% of it in a sweep of ten on Node2vec out of core (fewer waves per
    // CU, the per-step kernel and unwindowed searches were 10-100% slower)
    static std::vector<Entry> candidates()
    {
        const gw_launch c[] = {{256, 16, 2, 0}, {256, 24, 2, 0}, {512, 16, 2, 0}, {128, 16, 2, 0}};
        std::vector<Entry> v;
        for (const gw_launch& l : c) v.push_back(Entry{l, 0, 0});
        return v;
    }

    // the candidate the next walk of `k` runs
    uint32_t pick(const Key& k)
    {
        std::vector<Entry>& t = table(k);
        for (uint32_t i = 0; i < t.size(); ++i)
            if (!t[i].walks) return i;
        return best(k);
    }
    // times a walk of `k` with candidate i (the first walk of a workload only warms up)
    void record(const Key& k, uint32_t i, double kernel_ms, uint64_t steps)
    {
        std::vector<Entry>& t = table(k);
        if (!warm_[k]) {
            warm_[k] = true;
            return;
        }
        if (i >= t.size() || !steps) return;
        const double v = kernel_ms / ((double)steps * 1e-9);
        t[i].ms_per_gstep = (t[i].ms_per_gstep * t[i].walks + v) / (t[i].walks + 1);
        t[i].walks++;
    }
    // the fastest candidate timed so far (the first one if none is)
    uint32_t best(const Key& k)
    {
        std::vector<Entry>& t = table(k);
        uint32_t b = 0;
        bool have = false;
        for (uint32_t i = 0; i < t.size(); ++i)
            if (t[i].walks && (!have || t[i].ms_per_gstep < t[b].ms_per_gstep)) { b = i; have = true; }
        return b;
    }
    const std::vector<Entry>* find(const Key& k) const
    {
        auto it = tables_.find(k);
        return it == tables_.end() ? nullptr : &it->second;
    }
    const Key* last_key() const { return have_last_ ? &last_ : nullptr; }

private:
    std::vector<Entry>& table(const Key& k)
    {
        last_ = k;
        have_last_ = true;
        auto it = tables_.find(k);
        if (it == tables_.end()) it = tables_.emplace(k, candidates()).first;
        return it->second;
    }
    std::map<Key, std::vector<Entry>> tables_;
    std::map<Key, bool> warm_;
    Key last_{};
    bool have_last_ = false;
};

}  // namespace gwh
