// engine.hpp -- the out-of-core walk driver: GOWalker's graph_engine::pipe_run and
// pipeline (engine/engine.hpp:406-638) on top of libgw.so.
//
// Per iteration: the bucket matrix of unfinished walkers (left by the previous run, no
// extra pass), the scheduler's choice of resident blocks, placement into HBM slots
// (copyblocks, engine/schedule.hpp:144-197), uploads on the copy stream, optional
// prefetch of further blocks into free slots (copies that overlap the run), and one run
// of every walker whose (prev, cur) blocks are resident until it finishes or steps out.
// When the scheduler finds nothing worth loading, the remaining walkers finish by
// reading non-resident blocks from pinned host memory (the zero-copy tail of
// engine/engine.hpp:459-515, 593-634).  Walks do not depend on the schedule: every draw
// is keyed on (seed, walker, hop, draw), so only bytes moved and time change.
#pragma once
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "gw.h"
#include "schedule.hpp"

namespace gwh {

struct GwError : std::runtime_error {
    gw_status status;
    GwError(gw_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void ck(gw_status s, gw_ctx* c, const char* what)
{
    if (s != GW_OK) throw GwError(s, std::string(what) + ": " + gw_last_error(c));
}

// One block in the transfer encoding (pinned host memory), or data == nullptr: sent raw.
struct PackedBlock {
    const void* data = nullptr;
    uint64_t bytes = 0;
};

struct EngineConfig {
    int device = 0;
    uint64_t memory_bytes = 0;      // HBM budget for slots + walker state (0: every block fits)
    uint64_t zero_threshold = 0;
    bool zero_copy_tail = true;
    // block 0: 256 threads of `variant` (default: the lane kernel, which walks out-of-core
    // Node2vec fastest; other walker types run variant 0), search order from the probe model
    gw_launch launch{0, 0, 2, 0};
    // with a block given: its shape, but the search order from the probe model (the launch
    // tuner's candidates fix the shape only)
    bool auto_search = false;
    // block uploads in the transfer encoding (one entry per block), or nullptr: raw
    const std::vector<PackedBlock>* packed = nullptr;
    // copy lanes of packed uploads (gw_set_copy_lanes): 1 = one copy stream; 2 = two, with a
    // staging buffer of the largest block each when that costs no slot, else half of one
    // each.  One is the default: on C3 two lanes were no faster at 8 GB (3.10 vs 3.14 s) and
    // slower at 16 GB (380 vs 366 ms), where two copies sharing the link land the fill's
    // first blocks later
    uint32_t copy_lanes = 1;
    bool verbose = false;
};

struct WalkConfig {
    gw_walk_params params{};
    uint64_t nwalkers = 0;
    uint32_t wpv = 1;
    const uint32_t* sources = nullptr;
    uint32_t outputs = 0;
};

struct RunReport {
    uint64_t steps = 0, trials = 0, probes = 0, mix_steps = 0;
    uint64_t h2d_bytes = 0, h2d_raw_bytes = 0, uploads = 0, staged = 0, iterations = 0;
    uint64_t inplace_blocks = 0, ring_blocks = 0, inline_blocks = 0;   // packed blocks by decode path
    uint64_t kernel_launches = 0;   // step-kernel launches: iterations plus worklist-overflow rounds
    uint64_t zero_copy_runs = 0, zero_copy_steps = 0, zc_touches = 0;
    double wall_s = 0, kernel_ms = 0, filter_ms = 0, h2d_ms = 0, sched_ms = 0, tune_ms = 0;
    uint32_t nslots = 0, nblocks = 0, copy_lanes = 1;
    gw_launch tuned{0, 0, 0, 0};    // launch the autotuner picked (block 0: none)
};

// Probes of a bisection and of the kernel's interpolation order (same estimate rule)
// over `q` queries x in N(u), u the source of a random arc and x drawn from the list
// of a random neighbour of u -- the membership tests Node2vec makes.  Interpolation
// is preferred when it needs under 80% of the bisection's probes.
inline bool prefer_interpolation(const uint64_t* beg, const uint32_t* csr, uint64_t nv, uint32_t q = 4096)
{
    const uint64_t ne = nv ? beg[nv] - beg[0] : 0;
    if (!ne || nv > 0xFFFFFFFFull) return false;
    uint64_t z = 0x9E3779B97F4A7C15ULL, pb = 0, pi = 0;
    auto rnd = [&](uint64_t n) {
        z += 0x9E3779B97F4A7C15ULL;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
        return (x ^ (x >> 31)) % n;
    };
    auto src = [&](uint64_t e) {   // vertex whose list holds arc e
        uint64_t lo = 0, hi = nv;
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) / 2;
            if (beg[mid] <= e) lo = mid; else hi = mid;
        }
        return lo;
    };
    for (uint32_t i = 0; i < q; ++i) {
        const uint64_t e = beg[0] + rnd(ne);
        const uint64_t u = src(e);
        const uint32_t v = csr[e];
        if (v >= nv) return false;
        const uint64_t dv = beg[v + 1] - beg[v];
        if (!dv) continue;
        const uint32_t x = csr[beg[v] + rnd(dv)];
        const uint32_t* a = csr + beg[u];
        const int64_t n = (int64_t)(beg[u + 1] - beg[u]);
        int64_t lo = 0, hi = n - 1;
        while (lo <= hi) {
            const int64_t mid = lo + ((hi - lo) >> 1);
            ++pb;
            if (a[mid] == x) break;
            if (a[mid] < x) lo = mid + 1; else hi = mid - 1;
        }
        lo = 0;
        hi = n - 1;
        uint32_t vlo = 0, vhi = (uint32_t)nv;
        bool bisect = false;
        while (lo <= hi) {
            const uint32_t len = (uint32_t)(hi - lo + 1);
            int64_t mid;
            if (bisect || len <= 16 || vhi - vlo < 2) {
                mid = lo + ((hi - lo) >> 1);
            } else {
                const uint32_t off = (uint32_t)((float)(x - vlo) * (float)len / (float)(vhi - vlo));
                mid = lo + (off < len ? off : len - 1);
            }
            ++pi;
            if (a[mid] == x) break;
            if (a[mid] < x) { lo = mid + 1; vlo = a[mid] + 1; } else { hi = mid - 1; vhi = a[mid]; }
            bisect = (uint64_t)(hi - lo + 1) * 2 > len;
        }
    }
    return pb && pi * 5 < pb * 4;
}


class Engine {
public:
    // beg/csr: the whole graph in host memory (pinned for the zero-copy tail); blocks:
    // the block table over it.
    Engine(const EngineConfig& cfg, const uint64_t* beg, const uint32_t* csr, std::vector<gw_block_desc> blocks)
        : cfg_(cfg), beg_(beg), csr_(csr), blocks_(std::move(blocks))
    {
        gw_config gc{cfg.device, 0, cfg.memory_bytes};
        ck(gw_create(&gc, &ctx_), nullptr, "gw_create");
        ck(gw_set_block_table(ctx_, blocks_.data(), (uint32_t)blocks_.size()), ctx_, "gw_set_block_table");
    }
    ~Engine() { gw_destroy(ctx_); }
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    gw_ctx* ctx() { return ctx_; }
    uint64_t total_verts() const { return blocks_.empty() ? 0 : blocks_.back().start_vert + blocks_.back().nverts; }
    uint32_t nblocks() const { return (uint32_t)blocks_.size(); }

    // Number of HBM slots that fit next to the walker state under the budget.  The slots
    // this engine already holds count as free: a new reservation releases them first.
    uint32_t slots_for_budget()
    {
        const uint32_t nb = nblocks();
        if (!cfg_.memory_bytes) return nb;
        gw_counters c;
        ck(gw_counters_get(ctx_, &c), ctx_, "gw_counters_get");
        uint64_t per = 0;
        ck(gw_slot_bytes(ctx_, &per), ctx_, "gw_slot_bytes");
        const uint64_t mine = (uint64_t)nslots_ * per;
        const uint64_t held = c.hbm_bytes_held > mine ? c.hbm_bytes_held - mine : 0;
        uint64_t free_b = cfg_.memory_bytes > held ? cfg_.memory_bytes - held : 0;
        uint64_t n = free_b / per;
        return (uint32_t)(n < nb ? n : nb);
    }

    RunReport run(const WalkConfig& wc, Scheduler& sched)
    {
        using clk = std::chrono::steady_clock;
        auto t0 = clk::now();
        RunReport r;
        const uint32_t nb = nblocks();
        r.nblocks = nb;
        gw_counters c0;
        ck(gw_set_walk(ctx_, &wc.params), ctx_, "gw_set_walk");
        ck(gw_init_walkers(ctx_, wc.nwalkers, wc.wpv, wc.sources, wc.outputs), ctx_, "gw_init_walkers");
        // the staging buffer of packed uploads (the largest packed block) is counted against
        // the budget before slots.  Smaller buffers work too (payload in pieces), but each
        // piece adds a decode between copies on the copy stream: with 16 MB pieces the C3
        // 8 GB walk took 3.81 s instead of 3.43 s, and at that cap not even 25 MB of staging
        // left room for the slot a whole block's staging costs.
        // Two copy lanes (on request) let one lane's decode overlap the other's copy: with a
        // whole-block staging buffer each when the second buffer leaves the slot count as it
        // is, else sharing the one buffer (each lane sends its blocks in pieces of half of it).
        if (cfg_.packed && !pack_staging_) {
            uint64_t whole = 0, head = 0;
            for (uint32_t b = 0; b < nb; ++b)
                if (const PackedBlock* p = packed(b)) {
                    whole = std::max<uint64_t>(whole, p->bytes);
                    const uint32_t* w = (const uint32_t*)p->data;   // block header: magic, nvc, vpw
                    const uint64_t nch = (blocks_[b].nedges + 255) / 256;
                    head = std::max<uint64_t>(head, 4 * (4 + 4 * (uint64_t)w[1] + w[2] + 4 * nch));
                }
            whole = (whole + 255) & ~255ULL;
            if (whole) {
                ck(gw_set_copy_lanes(ctx_, 1), ctx_, "gw_set_copy_lanes");
                // one lane: the payloads land in their slots and decode in place, so the
                // buffer holds two head areas and a small ring for blocks that cannot
                // (C3: ~25 MB instead of the largest packed block, ~155 MB: one more slot at
                // 8 GB); GW_PACK_INPLACE=0 sends every payload through a ring of that size
                const char* ip = std::getenv("GW_PACK_INPLACE");
                // (head areas are a quarter each, rounded down to 256 bytes)
                const uint64_t one_lane = ip && std::atoi(ip) == 0
                                              ? whole
                                              : std::max<uint64_t>(4 * ((head + 511) & ~255ULL), 65536);
                ck(gw_set_pack_staging(ctx_, one_lane), ctx_, "gw_set_pack_staging");
                if (cfg_.copy_lanes == 2) {
                    const uint32_t one = slots_for_budget();
                    ck(gw_set_copy_lanes(ctx_, 2), ctx_, "gw_set_copy_lanes");
                    ck(gw_set_pack_staging(ctx_, 2 * whole), ctx_, "gw_set_pack_staging");
                    lanes_ = 2;
                    if (slots_for_budget() < one) {
                        // the halves must hold a block's headers plus its largest chunk
                        if (whole / 2 >= head + 1088 + 512) {
                            ck(gw_set_pack_staging(ctx_, whole), ctx_, "gw_set_pack_staging");
                        } else {
                            lanes_ = 1;
                            ck(gw_set_copy_lanes(ctx_, 1), ctx_, "gw_set_copy_lanes");
                            ck(gw_set_pack_staging(ctx_, one_lane), ctx_, "gw_set_pack_staging");
                        }
                    }
                }
            }
            pack_staging_ = true;
        }
        r.copy_lanes = lanes_;
        const uint32_t want = slots_for_budget();
        if (want == 0) throw GwError(GW_ERR_BUDGET, "the memory budget leaves no room for a single block slot");
        if (want != nslots_) {
            ck(gw_reserve_slots(ctx_, want), ctx_, "gw_reserve_slots");
            nslots_ = want;
        } else {
            for (uint32_t s = 0; s < nslots_; ++s) ck(gw_evict(ctx_, s), ctx_, "gw_evict");
        }
        r.nslots = nslots_;
        if (cfg_.verbose) {
            gw_counters hc;
            uint64_t per = 0;
            ck(gw_counters_get(ctx_, &hc), ctx_, "gw_counters_get");
            ck(gw_slot_bytes(ctx_, &per), ctx_, "gw_slot_bytes");
            std::fprintf(stderr, "[engine] budget %.3f GB: %u slots of %.1f MB, %.3f GB held, %.1f MB free\n",
                         cfg_.memory_bytes / 1e9, nslots_, per / 1e6, hc.hbm_bytes_held / 1e9,
                         cfg_.memory_bytes > hc.hbm_bytes_held ? (cfg_.memory_bytes - hc.hbm_bytes_held) / 1e6 : 0.0);
        }
        slots_.reset(nslots_, nb);
        sched.buckets.clear();
        ck(gw_counters_get(ctx_, &c0), ctx_, "gw_counters_get");
        ck(gw_upload_stats(ctx_, &up0_[0], &up0_[1], &up0_[2]), ctx_, "gw_upload_stats");
        uint64_t wcap = 0, rounds0 = 0;
        ck(gw_worklist_stats(ctx_, &wcap, &rounds0), ctx_, "gw_worklist_stats");
        std::vector<uint64_t> W((uint64_t)nb * nb);
        std::vector<uint8_t> staged(nb, 0);
        std::vector<uint32_t> pf;
        gw_launch base = cfg_.launch;
        bool pick_shape = base.block == 0 || cfg_.auto_search;
        // launch of the next run; on the first call with no shape given, the membership
        // search order is chosen from the probes both orders need on sampled Node2vec-style
        // queries of this graph (timing on the device would need a resident cache that is
        // not there yet).  That host work happens after the first copies are queued, so it
        // overlaps them.
        auto next_launch = [&]() {
            if (pick_shape) {
                auto tt = clk::now();
                // the lane kernel runs 16 waves per CU: each lane keeps a few adjacency lines
                // in use, and half the resident waves keep that working set nearer the L2 (out
                // of core: -2% walk time at a 16 GB cap, -13% step-kernel time at 8 GB)
                const uint32_t wpc = cfg_.launch.waves_per_cu ? cfg_.launch.waves_per_cu
                                                               : (cfg_.launch.variant == 2 ? 16u : 0u);
                // membership probes read 64-byte windows (a third of the reads of single
                // probes on C3: 1.9 vs 5.3 per step, -7% step-kernel time)
                base = gw_launch{cfg_.launch.block ? cfg_.launch.block : 256u, wpc, cfg_.launch.variant,
                                 (prefer_interpolation(beg_, csr_, total_verts()) ? (uint32_t)GW_RUN_INTERP : 0u) |
                                     (uint32_t)GW_RUN_WINDOW};
                r.tune_ms = std::chrono::duration<double, std::milli>(clk::now() - tt).count();
                r.tuned = base;
                pick_shape = false;
            }
            gw_launch l = base;
            l.flags |= sched.run_flags() | GW_RUN_BUCKETS;
            return l;
        };
        if (cfg_.verbose)
            std::fprintf(stderr, "[engine] %.1f ms walkers and slots ready\n",
                         std::chrono::duration<double, std::milli>(clk::now() - t0).count());
        for (uint64_t it = 0;; ++it) {
            ck(gw_bucket_counts(ctx_, W.data(), W.size()), ctx_, "gw_bucket_counts");
            uint64_t left = 0;
            for (uint64_t x : W) left += x;
            if (!left) break;
            auto ts = clk::now();
            for (uint32_t s = 0; s < nslots_; ++s) {
                const int b = slots_.slot_stg[s];
                if (b < 0) continue;
                int32_t sb = -1;
                uint32_t ready = 0;
                ck(gw_staged_block(ctx_, s, &sb, &ready), ctx_, "gw_staged_block");
                staged[b] = ready ? 2 : 1;
            }
            SchedInput in{W.data(), nb, nslots_, cfg_.zero_threshold, blocks_.data(), it, staged.data()};
            IterationOutcome out;
            out.iteration = it;
            gw_run_stats st{};
            if (sched.schedule(in)) {
                uint64_t bytes = 0;
                // staged copies the scheduler wants become resident (the run waits for them)
                for (uint32_t b : sched.buckets)
                    if (slots_.blk_stg[b] >= 0) {
                        ck(gw_commit_block(ctx_, (uint32_t)slots_.blk_stg[b]), ctx_, "gw_commit_block");
                        slots_.commit(b);
                    }
                for (uint32_t b : sched.buckets) {
                    int s = slots_.place(b, sched.buckets);
                    if (s == -2) throw GwError(GW_ERR_STATE, "scheduler wants more blocks than slots");
                    if (s < 0) continue;
                    const gw_block_desc& d = blocks_[b];
                    if (const PackedBlock* pk = packed(b))
                        ck(gw_upload_block_packed(ctx_, (uint32_t)s, b, pk->data, pk->bytes), ctx_,
                           "gw_upload_block_packed");
                    else
                        ck(gw_upload_block(ctx_, (uint32_t)s, b, beg_ + d.start_vert, csr_ + d.start_edge), ctx_,
                           "gw_upload_block");
                    bytes += ((uint64_t)d.nverts + 1) * 8 + d.nedges * 4;
                    out.loaded.push_back(b);
                    r.uploads++;
                }
                // prefetch: staged blocks the scheduler still wants stay, new ones go into
                // slots that hold neither a wanted block nor one of those
                pf.clear();
                sched.prefetch(in, nslots_ > sched.buckets.size() ? nslots_ - (uint32_t)sched.buckets.size() : 0u, pf);
                std::vector<uint32_t> keep = sched.buckets;
                for (uint32_t b : pf)
                    if (slots_.blk_stg[b] >= 0) keep.push_back(b);
                for (uint32_t b : pf) {
                    if (slots_.blk_stg[b] >= 0 || slots_.blk_slot[b] >= 0) continue;
                    const int s = slots_.stage(b, keep);
                    if (s < 0) continue;
                    keep.push_back(b);
                    const gw_block_desc& d = blocks_[b];
                    if (const PackedBlock* pk = packed(b))
                        ck(gw_stage_block_packed(ctx_, (uint32_t)s, b, pk->data, pk->bytes), ctx_,
                           "gw_stage_block_packed");
                    else
                        ck(gw_stage_block(ctx_, (uint32_t)s, b, beg_ + d.start_vert, csr_ + d.start_edge), ctx_,
                           "gw_stage_block");
                    r.staged++;
                }
                for (uint32_t b = 0; b < nb; ++b) staged[b] = slots_.blk_stg[b] >= 0 ? 1 : 0;
                r.sched_ms += std::chrono::duration<double, std::milli>(clk::now() - ts).count();
                const gw_launch launch = next_launch();
                ck(gw_run(ctx_, &launch, &st), ctx_, "gw_run");
                out.h2d_bytes = bytes;
            } else {
                r.sched_ms += std::chrono::duration<double, std::milli>(clk::now() - ts).count();
                if (!cfg_.zero_copy_tail)
                    throw GwError(GW_ERR_STATE, "walkers remain but no block is worth loading (zero-copy tail off)");
                for (uint32_t b = 0; b < nb; ++b) {
                    const gw_block_desc& d = blocks_[b];
                    ck(gw_map_host_block(ctx_, b, beg_ + d.start_vert, csr_ + d.start_edge), ctx_,
                       "gw_map_host_block");
                }
                const gw_launch launch = next_launch();
                ck(gw_run(ctx_, &launch, &st), ctx_, "gw_run (zero-copy)");
                ck(gw_unmap_host_blocks(ctx_), ctx_, "gw_unmap_host_blocks");
                r.zero_copy_runs++;
                r.zero_copy_steps += st.steps;
                r.zc_touches += st.zc_touches;
            }
            out.steps = st.steps;
            out.trials = st.trials;
            out.probes = st.probes;
            out.kernel_ms = st.kernel_ms;
            out.walkers_run = st.walkers_run;
            out.zc_touches = st.zc_touches;
            out.zero_copy = st.zc_touches > 0;
            auto tf = clk::now();
            sched.feedback(in, out);
            r.sched_ms += std::chrono::duration<double, std::milli>(clk::now() - tf).count();
            r.iterations++;
            r.steps += st.steps;
            r.trials += st.trials;
            r.probes += st.probes;
            r.mix_steps += st.mix_steps;
            r.kernel_ms += st.kernel_ms;
            r.filter_ms += st.filter_ms;
            if (cfg_.verbose)
                std::fprintf(stderr, "[engine] %.1f ms it %llu: %zu resident, %zu loaded, %zu staged, steps %llu, "
                             "walkers run %llu, left %llu, kernel %.2f ms\n",
                             std::chrono::duration<double, std::milli>(clk::now() - t0).count(), (unsigned long long)it,
                             sched.buckets.size(), out.loaded.size(), pf.size(), (unsigned long long)st.steps,
                             (unsigned long long)st.walkers_run, (unsigned long long)st.walkers_left, st.kernel_ms);
        }
        sched.finish();
        ck(gw_wait_uploads(ctx_), ctx_, "gw_wait_uploads");
        gw_counters c1;
        ck(gw_counters_get(ctx_, &c1), ctx_, "gw_counters_get");
        r.h2d_bytes = c1.h2d_bytes - c0.h2d_bytes;
        r.h2d_raw_bytes = c1.h2d_raw_bytes - c0.h2d_raw_bytes;
        {
            uint64_t a = 0, b = 0, d = 0;
            ck(gw_upload_stats(ctx_, &a, &b, &d), ctx_, "gw_upload_stats");
            r.inplace_blocks = a - up0_[0];
            r.ring_blocks = b - up0_[1];
            r.inline_blocks = d - up0_[2];
        }
        {
            uint64_t rounds1 = 0;
            ck(gw_worklist_stats(ctx_, &wcap, &rounds1), ctx_, "gw_worklist_stats");
            r.kernel_launches = r.iterations + (rounds1 - rounds0);
        }
        r.h2d_ms = c1.h2d_ms - c0.h2d_ms;
        r.wall_s = std::chrono::duration<double>(clk::now() - t0).count();
        return r;
    }

private:
    const PackedBlock* packed(uint32_t b) const
    {
        if (!cfg_.packed || b >= cfg_.packed->size()) return nullptr;
        const PackedBlock& p = (*cfg_.packed)[b];
        return p.data ? &p : nullptr;
    }

    EngineConfig cfg_;
    bool pack_staging_ = false;
    uint32_t lanes_ = 1;
    const uint64_t* beg_;
    const uint32_t* csr_;
    std::vector<gw_block_desc> blocks_;
    gw_ctx* ctx_ = nullptr;
    uint32_t nslots_ = 0;
    uint64_t up0_[3] = {0, 0, 0};   // gw_upload_stats at the walk's start
    SlotMap slots_;
};

}  // namespace gwh
