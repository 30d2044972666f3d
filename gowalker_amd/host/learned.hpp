// learned.hpp -- learned block scheduler: a 12-64-64-1 MLP (BF16 weights, FP32
// accumulation, trained on the GPU by libgw.so's kernels) predicts the walk steps that
// loading a block buys over its stay in the cache (log1p scale), from the block's 12
// features at decision time (gw_block_features).  It is trained online on what the
// engine records: a load opens a sample (the features, the steps of the run that
// followed it) and every later run adds the steps the block served as the walkers'
// current block (the block-access trace, gw_trace_read) until its eviction closes it.
//
// Decisions, given the predictions y(b):
//   fill   (free slots left): commit staged blocks whose copies are done, else the best
//          staged one (the run waits for it), else upload the best candidate; then stage
//          the next-best blocks into the free slots so their copies overlap the run;
//   swap   (cache full): the plain rule's load and victim;
//   tail   when y(c) of the block the plain rule would load says the upload buys fewer
//          steps than its bytes buy in zero-copy reads (steps x reads per step x 64 B <
//          block bytes, reads per step measured on the walk so far) and all the work left
//          is worth at most tail_guard uploads in zero-copy bytes, finish the walk by
//          zero-copy instead; the reference's fixed walker-count threshold
//          (apps/Node2vec.cu:100-104, the plain rule's version of this decision) is the
//          latest point it switches at.
// The plain walker-count rule (engine/schedule.hpp:70-138) is its fallback: the model is
// ignored while it has too few samples and, for a cool-down period, whenever its recent
// predictions of the steps a load buys are off by more than a factor e on average (the
// error monitor).  Walks stay bit-identical either way (counter-based
// RNG).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "gw.h"
#include "schedule.hpp"

namespace gwh {

struct LearnedConfig {
    uint32_t h1 = 64, h2 = 64;
    uint64_t seed = 7;
    uint32_t min_samples = 24;      // plain swap decisions until this many upload samples
    uint32_t train_steps = 4;       // Adam steps per training round
    uint32_t train_every = 8;       // new samples between training rounds
    uint32_t batch = 256;
    uint32_t max_samples = 16384;   // ring buffer of trace samples
    float lr = 2e-3f;
    double zc_read_bytes = 64;      // PCIe bytes of one zero-copy read (row, candidate or probe)
    double tail_margin = 1.0;       // zero-copy once steps x zc bytes x margin < block bytes
    double tail_guard = 64;         // ... and the work left x zc bytes < guard x block bytes
    uint32_t max_inflight = 8;      // staged copies in flight at once: more blocks land per
                                    // run than with 4 (C3 16 GB: 0.471 s vs 0.478 s); deeper
                                    // queues make a run wait for a wanted block queued behind
                                    // others (0.49 s at 16 and 53)
    bool swap_prefetch = true;      // full cache: keep one slot for staging the next load
                                    // (overlaps every upload, costs a slot and missed copies;
                                    // C5 29.4 -> 22.7 s, C3 8 GB 4.31 -> 3.82 s, 11-13% more
                                    // PCIe bytes)
    bool myopic_label = false;      // label a load with the next run's steps only
    double error_limit = 1.0;       // mean |log error| of recent predictions above this => fall back
    uint32_t cooldown = 8;
};

class LearnedScheduler : public Scheduler {
public:
    LearnedScheduler(gw_ctx* ctx, const LearnedConfig& cfg) : ctx_(ctx), cfg_(cfg)
    {
        if (gw_mlp_create(ctx_, kF, cfg_.h1, cfg_.h2, cfg_.seed, &mlp_) != GW_OK) mlp_ = nullptr;
    }
    ~LearnedScheduler() override
    {
        if (mlp_) gw_mlp_destroy(ctx_, mlp_);
    }
    const char* name() const override { return "learned"; }
    uint32_t run_flags() const override { return GW_RUN_TRACE; }

    bool schedule(const SchedInput& in) override
    {
        const uint32_t nb = in.nblocks;
        if (age_.size() != nb) age_.assign(nb, 64.f);
        prepare(in);
        last_kind_ = kPlain;
        std::vector<char> cached(nb, 0);
        for (uint32_t b : buckets) cached[b] = 1;
        const uint32_t cap = capacity(in);

        if (buckets.size() < cap) {   // ---- fill: free slots remain
            bool took = false;
            for (uint32_t b = 0; b < nb && buckets.size() < cap; ++b)
                if (!cached[b] && in.staged && in.staged[b] == 2) { buckets.push_back(b); cached[b] = 1; took = true; }
            if (!took && in.staged) {
                int best = -1;
                for (uint32_t b = 0; b < nb; ++b)
                    if (!cached[b] && in.staged[b] == 1 && (best < 0 || rank_[b] > rank_[best])) best = (int)b;
                if (best >= 0) { buckets.push_back((uint32_t)best); cached[best] = 1; took = true; }
            }
            if (took) { last_kind_ = kFill; return true; }
            const int c = best_candidate(cached);
            if (c < 0) return plain_decide(in);
            if (tail_now(in, (uint32_t)c)) return false;
            buckets.push_back((uint32_t)c);
            last_kind_ = kFill;
            return true;
        }

        // ---- swap: the cache is full.  The plain rule picks the load and the victim (its
        // pairing-aware eviction moved fewer bytes than evicting by predicted value in our
        // measurements); the model decides when uploads stop paying for themselves.
        if (cooldown_ > 0) --cooldown_;
        const std::vector<uint32_t> before = buckets;
        if (!plain_decide(in)) return false;
        int c = -1;
        for (uint32_t b : buckets)
            if (std::find(before.begin(), before.end(), b) == before.end()) c = (int)b;
        if (c >= 0 && tail_now(in, (uint32_t)c)) {
            buckets = before;
            return false;
        }
        return true;
    }

    void prefetch(const SchedInput& in, uint32_t spare, std::vector<uint32_t>& out) override
    {
        const uint32_t nb = in.nblocks;
        const uint32_t cap = capacity(in);
        if (!spare) return;
        std::vector<char> skip(nb, 0);
        for (uint32_t b : buckets) skip[b] = 1;
        std::vector<uint32_t> cand;
        for (uint32_t b = 0; b < nb; ++b)
            if (!skip[b] && demand_[b] > 0) cand.push_back(b);
        std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return rank_[a] > rank_[b]; });
        if (buckets.size() < cap) {
            // fill: keep the copies in flight, top up to max_inflight with the best blocks
            uint32_t n = 0;
            for (uint32_t b : cand)
                if (in.staged && in.staged[b] == 1) { out.push_back(b); ++n; }
            for (uint32_t b : cand) {
                if (n >= cfg_.max_inflight || out.size() >= std::min<uint32_t>(spare, cap - (uint32_t)buckets.size())) break;
                if (in.staged && in.staged[b]) continue;
                out.push_back(b);
                ++n;
            }
            return;
        }
        // swap: stage into the spare slot the block the plain rule would load next on the
        // current bucket matrix, so that when it does, the copy has overlapped this run
        if (!cfg_.swap_prefetch) return;
        PlainScheduler next;
        next.buckets = buckets;
        SchedInput pin = in;
        pin.ncblock = cap;
        uint64_t mx = 0;
        for (uint32_t b = 0; b < nb; ++b) {
            const gw_block_desc& d = in.blocks[b];
            mx = std::max<uint64_t>(mx, ((uint64_t)d.nverts + 1) * 8 + d.nedges * 4);
        }
        pin.zero_threshold = std::max<uint64_t>(in.zero_threshold, (uint64_t)zero_floor((double)mx));
        if (!next.schedule(pin)) return;
        for (uint32_t b : next.buckets)
            if (!skip[b]) { out.push_back(b); return; }
    }

    void feedback(const SchedInput& in, const IterationOutcome& out) override
    {
        const uint32_t nb = in.nblocks;
        std::vector<uint64_t> trace(nb, 0);
        if (gw_trace_read(ctx_, trace.data(), nb) != GW_OK) return;
        if (open_.size() != nb) open_.assign(nb, Open{});
        std::vector<char> now(nb, 0);
        for (uint32_t b : buckets) now[b] = 1;
        // a block that left the resident set closes its sample: the steps it served from
        // its upload to its eviction
        for (uint32_t b = 0; b < nb; ++b)
            if (open_[b].live && !now[b]) close_sample(b);
        // blocks that became resident open one: (features when the load was decided, the
        // run's steps if it was the only new block, else the steps it served as the
        // walkers' current block)
        std::vector<uint32_t> fresh;
        for (uint32_t b : buckets)
            if (!was_resident_[b]) fresh.push_back(b);
        for (uint32_t b : fresh) {
            if (!have_feats_ || feats_.size() < (size_t)(b + 1) * kF) continue;
            Open& o = open_[b];
            o.live = true;
            o.f.assign(feats_.begin() + (size_t)b * kF, feats_.begin() + (size_t)(b + 1) * kF);
            o.steps = fresh.size() == 1 ? (double)out.steps : (double)trace[b];
            o.pred = use_model_ ? pred_[b] : -1.f;
        }
        // reads per step (row + candidates + probes) of the walks so far: a zero-copy step
        // costs about that many PCIe reads
        if (out.steps) {
            const double r = 1.0 + (double)(out.trials + out.probes) / (double)out.steps;
            reads_per_step_ = reads_n_ ? 0.9 * reads_per_step_ + 0.1 * r : r;
            ++reads_n_;
        }
        for (uint32_t b = 0; b < nb; ++b) {
            if (now[b] && open_[b].live && !cfg_.myopic_label &&
                std::find(fresh.begin(), fresh.end(), b) == fresh.end())
                open_[b].steps += (double)trace[b];
            age_[b] = trace[b] ? 0.f : std::min(age_[b] + 1.f, 1024.f);
        }
        if (mlp_ && samples_ >= 8 && samples_ >= trained_at_ + cfg_.train_every) {
            trained_at_ = samples_;
            train();
        }
    }

    uint64_t learned_decisions() const { return learned_decisions_; }
    uint64_t fallbacks() const { return fallbacks_; }
    uint64_t tail_decisions() const { return tail_decisions_; }
    float last_loss() const { return last_loss_; }

private:
    static constexpr uint32_t kF = 12;
    enum Kind { kPlain, kFill, kLearnedSwap };

    bool trained() const { return mlp_ && samples_ >= cfg_.min_samples; }

    // resident blocks allowed: every slot, or all but the staging slot when the graph does
    // not fit and the next load is prefetched
    uint32_t capacity(const SchedInput& in) const
    {
        return cfg_.swap_prefetch && in.nblocks > in.ncblock && in.ncblock > 1 ? in.ncblock - 1 : in.ncblock;
    }

    double zc_bytes() const { return cfg_.zc_read_bytes * (reads_n_ ? reads_per_step_ : 6.0); }

    struct Open {
        bool live = false;
        std::vector<float> f;
        double steps = 0;
        float pred = -1.f;
    };

    void close_sample(uint32_t b)
    {
        Open& o = open_[b];
        o.live = false;
        const float y = std::log1p((float)o.steps);
        add_sample(o.f.data(), y);
        // error monitor: how far the model's prediction for this load was from what it bought
        if (o.pred >= 0.f) {
            const double e = std::fabs((double)o.pred - (double)y);
            err_ = err_n_ ? 0.8 * err_ + 0.2 * e : e;
            ++err_n_;
            if (err_n_ >= 4 && err_ > cfg_.error_limit) {
                cooldown_ = cfg_.cooldown;
                err_n_ = 0;
                ++fallbacks_;
            }
        }
    }

    // features, walker demand and the ranking score of every block
    void prepare(const SchedInput& in)
    {
        const uint32_t nb = in.nblocks;
        const uint64_t n = nb;
        std::vector<char> cached(nb, 0);
        for (uint32_t b : buckets) cached[b] = 1;
        was_resident_ = cached;
        uint64_t wl = 0;
        if (gw_pending_work(ctx_, &wl, &pending_hops_) != GW_OK) pending_hops_ = ~0ULL;   // host copy, no device trip
        add_.assign(nb, 0);
        demand_.assign(nb, 0);
        for (uint32_t b = 0; b < nb; ++b) {
            uint64_t a = in.W[b * n + b], d = in.W[b * n + b];
            for (uint32_t j = 0; j < nb; ++j) {
                if (j == b) continue;
                const uint64_t x = in.W[b * n + j] + in.W[(uint64_t)j * n + b];
                d += x;
                if (cached[j]) a += x;
            }
            add_[b] = a;
            demand_[b] = d;
        }
        // the model is consulted while free slots are being filled and once its tail rule
        // can fire (all the work left is worth at most tail_guard uploads in zero-copy
        // reads); elsewhere the plain rule decides and a forward pass would be wasted
        uint64_t max_bytes = 0;
        for (uint32_t b = 0; b < nb; ++b) {
            const gw_block_desc& d = in.blocks[b];
            max_bytes = std::max<uint64_t>(max_bytes, ((uint64_t)d.nverts + 1) * 8 + d.nedges * 4);
        }
        const bool need = buckets.size() < capacity(in) ||
                          (double)pending_hops_ * zc_bytes() < cfg_.tail_guard * (double)max_bytes;
        // features (a kernel and a readback): whenever the model is consulted, and on every
        // 4th iteration in between so that loads keep opening training samples
        have_feats_ = false;
        if (need || in.iteration % 4 == 0) {
            feats_.assign((size_t)nb * kF, 0.f);
            have_feats_ = gw_block_features(ctx_, age_.data(), feats_.data(), feats_.size()) == GW_OK;
        }
        pred_.assign(nb, 0.f);
        use_model_ = need && trained() && have_feats_ &&
                     gw_mlp_forward(ctx_, mlp_, feats_.data(), nb, pred_.data()) == GW_OK;
        if (use_model_) ++learned_decisions_;
        rank_.resize(nb);
        // before the model is trained: the walkers a load would release (log scale), with
        // the walkers that need the block at all as a tie-break
        for (uint32_t b = 0; b < nb; ++b)
            rank_[b] = use_model_ ? pred_[b] : std::log1p((float)add_[b]) + 1e-3f * std::log1p((float)demand_[b]);
    }

    // the non-resident block with the best rank among those whose walkers could run now
    int best_candidate(const std::vector<char>& cached) const
    {
        int best = -1;
        for (uint32_t c = 0; c < add_.size(); ++c) {
            if (cached[c] || add_[c] == 0) continue;
            if (best < 0 || rank_[c] > rank_[best]) best = (int)c;
        }
        return best;
    }

    // the reference's zero-copy threshold for a block of this size, in waiting walkers
    // (blocksize / 128 / 24, apps/Node2vec.cu:100-104): the learned tail never comes later
    static double zero_floor(double bytes) { return bytes / 3072.0; }

    bool tail_now(const SchedInput& in, uint32_t c)
    {
        const gw_block_desc& d = in.blocks[c];
        const double bytes = ((double)d.nverts + 1) * 8 + (double)d.nedges * 4;
        // the reference's rule looks at the best walker count of all candidates
        uint64_t best_add = 0;
        for (uint32_t b = 0; b < in.nblocks; ++b)
            if (!was_resident_[b]) best_add = std::max(best_add, add_[b]);
        if ((double)best_add <= zero_floor(bytes)) {
            ++tail_decisions_;
            return true;
        }
        if (!use_model_ || cooldown_ > 0) return false;
        const double steps = std::expm1((double)std::max(0.f, pred_[c]));
        const double zc = zc_bytes();
        if (steps * zc * cfg_.tail_margin < bytes && (double)pending_hops_ * zc < cfg_.tail_guard * bytes) {
            ++tail_decisions_;
            return true;
        }
        return false;
    }

    bool plain_decide(const SchedInput& in)
    {
        plain_.buckets = buckets;
        SchedInput pin = in;
        pin.ncblock = capacity(in);
        uint64_t mx = 0;
        for (uint32_t b = 0; b < in.nblocks; ++b) {
            const gw_block_desc& d = in.blocks[b];
            mx = std::max<uint64_t>(mx, ((uint64_t)d.nverts + 1) * 8 + d.nedges * 4);
        }
        pin.zero_threshold = std::max<uint64_t>(in.zero_threshold, (uint64_t)zero_floor((double)mx));
        bool r = plain_.schedule(pin);
        buckets = plain_.buckets;
        last_kind_ = kPlain;
        return r;
    }

    void add_sample(const float* f, float y)
    {
        if (xs_.size() < (size_t)cfg_.max_samples * kF) {
            xs_.insert(xs_.end(), f, f + kF);
            ys_.push_back(y);
        } else {
            const size_t i = ring_++ % cfg_.max_samples;
            std::copy(f, f + kF, xs_.begin() + i * kF);
            ys_[i] = y;
        }
        ++samples_;
    }

    void train()
    {
        const uint32_t n = (uint32_t)ys_.size();
        if (!n) return;
        const uint32_t bs = std::min(cfg_.batch, n);
        std::vector<float> bx((size_t)bs * kF), by(bs);
        for (uint32_t e = 0; e < cfg_.train_steps; ++e) {
            for (uint32_t i = 0; i < bs; ++i) {
                rng_ = rng_ * 6364136223846793005ULL + 1442695040888963407ULL;
                const uint32_t k = (uint32_t)((rng_ >> 33) % n);
                std::copy(xs_.begin() + (size_t)k * kF, xs_.begin() + (size_t)(k + 1) * kF, bx.begin() + (size_t)i * kF);
                by[i] = ys_[k];
            }
            float loss = 0.f;
            if (gw_mlp_step(ctx_, mlp_, bx.data(), by.data(), bs, cfg_.lr, &loss) != GW_OK) return;
            last_loss_ = loss;
        }
    }

    gw_ctx* ctx_;
    LearnedConfig cfg_;
    gw_mlp* mlp_ = nullptr;
    PlainScheduler plain_;
    std::vector<float> age_, feats_, pred_, rank_, xs_, ys_;
    std::vector<uint64_t> add_, demand_;
    std::vector<char> was_resident_;
    std::vector<Open> open_;
    double reads_per_step_ = 0;
    uint64_t reads_n_ = 0;
    uint64_t trained_at_ = 0;
    uint64_t pending_hops_ = 0;
    bool have_feats_ = false, use_model_ = false;
    uint64_t samples_ = 0, ring_ = 0, rng_ = 0x853c49e6748fea9bULL;
    uint32_t cooldown_ = 0;
    Kind last_kind_ = kPlain;
    uint64_t learned_decisions_ = 0, fallbacks_ = 0, tail_decisions_ = 0, err_n_ = 0;
    double err_ = 0;
    float last_loss_ = 0.f;
};

}  // namespace gwh
