// learned.hpp -- learned block scheduler: a 12-64-64-1 MLP (BF16 weights, FP32
// accumulation, trained on the GPU by libgw.so's kernels) predicts V(b), the walk steps
// block b serves from a decision until its eviction (log1p scale), from the block's 12
// features at that decision (gw_block_features).  It is trained online on what the
// engine records (the block-access trace, gw_trace_read):
//   * a load opens a sample: the features when the load was decided, the steps of the run
//     that followed, then every later run's steps with the block as the walkers' current
//     block, until its eviction closes it;
//   * every keep_every-th iteration each resident block without an open keep sample opens
//     one the same way (features with residency 1), so V is also learned for blocks that
//     are already resident -- the value an eviction gives up.
// The samples and the parameters outlive a walk (ModelState): walks of the same workload
// train one network, as the plain scheduler's runs would record traces for it.
//
// Decisions, given V:
//   fill   (free slots left): commit staged blocks whose copies are done, else the best
//          staged one (the run waits for it), else upload the best-ranked candidate; then
//          stage the next-best blocks into the free slots so their copies overlap the run;
//          rank = V(b) (the fill picks the blocks the swap phase starts from);
//   swap   (cache full): the plain rule's scores (walkers a load makes runnable, minus the
//          pairing its victim takes away) form a shortlist -- candidates scoring at least
//          shortlist_frac of the best, at most `shortlist` of them; the network picks the
//          load c with the largest V(c) - V(m(c)), m(c) the plain rule's victim for c (its
//          least-paired resident block; among victims paired as little, the one with the
//          smallest V); with value_victim, m(c) minimises the steps its eviction gives up,
//          s x pairing(c, m) + V(m), s the measured steps per walker and run;
//   tail   when V(c) of the block that would be loaded says the upload takes longer than
//          zero-copy reads would take for the steps it buys (V x reads per step / the
//          zero-copy read rate > block bytes / the copy rate; reads per step from the walk
//          so far, both rates measured: the zero-copy rate on earlier walks' tails) and
//          all the work left would take at most tail_guard uploads' time in zero-copy
//          reads, finish the walk by zero-copy instead; the reference's fixed walker-count
//          threshold (apps/Node2vec.cu:100-104, the plain rule's version of this decision)
//          is the latest point it switches at.
// The plain walker-count rule (engine/schedule.hpp:70-138) is its fallback: the network is
// ignored while it has too few samples and, for a cool-down period, whenever its recent
// predictions are off by more than a factor e on average (the error monitor).  When every
// block fits the cache, each is loaded once and never evicted and the fill only wants steps
// while copies are in flight, which the walker count measures: the network, its samples
// and its training are left out of that walk (the staging ablation's decisions).
//
// With use_model false the same code runs with the network left out (the staging
// ablation): fill rank = the walkers a load would release (ties: the walkers that need the
// block at all), swap = the plain rule, tail
// = the reference's threshold only.  Walks stay bit-identical either way (counter-based
// RNG): only bytes moved and time change.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "gw.h"
#include "schedule.hpp"

namespace gwh {

// What carries over from one walk to the next: the network and its samples.
struct ModelState {
    std::vector<float> params;          // fp32 image of the bf16 weights (empty: untrained)
    std::vector<float> xs, ys;          // sample ring: features (kF each) and log1p labels
    uint64_t samples = 0, ring = 0, trained_at = 0, adam_steps = 0, walks = 0;
    uint64_t rng = 0x853c49e6748fea9bULL;
    float loss = 0.f;
    double zc_reads_per_ms = 0;         // zero-copy reads per ms measured on tail runs (0: none yet)
    // prediction quality on samples the network has not been trained on: when a sample
    // closes (its label is known, it is not yet in the ring), the current network predicts
    // it from its features; its mean |log error| against that of the running mean of all
    // labels seen before (a predictor with no features).  Scored once the network has
    // trained on score_after samples.
    uint64_t scored = 0;
    double net_abs_err = 0, mean_abs_err = 0, label_sum = 0;
    uint64_t labels = 0;
    // the network is trained on standardized labels (y - y_mu) / y_sigma, the mean and
    // deviation of the sample ring at the last training round (y_sigma 0: none yet), and
    // predicts y_mu + y_sigma x its output: an untrained network predicts the mean label,
    // and Adam need not first walk the output bias to the label scale (log1p of millions
    // of steps is ~15; at a 2e-3 rate that alone takes thousands of steps); recomputing
    // both per round follows the labels as the ring turns over (the first samples of a
    // workload can be unlike the later ones)
    float y_mu = 0.f, y_sigma = 0.f;
    // inputs are standardized the same way: per feature, the mean and deviation of the ring
    // at the last training round (the features mix log counts of up to ~17 with fractions)
    std::vector<float> x_mu, x_sigma;
};

struct LearnedConfig {
    bool use_model = true;          // false: the staging ablation (no network)
    uint32_t h1 = 64, h2 = 64;
    uint64_t seed = 7;
    uint32_t min_samples = 24;      // plain decisions until this many samples
    uint32_t train_steps = 2;       // Adam steps per training round (queued, one host wait)
    uint32_t train_every = 32;      // new samples between training rounds (every 8 with 4 steps
                                    // cost 11 ms of host time per 0.19 s RMAT-24 walk)
    uint32_t finish_steps = 16;     // Adam steps when a walk ends (~50 us each).  4 and 48 steps
                                    // gave held-out errors 0.41 vs 1.12 for the running mean on
                                    // C3 (these: 0.74 vs 1.36) but cost the 16 GB walk 8 ms
    uint32_t score_after = 256;     // samples trained on before predictions are scored
    uint32_t batch = 256;
    uint32_t max_samples = 16384;   // ring buffer of trace samples
    uint32_t keep_every = 4;        // iterations between keep samples of resident blocks
    float lr = 2e-3f;
    double zc_prior_reads_per_ms = 375e3;   // zero-copy read rate before one is measured
                                            // (64-byte reads at ~24 GB/s)
    double copy_prior_bytes_per_ms = 55e6;  // block copy rate before one is measured
    double tail_margin = 1.0;       // zero-copy once its time for V(c) steps x margin < upload time
    double tail_guard = 8;          // ... and the whole tail's zero-copy time < guard uploads' time
    uint32_t max_inflight = 8;      // staged copies in flight at once: more blocks land per
                                    // run than with 4 (C3 16 GB: 0.471 s vs 0.478 s); deeper
                                    // queues make a run wait for a wanted block queued behind
                                    // others (0.49 s at 16 and 53)
    bool swap_prefetch = true;      // full cache: keep one slot for staging the next load
                                    // (overlaps every upload, costs a slot and missed copies;
                                    // C5 29.4 -> 22.7 s, C3 8 GB 4.31 -> 3.82 s, 11-13% more
                                    // PCIe bytes)
    bool model_swap = false;        // the network chooses swap loads among near-best candidates
                                    // (else the plain rule; the plain rule moved 1-5% fewer
                                    // bytes in every measurement)
    bool value_victim = false;      // ... and victims by the steps an eviction gives up
    bool sticky = true;             // a swap takes the block already staged when its plain
                                    // score is within the shortlist (its copy has overlapped
                                    // the last run; any other load waits for a fresh copy)
    uint32_t shortlist = 4;         // swap candidates the network chooses among
    double shortlist_frac = 0.8;    // ... each scoring at least this share of the best
    bool myopic_label = false;      // label a load with the next run's steps only
    double error_limit = 1.0;       // mean |log error| of recent predictions above this => fall back
    uint32_t cooldown = 8;
};

class LearnedScheduler : public Scheduler {
public:
    LearnedScheduler(gw_ctx* ctx, const LearnedConfig& cfg, ModelState* state = nullptr)
        : ctx_(ctx), cfg_(cfg), state_(state)
    {
        if (!cfg_.use_model) return;
        if (gw_mlp_create(ctx_, kF, cfg_.h1, cfg_.h2, cfg_.seed, &mlp_) != GW_OK) mlp_ = nullptr;
        if (state_) m_ = *state_;
        uint32_t np = 0;
        if (!mlp_ || gw_mlp_num_params(mlp_, &np) != GW_OK) return;
        if (!m_.params.empty() && np == m_.params.size()) {
            (void)gw_mlp_set_params(ctx_, mlp_, m_.params.data());
        } else {
            // a new network: its output layer (w3, b3: the last h2 + 1 parameters) starts at
            // zero, so before training it predicts the mean label exactly instead of the mean
            // plus the noise of a random output layer; the gradients of w3 start the rest
            std::vector<float> p(np);
            if (gw_mlp_get_params(ctx_, mlp_, p.data()) == GW_OK && np > cfg_.h2) {
                std::fill(p.end() - (cfg_.h2 + 1), p.end(), 0.f);
                (void)gw_mlp_set_params(ctx_, mlp_, p.data());
            }
        }
    }
    ~LearnedScheduler() override
    {
        if (mlp_) gw_mlp_destroy(ctx_, mlp_);
    }
    const char* name() const override { return cfg_.use_model ? "learned" : "staging"; }
    // the block-access trace feeds the samples; none are taken when every block fits
    uint32_t run_flags() const override { return cfg_.use_model && !fits_ ? (uint32_t)GW_RUN_TRACE : 0u; }

    // Stores the network and its samples for the next walk (call once the walk is done).
    void save()
    {
        if (!state_ || !mlp_) return;
        uint32_t np = 0;
        if (gw_mlp_num_params(mlp_, &np) == GW_OK) {
            m_.params.resize(np);
            if (gw_mlp_get_params(ctx_, mlp_, m_.params.data()) != GW_OK) m_.params.clear();
        }
        m_.walks++;
        *state_ = m_;
    }

    // The walk is over: samples still open are closed with the steps their blocks served
    // until the end (at a cap where nothing is evicted, every sample ends this way).
    void finish() override
    {
        for (Open& o : open_)
            if (o.live) close_sample(o);
        for (Open& o : keep_)
            if (o.live) close_sample(o);
        flush_closed();
        // the walk is over: a longer training round on every sample so far, off the
        // per-iteration path (the next walk decides with it)
        if (mlp_ && !fits_ && m_.samples >= cfg_.min_samples) {
            m_.trained_at = m_.samples;
            train(cfg_.finish_steps);
        }
    }

    bool schedule(const SchedInput& in) override
    {
        const uint32_t nb = in.nblocks;
        if (age_.size() != nb) age_.assign(nb, 64.f);
        prepare(in);
        std::vector<char> cached(nb, 0);
        for (uint32_t b : buckets) cached[b] = 1;
        const uint32_t cap = capacity(in);

        if (buckets.size() < cap) {   // ---- fill: free slots remain
            // the network decided this fill step when its rank (V) picks another block than
            // the walker rank would
            if (use_model_) {
                const int a = best_candidate(cached, rank_), b = best_candidate(cached, rank0_);
                if (a != b) ++fill_model_;
            }
            bool took = false;
            for (uint32_t b = 0; b < nb && buckets.size() < cap; ++b)
                if (!cached[b] && in.staged && in.staged[b] == 2) { buckets.push_back(b); cached[b] = 1; took = true; }
            if (!took && in.staged) {
                int best = -1;
                for (uint32_t b = 0; b < nb; ++b)
                    if (!cached[b] && in.staged[b] == 1 && (best < 0 || rank_[b] > rank_[best])) best = (int)b;
                if (best >= 0) { buckets.push_back((uint32_t)best); cached[best] = 1; took = true; }
            }
            if (took) return true;
            const int c = best_candidate(cached, rank_);
            if (c < 0) {
                if (plain_decide(in)) return true;
                ++tail_threshold_;
                return false;
            }
            if (tail_now(in, (uint32_t)c)) return false;
            buckets.push_back((uint32_t)c);
            return true;
        }

        // ---- swap: the cache is full
        if (cooldown_ > 0) --cooldown_;
        const std::vector<uint32_t> before = buckets;
        if (!plain_decide(in)) {                // the plain rule finds nothing worth loading
            ++tail_threshold_;
            return false;
        }
        const std::vector<uint32_t> plain_choice = buckets;
        {
            uint32_t c, pos;
            int kind = 0;
            if (swap_choice(in, before, in.staged, &c, &pos, &kind)) {
                buckets = before;
                buckets[pos] = c;
                if (kind == kSticky) ++sticky_;
                if (kind == kModel) {
                    ++swap_model_;
                    if (buckets != plain_choice) ++swap_differs_;
                }
            }
        }
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
        // swap: stage into the spare slot the block this scheduler would load next on the
        // current bucket matrix, so that when it does, the copy has overlapped this run
        if (!cfg_.swap_prefetch) return;
        PlainScheduler next;
        next.buckets = buckets;
        SchedInput pin = in;
        pin.ncblock = cap;
        pin.zero_threshold = std::max<uint64_t>(in.zero_threshold, (uint64_t)zero_floor((double)max_block_bytes(in)));
        if (!next.schedule(pin)) return;
        {
            uint32_t c, pos;
            int kind = 0;
            if (swap_choice(in, buckets, nullptr, &c, &pos, &kind) && !skip[c]) { out.push_back(c); return; }
        }
        for (uint32_t b : next.buckets)
            if (!skip[b]) { out.push_back(b); return; }
    }

    void feedback(const SchedInput& in, const IterationOutcome& out) override
    {
        if (!cfg_.use_model || fits_) return;
        const uint32_t nb = in.nblocks;
        std::vector<uint64_t> trace(nb, 0);
        if (gw_trace_read(ctx_, trace.data(), nb) != GW_OK) return;
        if (open_.size() != nb) { open_.assign(nb, Open{}); keep_.assign(nb, Open{}); }
        std::vector<char> now(nb, 0);
        for (uint32_t b : buckets) now[b] = 1;
        // a block that left the resident set closes its samples: the steps it served from
        // the decision to its eviction
        for (uint32_t b = 0; b < nb; ++b)
            if (!now[b]) {
                if (open_[b].live) close_sample(open_[b]);
                if (keep_[b].live) close_sample(keep_[b]);
            }
        // blocks that became resident open a load sample: (features when the load was
        // decided, the run's steps if it was the only new block, else the steps it served
        // as the walkers' current block)
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
        // resident blocks that stayed open a keep sample now and then (features with
        // residency 1; label = the steps they serve from here to their eviction)
        if (have_feats_ && cfg_.keep_every && in.iteration % cfg_.keep_every == 0)
            for (uint32_t b : buckets)
                if (was_resident_[b] && !keep_[b].live && feats_.size() >= (size_t)(b + 1) * kF) {
                    Open& o = keep_[b];
                    o.live = true;
                    o.f.assign(feats_.begin() + (size_t)b * kF, feats_.begin() + (size_t)(b + 1) * kF);
                    o.steps = (double)trace[b];
                    o.pred = use_model_ ? pred_[b] : -1.f;
                }
        // reads per step (row + candidates + probes) of the walks so far: a zero-copy step
        // costs about that many PCIe reads; steps per walker and run
        if (out.steps && !out.zero_copy) {
            const double r = 1.0 + (double)(out.trials + out.probes) / (double)out.steps;
            reads_per_step_ = reads_n_ ? 0.9 * reads_per_step_ + 0.1 * r : r;
            ++reads_n_;
            if (out.walkers_run) {
                const double sw = (double)out.steps / (double)out.walkers_run;
                steps_per_walker_ = sw_n_ ? 0.9 * steps_per_walker_ + 0.1 * sw : sw;
                ++sw_n_;
            }
        }
        // the zero-copy read rate of a tail run (kept for later walks)
        if (out.zero_copy && out.kernel_ms > 0 && out.zc_touches >= 1000000) {
            const double rate = (double)out.zc_touches / out.kernel_ms;
            m_.zc_reads_per_ms = m_.zc_reads_per_ms > 0 ? 0.5 * m_.zc_reads_per_ms + 0.5 * rate : rate;
        }
        for (uint32_t b = 0; b < nb; ++b) {
            if (now[b] && !cfg_.myopic_label && std::find(fresh.begin(), fresh.end(), b) == fresh.end()) {
                if (open_[b].live) open_[b].steps += (double)trace[b];
                if (keep_[b].live) keep_[b].steps += (double)trace[b];
            }
            age_[b] = trace[b] ? 0.f : std::min(age_[b] + 1.f, 1024.f);
        }
        const bool first = m_.y_sigma <= 0.f && m_.samples >= cfg_.min_samples;
        if (mlp_ && (first || (m_.y_sigma > 0.f && m_.samples >= m_.trained_at + cfg_.train_every))) {
            m_.trained_at = m_.samples;
            train();
        }
    }

    // iterations in which the network's predictions ranked the fill, chose a swap or the
    // tail
    uint64_t learned_decisions() const { return fill_model_ + swap_model_ + tail_model_; }
    uint64_t fallbacks() const { return fallbacks_; }
    uint64_t tail_decisions() const { return tail_model_ + tail_threshold_; }
    uint64_t tail_model() const { return tail_model_; }
    uint64_t tail_threshold() const { return tail_threshold_; }
    uint64_t swap_model() const { return swap_model_; }
    uint64_t swap_differs() const { return swap_differs_; }
    uint64_t fill_model() const { return fill_model_; }
    uint64_t sticky_swaps() const { return sticky_; }
    uint64_t samples() const { return m_.samples; }
    float last_loss() const { return m_.loss; }

private:
    static constexpr uint32_t kF = 12;

    bool trained() const { return mlp_ && m_.y_sigma > 0.f; }   // trained at least once
    bool model_decides() const { return use_model_ && cooldown_ == 0 && cfg_.model_swap; }

    // resident blocks allowed: every slot, or all but the staging slot when the graph does
    // not fit and the next load is prefetched
    uint32_t capacity(const SchedInput& in) const
    {
        return cfg_.swap_prefetch && in.nblocks > in.ncblock && in.ncblock > 1 ? in.ncblock - 1 : in.ncblock;
    }

    static uint64_t block_bytes(const gw_block_desc& d) { return ((uint64_t)d.nverts + 1) * 8 + d.nedges * 4; }
    static uint64_t max_block_bytes(const SchedInput& in)
    {
        uint64_t mx = 0;
        for (uint32_t b = 0; b < in.nblocks; ++b) mx = std::max<uint64_t>(mx, block_bytes(in.blocks[b]));
        return mx;
    }

    double reads_per_step() const { return reads_n_ ? reads_per_step_ : 6.0; }
    double zc_reads_per_ms() const { return m_.zc_reads_per_ms > 0 ? m_.zc_reads_per_ms : cfg_.zc_prior_reads_per_ms; }
    double copy_bytes_per_ms() const
    {
        gw_counters c;
        if (gw_counters_get(ctx_, &c) == GW_OK && c.h2d_ms > 10 && c.h2d_bytes)
            return (double)c.h2d_bytes / c.h2d_ms;
        return cfg_.copy_prior_bytes_per_ms;
    }

    struct Open {
        bool live = false;
        std::vector<float> f;
        double steps = 0;
        float pred = -1.f;
    };

    // The network's predictions (log1p steps) for n feature rows: inputs and outputs in the
    // units the network was trained on (standardized), the result back in label units.
    bool predict(const std::vector<float>& f, uint32_t n, std::vector<float>& out)
    {
        std::vector<float> x(f.begin(), f.begin() + (size_t)n * kF);
        standardize(x.data(), n);
        out.resize(n);
        if (gw_mlp_forward(ctx_, mlp_, x.data(), n, out.data()) != GW_OK) return false;
        for (float& p : out) p = std::max(0.f, m_.y_mu + m_.y_sigma * p);
        return true;
    }

    void standardize(float* x, uint32_t n) const
    {
        if (m_.x_mu.size() != kF) return;
        for (uint32_t i = 0; i < n; ++i)
            for (uint32_t j = 0; j < kF; ++j) x[(size_t)i * kF + j] = (x[(size_t)i * kF + j] - m_.x_mu[j]) / m_.x_sigma[j];
    }

    // Closed samples wait here (a few dozen at a time) to be scored by the current network
    // -- which has never trained on them -- and then join the ring.
    void flush_closed()
    {
        if (closed_y_.empty()) return;
        const uint32_t n = (uint32_t)closed_y_.size();
        std::vector<float> p;
        if (trained() && m_.trained_at >= cfg_.score_after && m_.labels && predict(closed_f_, n, p)) {
            const double mean = m_.label_sum / (double)m_.labels;
            for (uint32_t i = 0; i < n; ++i) {
                ++m_.scored;
                m_.net_abs_err += std::fabs((double)p[i] - (double)closed_y_[i]);
                m_.mean_abs_err += std::fabs(mean - (double)closed_y_[i]);
            }
        }
        for (uint32_t i = 0; i < n; ++i) {
            m_.label_sum += closed_y_[i];
            ++m_.labels;
            add_sample(closed_f_.data() + (size_t)i * kF, closed_y_[i]);
        }
        closed_f_.clear();
        closed_y_.clear();
    }

    void close_sample(Open& o)
    {
        o.live = false;
        const float y = std::log1p((float)o.steps);
        closed_f_.insert(closed_f_.end(), o.f.begin(), o.f.end());
        closed_y_.push_back(y);
        if (closed_y_.size() >= cfg_.train_every) flush_closed();
        // error monitor: how far the network's prediction was from what the block served
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
        // features (a kernel and a readback) and a forward pass whenever the network may
        // decide, and on every keep_every-th iteration for samples.  When every block fits
        // the cache, each is loaded once and never evicted: no decision is left that V could
        // make better than the walker count (ranking that fill by V loaded blocks whose
        // walkers could not run yet: C3 16 GB 236-245 ms of step kernel per walk against
        // 234 ms), so the network, its samples and its host waits are left out
        have_feats_ = false;
        use_model_ = false;
        pred_.assign(nb, 0.f);
        fits_ = nb <= capacity(in);
        if (cfg_.use_model && !fits_) {
            // predictions are used by the fill, by network swaps, and by the tail rule once
            // the work left could be done in zero-copy reads within tail_guard uploads' time;
            // features are also taken every keep_every-th iteration for samples
            uint64_t max_bytes = 0;
            for (uint32_t b = 0; b < nb; ++b) max_bytes = std::max<uint64_t>(max_bytes, block_bytes(in.blocks[b]));
            const bool tail_near = (double)pending_hops_ * reads_per_step() / zc_reads_per_ms() <
                                   cfg_.tail_guard * (double)max_bytes / copy_bytes_per_ms();
            const bool fill = buckets.size() < capacity(in);
            const bool score = trained() && m_.x_mu.size() == kF && (fill || cfg_.model_swap || tail_near);
            const bool want = score || fill || in.iteration % std::max(1u, cfg_.keep_every) == 0;
            if (want) {
                feats_.assign((size_t)nb * kF, 0.f);
                if (score) {   // features and the network's scores with one host wait
                    have_feats_ = use_model_ = gw_block_scores(ctx_, mlp_, age_.data(), m_.x_mu.data(), m_.x_sigma.data(),
                                                               feats_.data(), pred_.data(), nb) == GW_OK;
                    if (use_model_)
                        for (float& p : pred_) p = std::max(0.f, m_.y_mu + m_.y_sigma * p);
                } else {
                    have_feats_ = gw_block_features(ctx_, age_.data(), feats_.data(), feats_.size()) == GW_OK;
                }
            }
        }
        // fill rank: with the network (the cache cannot hold every block), V -- the fill
        // chooses the blocks the swap phase starts from, and V is what a block serves until
        // it is evicted; without it, the walkers a load would release now (an integer) and,
        // as a tie-break in [0, 0.5), the walkers that need the block at all
        rank_.resize(nb);
        rank0_.resize(nb);
        uint64_t dmax = 0;
        for (uint32_t b = 0; b < nb; ++b) dmax = std::max(dmax, demand_[b]);
        for (uint32_t b = 0; b < nb; ++b) {
            rank0_[b] = (double)add_[b] + 0.5 * (double)demand_[b] / ((double)dmax + 1.0);
            rank_[b] = use_model_ ? (double)pred_[b] : rank0_[b];
        }
    }

    // the non-resident block with the best rank among those whose walkers could run now
    int best_candidate(const std::vector<char>& cached, const std::vector<double>& rank) const
    {
        int best = -1;
        for (uint32_t c = 0; c < add_.size(); ++c) {
            if (cached[c] || add_[c] == 0) continue;
            if (best < 0 || rank[c] > rank[best]) best = (int)c;
        }
        return best;
    }

    // The network's swap on resident set C: among the plain rule's shortlist (scores at
    // least shortlist_frac of the best, at most `shortlist` candidates), the load c with
    // the largest V(c) - V(m(c)); m(c) is c's least-paired resident block, ties broken by
    // the smaller V.  Returns the load and the position in C it replaces.
    enum { kPlain = 0, kSticky = 1, kModel = 2 };

    bool swap_choice(const SchedInput& in, const std::vector<uint32_t>& C, const uint8_t* staged, uint32_t* load,
                     uint32_t* pos, int* kind) const
    {
        const uint32_t nb = in.nblocks;
        const uint64_t n = nb;
        if (C.empty()) return false;
        std::vector<char> cached(nb, 0);
        for (uint32_t b : C) cached[b] = 1;
        struct Cand { uint64_t score; uint32_t c, pos; };
        std::vector<Cand> cs;
        uint64_t best = 0;
        for (uint32_t c = 0; c < nb; ++c) {
            if (cached[c]) continue;
            uint64_t mn = ~0ull, all = 0;
            uint32_t mpos = 0;
            double mcost = 0;
            const double sw = sw_n_ ? steps_per_walker_ : 2.0;
            for (uint32_t j = 0; j < C.size(); ++j) {
                const uint64_t x = in.W[c * n + C[j]] + in.W[(uint64_t)C[j] * n + c];
                all += x;
                if (cfg_.value_victim) {
                    // steps the eviction of C[j] gives up: its walkers paired with c, and
                    // what it would serve if it stayed
                    const double cost = sw * (double)x + std::expm1((double)pred_[C[j]]);
                    if (j == 0 || cost < mcost) { mcost = cost; mpos = j; mn = x; }
                } else if (x < mn || (x == mn && pred_[C[j]] < pred_[C[mpos]])) {
                    mn = x;
                    mpos = j;
                }
            }
            const uint64_t s = all - mn;
            if (s == 0) continue;
            cs.push_back({s, c, mpos});
            best = std::max(best, s);
        }
        if (cs.empty()) return false;
        // best first; equal scores keep the lower block id first, as the plain rule does
        std::stable_sort(cs.begin(), cs.end(), [](const Cand& a, const Cand& b) { return a.score > b.score; });
        uint32_t len = 0;
        while (len < cs.size() && len < cfg_.shortlist && (double)cs[len].score >= cfg_.shortlist_frac * (double)best)
            ++len;
        int pick = 0;
        *kind = kPlain;
        bool sticky = false;
        if (cfg_.sticky && staged)
            for (uint32_t i = 0; i < len && !sticky; ++i)
                if (staged[cs[i].c]) { pick = (int)i; sticky = true; *kind = kSticky; }
        if (!sticky && model_decides()) {
            double pv = 0;
            for (uint32_t i = 0; i < len; ++i) {
                const double v = std::expm1((double)pred_[cs[i].c]) - std::expm1((double)pred_[C[cs[i].pos]]);
                if (i == 0 || v > pv) { pick = (int)i; pv = v; }
            }
            *kind = kModel;
        }
        *load = cs[pick].c;
        *pos = cs[pick].pos;
        return true;
    }

    // the reference's zero-copy threshold for a block of this size, in waiting walkers
    // (blocksize / 128 / 24, apps/Node2vec.cu:100-104): the learned tail never comes later
    static double zero_floor(double bytes) { return bytes / 3072.0; }

    bool tail_now(const SchedInput& in, uint32_t c)
    {
        const double bytes = (double)block_bytes(in.blocks[c]);
        // the reference's rule looks at the best walker count of all candidates
        uint64_t best_add = 0;
        for (uint32_t b = 0; b < in.nblocks; ++b)
            if (!was_resident_[b]) best_add = std::max(best_add, add_[b]);
        if ((double)best_add <= zero_floor(bytes)) {
            ++tail_threshold_;
            return true;
        }
        if (!use_model_ || cooldown_ > 0) return false;
        // time of the upload against the time zero-copy reads need for the steps it buys
        const double steps = std::expm1((double)std::max(0.f, pred_[c]));
        const double t_up = bytes / copy_bytes_per_ms();
        const double per_step = reads_per_step() / zc_reads_per_ms();
        if (steps * per_step * cfg_.tail_margin < t_up && (double)pending_hops_ * per_step < cfg_.tail_guard * t_up) {
            ++tail_model_;
            return true;
        }
        return false;
    }

    bool plain_decide(const SchedInput& in)
    {
        plain_.buckets = buckets;
        SchedInput pin = in;
        pin.ncblock = capacity(in);
        pin.zero_threshold = std::max<uint64_t>(in.zero_threshold, (uint64_t)zero_floor((double)max_block_bytes(in)));
        bool r = plain_.schedule(pin);
        buckets = plain_.buckets;
        return r;
    }

    void add_sample(const float* f, float y)
    {
        if (m_.xs.size() < (size_t)cfg_.max_samples * kF) {
            m_.xs.insert(m_.xs.end(), f, f + kF);
            m_.ys.push_back(y);
        } else {
            const size_t i = m_.ring++ % cfg_.max_samples;
            std::copy(f, f + kF, m_.xs.begin() + i * kF);
            m_.ys[i] = y;
        }
        ++m_.samples;
    }

    void train() { train(cfg_.train_steps); }
    void train(uint32_t steps)
    {
        const uint32_t n = (uint32_t)m_.ys.size();
        if (!n) return;
        {
            double s = 0, s2 = 0;
            for (float y : m_.ys) {
                s += y;
                s2 += (double)y * y;
            }
            const double mu = s / n;
            m_.y_mu = (float)mu;
            // floor: labels that barely vary would otherwise blow small errors up
            m_.y_sigma = (float)std::max(0.05, std::sqrt(std::max(0.0, s2 / n - mu * mu)));
            std::vector<double> xs(kF, 0.0), xs2(kF, 0.0);
            for (uint32_t i = 0; i < n; ++i)
                for (uint32_t j = 0; j < kF; ++j) {
                    const double v = m_.xs[(size_t)i * kF + j];
                    xs[j] += v;
                    xs2[j] += v * v;
                }
            m_.x_mu.assign(kF, 0.f);
            m_.x_sigma.assign(kF, 1.f);
            for (uint32_t j = 0; j < kF; ++j) {
                const double m = xs[j] / n;
                m_.x_mu[j] = (float)m;
                m_.x_sigma[j] = (float)std::max(0.05, std::sqrt(std::max(0.0, xs2[j] / n - m * m)));
            }
        }
        const uint32_t bs = std::min(cfg_.batch, n);
        std::vector<float> bx((size_t)bs * kF), by(bs);
        for (uint32_t e = 0; e < steps; ++e) {
            for (uint32_t i = 0; i < bs; ++i) {
                m_.rng = m_.rng * 6364136223846793005ULL + 1442695040888963407ULL;
                const uint32_t k = (uint32_t)((m_.rng >> 33) % n);
                std::copy(m_.xs.begin() + (size_t)k * kF, m_.xs.begin() + (size_t)(k + 1) * kF, bx.begin() + (size_t)i * kF);
                by[i] = (m_.ys[k] - m_.y_mu) / m_.y_sigma;
            }
            standardize(bx.data(), bs);
            // the steps of a round are queued back to back; only the last one waits (for its loss)
            float loss = 0.f;
            const bool last = e + 1 == steps;
            if (gw_mlp_step(ctx_, mlp_, bx.data(), by.data(), bs, cfg_.lr, last ? &loss : nullptr) != GW_OK) return;
            if (last) m_.loss = loss;
            ++m_.adam_steps;
        }
    }

    gw_ctx* ctx_;
    LearnedConfig cfg_;
    ModelState* state_;
    ModelState m_;
    gw_mlp* mlp_ = nullptr;
    PlainScheduler plain_;
    std::vector<float> age_, feats_, pred_;
    std::vector<double> rank_, rank0_;   // fill ranks: the one in use (V with the network), the walker rank
    std::vector<uint64_t> add_, demand_;
    std::vector<char> was_resident_;
    std::vector<Open> open_, keep_;
    std::vector<float> closed_f_, closed_y_;   // closed samples not yet scored and in the ring
    double reads_per_step_ = 0, steps_per_walker_ = 0;
    uint64_t reads_n_ = 0, sw_n_ = 0;
    uint64_t pending_hops_ = 0;
    bool have_feats_ = false, use_model_ = false;
    bool fits_ = false;   // every block fits the cache (this walk): the network is left out
    uint32_t cooldown_ = 0;
    uint64_t fill_model_ = 0, swap_model_ = 0, sticky_ = 0, swap_differs_ = 0, tail_model_ = 0, tail_threshold_ = 0;
    uint64_t fallbacks_ = 0, err_n_ = 0;
    double err_ = 0;
};

}  // namespace gwh
