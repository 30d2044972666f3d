// learned.hpp -- learned block scheduler: a 12-64-64-1 MLP (BF16 weights, FP32
// accumulation, trained on the GPU by libgw.so's kernels) scores blocks by the walk steps
// they will serve in the next iteration if resident.  It is trained online on the
// block-access trace the engine records (steps per block per iteration, gw_trace_read).
//
// Decision, given the scores of every block:
//   load  c* = the non-resident block with the highest score among those that have
//              walkers able to run right now (plain add score > zero_threshold);
//   evict j* = (cache full) the resident block with the lowest score; swap only when
//              score(c*) > score(j*).
// The plain walker-count rule (engine/schedule.hpp:70-138) is its fallback and judge: it
// decides while the model has too few samples, and it takes over for a cool-down period
// whenever the learned decisions' realised steps per uploaded byte fall below the plain
// decisions' running average.  Walks stay bit-identical either way (counter-based RNG).
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
    uint32_t min_samples = 256;    // plain decisions until this many trace samples
    uint32_t train_every = 2;      // iterations between training rounds
    uint32_t epochs = 16;          // Adam steps per training round
    uint32_t batch = 1024;
    uint32_t max_samples = 16384;  // ring buffer of trace samples
    float lr = 2e-3f;
    double judge_ratio = 0.8;      // learned yield/byte below ratio x plain's => fall back
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
        feats_.assign((size_t)nb * kF, 0.f);
        if (gw_block_features(ctx_, age_.data(), feats_.data(), feats_.size()) != GW_OK) return plain_decide(in);
        last_decision_learned_ = false;
        if (!mlp_ || samples_ < cfg_.min_samples || cooldown_ > 0) {
            if (cooldown_ > 0) --cooldown_;
            return plain_decide(in);
        }
        pred_.assign(nb, 0.f);
        if (gw_mlp_forward(ctx_, mlp_, feats_.data(), nb, pred_.data()) != GW_OK) return plain_decide(in);
        std::vector<char> cached(nb, 0);
        for (uint32_t b : buckets) cached[b] = 1;
        int best = -1;
        for (uint32_t c = 0; c < nb; ++c) {
            if (cached[c]) continue;
            // walkers that could run at once if c joined the resident set
            uint64_t gain = in.W[(uint64_t)c * nb + c];
            for (uint32_t j : buckets) gain += in.W[(uint64_t)c * nb + j] + in.W[(uint64_t)j * nb + c];
            if (gain == 0 || gain <= in.zero_threshold) continue;
            if (best < 0 || pred_[c] > pred_[best]) best = (int)c;
        }
        if (best < 0) return plain_decide(in);
        if (buckets.size() < in.ncblock) {
            buckets.push_back((uint32_t)best);
        } else {
            int victim = -1;
            for (uint32_t p = 0; p < buckets.size(); ++p)
                if (victim < 0 || pred_[buckets[p]] < pred_[buckets[victim]]) victim = (int)p;
            if (!(pred_[best] > pred_[buckets[victim]])) return plain_decide(in);
            buckets[victim] = (uint32_t)best;
        }
        last_decision_learned_ = true;
        ++learned_decisions_;
        return true;
    }

    void feedback(const SchedInput& in, const IterationOutcome& out) override
    {
        const uint32_t nb = in.nblocks;
        std::vector<uint64_t> trace(nb, 0);
        if (gw_trace_read(ctx_, trace.data(), nb) != GW_OK) return;
        // samples: (features at decision time, log1p(steps served while resident)) for
        // every block that was resident during the iteration
        std::vector<char> res(nb, 0);
        for (uint32_t b : buckets) res[b] = 1;
        for (uint32_t b = 0; b < nb; ++b) {
            if (!res[b] || feats_.size() < (size_t)(b + 1) * kF) continue;
            add_sample(&feats_[(size_t)b * kF], std::log1p((float)trace[b]));
        }
        for (uint32_t b = 0; b < nb; ++b) age_[b] = trace[b] ? 0.f : std::min(age_[b] + 1.f, 1024.f);
        // judge: realised steps per uploaded byte, learned vs plain decisions
        if (out.h2d_bytes) {
            const double y = (double)out.steps / (double)out.h2d_bytes;
            if (last_decision_learned_) {
                learned_yield_ = learned_n_ ? 0.8 * learned_yield_ + 0.2 * y : y;
                ++learned_n_;
                if (plain_n_ >= 4 && learned_n_ >= 4 && learned_yield_ < cfg_.judge_ratio * plain_yield_) {
                    cooldown_ = cfg_.cooldown;
                    learned_n_ = 0;
                    ++fallbacks_;
                }
            } else {
                plain_yield_ = plain_n_ ? 0.8 * plain_yield_ + 0.2 * y : y;
                ++plain_n_;
            }
        }
        if (mlp_ && ++since_train_ >= cfg_.train_every && samples_ >= std::min<uint32_t>(cfg_.min_samples, 64)) {
            since_train_ = 0;
            train();
        }
    }

    uint64_t learned_decisions() const { return learned_decisions_; }
    uint64_t fallbacks() const { return fallbacks_; }
    float last_loss() const { return last_loss_; }

private:
    static constexpr uint32_t kF = 12;

    bool plain_decide(const SchedInput& in)
    {
        plain_.buckets = buckets;
        bool r = plain_.schedule(in);
        buckets = plain_.buckets;
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
        for (uint32_t e = 0; e < cfg_.epochs; ++e) {
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
    std::vector<float> age_, feats_, pred_, xs_, ys_;
    uint64_t samples_ = 0, ring_ = 0, rng_ = 0x853c49e6748fea9bULL;
    uint32_t since_train_ = 0, cooldown_ = 0;
    bool last_decision_learned_ = false;
    uint64_t learned_decisions_ = 0, fallbacks_ = 0, learned_n_ = 0, plain_n_ = 0;
    double learned_yield_ = 0, plain_yield_ = 0;
    float last_loss_ = 0.f;
};

}  // namespace gwh
