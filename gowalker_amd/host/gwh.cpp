// gwh.cpp -- libgwhost.so: the C entry of include/gwh.h over the C++ engine.
#include "gwh.h"

#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"
#include "format.hpp"
#include "learned.hpp"
#include "pack.hpp"
#include "schedule.hpp"
#include "tuner.hpp"

static std::string g_err;

struct gwh_model {
    gwh::ModelState state;     // the learned scheduler's network and samples
    gwh::LaunchTuner tuner;    // step-kernel launch timings per workload
};

struct gwh_packed {
    const uint64_t* beg = nullptr;
    const uint32_t* csr = nullptr;
    uint64_t entries = 0, block_bytes = 0;
    gwh::PackedGraph g;
};

extern "C" const char* gwh_last_error(void) { return g_err.c_str(); }

// Splits the graph as gwh_run does and packs every block whose encoding is smaller than its
// csr.
extern "C" gw_status gwh_pack_graph(const uint64_t* beg, const uint32_t* csr, uint64_t entries, uint64_t block_bytes,
                                    gwh_packed** out)
{
    using namespace gwh;
    if (!beg || !csr || entries < 2 || !out || block_bytes < 64) {
        g_err = "gwh_pack_graph: bad arguments";
        return GW_ERR_ARG;
    }
    *out = nullptr;
    std::unique_ptr<gwh_packed> pk(new gwh_packed());
    try {
        std::vector<uint32_t> vb(entries + 2);
        std::vector<uint64_t> eb(entries + 2);
        uint64_t n = 0;
        ck(gw_split_blocks(beg, entries, block_bytes, vb.data(), eb.data(), vb.size(), &n), nullptr, "gw_split_blocks");
        vb.resize(n);
        eb.resize(n);
        pk->g.build(beg, csr, make_block_table(vb, eb, block_bytes));
    } catch (const GwError& e) {
        g_err = e.what();
        return e.status;
    } catch (const std::exception& e) {
        g_err = e.what();
        return GW_ERR_NOMEM;
    }
    pk->beg = beg;
    pk->csr = csr;
    pk->entries = entries;
    pk->block_bytes = block_bytes;
    *out = pk.release();
    return GW_OK;
}

extern "C" gw_status gwh_packed_info(const gwh_packed* p, uint64_t* raw_bytes, uint64_t* packed_bytes, uint32_t* nblocks,
                                     uint32_t* blocks_packed)
{
    if (!p) {
        g_err = "gwh_packed_info: null argument";
        return GW_ERR_ARG;
    }
    if (raw_bytes) *raw_bytes = p->g.raw;
    if (packed_bytes) *packed_bytes = p->g.bytes;
    if (nblocks) *nblocks = (uint32_t)p->g.blocks.size();
    if (blocks_packed) *blocks_packed = p->g.npacked;
    return GW_OK;
}

extern "C" gw_status gwh_packed_destroy(gwh_packed* p)
{
    delete p;
    return GW_OK;
}

extern "C" gw_status gwh_model_create(gwh_model** out)
{
    if (!out) {
        g_err = "gwh_model_create: null argument";
        return GW_ERR_ARG;
    }
    *out = new gwh_model();
    return GW_OK;
}

extern "C" gw_status gwh_model_destroy(gwh_model* m)
{
    delete m;
    return GW_OK;
}

extern "C" gw_status gwh_model_info_get(const gwh_model* m, gwh_model_info* out)
{
    if (!m || !out) {
        g_err = "gwh_model_info_get: null argument";
        return GW_ERR_ARG;
    }
    *out = gwh_model_info{};
    out->samples = m->state.samples;
    out->walks = m->state.walks;
    out->adam_steps = m->state.adam_steps;
    out->loss = m->state.loss;
    out->has_params = m->state.params.empty() ? 0u : 1u;
    out->scored = m->state.scored;
    out->net_abs_err = m->state.scored ? m->state.net_abs_err / (double)m->state.scored : 0.0;
    out->mean_abs_err = m->state.scored ? m->state.mean_abs_err / (double)m->state.scored : 0.0;
    return GW_OK;
}

extern "C" gw_status gwh_model_tuning(const gwh_model* m, gwh_tune_entry* out, uint32_t cap, uint32_t* n,
                                      uint32_t* chosen)
{
    if (!m || !n) {
        g_err = "gwh_model_tuning: null argument";
        return GW_ERR_ARG;
    }
    *n = 0;
    const gwh::LaunchTuner::Key* k = m->tuner.last_key();
    const auto* t = k ? m->tuner.find(*k) : nullptr;
    if (!t) return GW_OK;
    uint32_t best = 0;
    bool have = false;
    for (uint32_t i = 0; i < t->size(); ++i) {
        const auto& e = (*t)[i];
        if (e.walks && (!have || e.ms_per_gstep < (*t)[best].ms_per_gstep)) { best = i; have = true; }
        if (out && i < cap) out[i] = gwh_tune_entry{e.launch, e.ms_per_gstep, e.walks, 0};
    }
    *n = (uint32_t)t->size();
    if (chosen) *chosen = best;
    return GW_OK;
}

extern "C" gw_status gwh_run(const uint64_t* beg, const uint32_t* csr, uint64_t entries, const gwh_engine_opts* opts,
                             const gw_walk_params* params, uint64_t nwalkers, uint32_t wpv, const uint32_t* sources,
                             gwh_report* report, uint64_t* hashes, uint32_t* visits)
{
    using namespace gwh;
    if (!beg || !csr || entries < 2 || !opts || !params || !report || opts->block_bytes < 64) {
        g_err = "gwh_run: bad arguments";
        return GW_ERR_ARG;
    }
    try {
        std::vector<uint32_t> vb(entries + 2);
        std::vector<uint64_t> eb(entries + 2);
        uint64_t n = 0;
        ck(gw_split_blocks(beg, entries, opts->block_bytes, vb.data(), eb.data(), vb.size(), &n), nullptr,
           "gw_split_blocks");
        vb.resize(n);
        eb.resize(n);
        EngineConfig ec;
        ec.device = opts->device;
        ec.memory_bytes = opts->memory_bytes;
        ec.zero_threshold = opts->zero_threshold;
        ec.launch = opts->launch;
        ec.verbose = opts->verbose != 0;
        if (opts->packed) {
            const gwh_packed* pk = opts->packed;
            if (pk->beg != beg || pk->csr != csr || pk->entries != entries || pk->block_bytes != opts->block_bytes) {
                g_err = "gwh_run: the packed graph was built from other arrays or another block size";
                return GW_ERR_ARG;
            }
            ec.packed = &pk->g.blocks;
        }
        ec.copy_lanes = (opts->flags & GWH_TWO_LANES) ? 2u : 1u;
        // launch tuning across walks: with a model and no launch given, this walk runs the
        // tuner's next candidate shape
        const bool tune = opts->model && opts->launch.block == 0 && (opts->flags & GWH_NO_TUNE) == 0;
        const LaunchTuner::Key tkey{opts->memory_bytes, opts->block_bytes, nwalkers, entries - 1};
        uint32_t tcand = 0;
        if (tune) {
            tcand = (opts->flags & GWH_TUNE_FREEZE) ? opts->model->tuner.best(tkey) : opts->model->tuner.pick(tkey);
            const gw_launch l = LaunchTuner::candidates()[tcand].launch;
            ec.launch = gw_launch{l.block, l.waves_per_cu, l.variant, 0};
            ec.auto_search = true;
        }
        Engine engine(ec, beg, csr, make_block_table(vb, eb, opts->block_bytes));
        WalkConfig wc;
        wc.params = *params;
        wc.nwalkers = nwalkers;
        wc.wpv = wpv;
        wc.sources = sources;
        wc.outputs = (hashes ? (uint32_t)GW_OUT_HASH : 0u) | (visits ? (uint32_t)GW_OUT_VISITS : 0u);
        std::unique_ptr<Scheduler> sched;
        LearnedScheduler* ls = nullptr;
        if (opts->scheduler == GWH_SCHED_LEARNED || opts->scheduler == GWH_SCHED_STAGING) {
            LearnedConfig lc;
            lc.use_model = opts->scheduler == GWH_SCHED_LEARNED;
            lc.swap_prefetch = (opts->flags & GWH_NO_SWAP_PREFETCH) == 0;
            lc.myopic_label = (opts->flags & GWH_MYOPIC_LABEL) != 0;
            lc.model_swap = (opts->flags & GWH_MODEL_SWAP) != 0;
            lc.value_victim = (opts->flags & GWH_VALUE_VICTIM) != 0;
            lc.sticky = (opts->flags & GWH_NO_STICKY) == 0;
            if (opts->prefetch_depth) lc.max_inflight = opts->prefetch_depth;
            ls = new LearnedScheduler(engine.ctx(), lc, opts->model ? &opts->model->state : nullptr);
            sched.reset(ls);
        } else if (opts->scheduler == GWH_SCHED_SOWALKER) {
            sched.reset(new SOWalkerScheduler());
        } else {
            sched.reset(new PlainScheduler());
        }
        RunReport r = engine.run(wc, *sched);
        if (ls && opts->model) ls->save();
        if (tune && !(opts->flags & GWH_TUNE_FREEZE)) opts->model->tuner.record(tkey, tcand, r.kernel_ms, r.steps);
        *report = gwh_report{};
        report->steps = r.steps;
        report->trials = r.trials;
        report->probes = r.probes;
        report->mix_steps = r.mix_steps;
        report->h2d_bytes = r.h2d_bytes;
        report->h2d_raw_bytes = r.h2d_raw_bytes;
        report->inplace_blocks = r.inplace_blocks;
        report->ring_blocks = r.ring_blocks;
        report->inline_blocks = r.inline_blocks;
        report->kernel_launches = r.kernel_launches;
        report->uploads = r.uploads;
        report->iterations = r.iterations;
        report->zero_copy_runs = r.zero_copy_runs;
        report->zero_copy_steps = r.zero_copy_steps;
        report->wall_s = r.wall_s;
        report->kernel_ms = r.kernel_ms;
        report->filter_ms = r.filter_ms;
        report->h2d_ms = r.h2d_ms;
        report->nslots = r.nslots;
        report->nblocks = r.nblocks;
        report->staged = r.staged;
        report->zc_touches = r.zc_touches;
        report->sched_ms = r.sched_ms;
        report->tune_ms = r.tune_ms;
        report->tuned = r.tuned;
        report->tune_candidate = tune ? tcand : 0xFFFFFFFFu;
        report->copy_lanes = r.copy_lanes;
        if (ls) {
            report->learned_decisions = ls->learned_decisions();
            report->fallbacks = ls->fallbacks();
            report->final_loss = ls->last_loss();
            report->tail_decisions = ls->tail_decisions();
            report->tail_model = ls->tail_model();
            report->tail_threshold = ls->tail_threshold();
            report->swap_model = ls->swap_model();
            report->swap_differs = ls->swap_differs();
            report->fill_model = ls->fill_model();
            report->model_samples = ls->samples();
            report->sticky_swaps = ls->sticky_swaps();
        }
        if (hashes) ck(gw_read_hashes(engine.ctx(), hashes, nwalkers), engine.ctx(), "gw_read_hashes");
        if (visits) ck(gw_read_visits(engine.ctx(), visits, entries - 1), engine.ctx(), "gw_read_visits");
    } catch (const GwError& e) {
        g_err = e.what();
        return e.status;
    } catch (const std::exception& e) {
        g_err = e.what();
        return GW_ERR_IO;
    }
    return GW_OK;
}
