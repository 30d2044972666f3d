// app.hpp -- the Node2vec / SOPR / SOSR entry points (apps/Node2vec.cu, apps/SOPR.cu,
// apps/SOSR.cu) over libgw.so.  Same option names and defaults as the reference
// (apps/Node2vec.cu:35-59); additions: seed, term, sched (plain|learned|staging), source
// (reference|pervertex, SOPR/SOSR), pack, visits <file>, hashes <file>, verbose.
#pragma once
#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "cmdopts.hpp"
#include "engine.hpp"
#include "format.hpp"
#include "learned.hpp"
#include "metrics.hpp"
#include "pack.hpp"
#include "schedule.hpp"

namespace gwh {

enum class App { Node2vec, SOPR, SOSR };

// Whole graph in pinned host memory (the reference preads every block into pinned
// buffers at start-up, engine/engine.hpp:74-77).
struct HostGraph {
    uint64_t* beg = nullptr;
    uint32_t* csr = nullptr;
    uint64_t entries = 0, nedges = 0;
    ~HostGraph()
    {
        if (beg) gw_free_pinned(beg);
        if (csr) gw_free_pinned(csr);
    }
    void load(const std::string& base)
    {
        const std::string bn = beg_name(base), cn = csr_name(base);
        uint64_t bsz = file_size(bn), csz = file_size(cn);
        if (bsz % 8 || bsz < 16) throw IoError(bn + " is not a u64 offset array");
        if (csz % 4) throw IoError(cn + " is not a u32 array");
        entries = bsz / 8;
        nedges = csz / 4;
        if (gw_alloc_pinned(bsz, (void**)&beg) != GW_OK || gw_alloc_pinned(csz ? csz : 4, (void**)&csr) != GW_OK)
            throw IoError(std::string("pinned allocation failed: ") + gw_last_error(nullptr));
        int fb = ::open(bn.c_str(), O_RDONLY), fc = ::open(cn.c_str(), O_RDONLY);
        if (fb < 0 || fc < 0) {
            if (fb >= 0) ::close(fb);
            if (fc >= 0) ::close(fc);
            throw IoError("cannot open " + bn + " / " + cn);
        }
        try {
            pread_all(fb, beg, bsz, 0, bn);
            const uint64_t chunk = 1ull << 28;
#pragma omp parallel for schedule(dynamic, 1)
            for (int64_t i = 0; i < (int64_t)((csz + chunk - 1) / chunk); ++i) {
                uint64_t off = (uint64_t)i * chunk, len = std::min(chunk, csz - off);
                pread_all(fc, (char*)csr + off, len, off, cn);
            }
        } catch (...) {
            ::close(fb);
            ::close(fc);
            throw;
        }
        ::close(fb);
        ::close(fc);
        if (beg[0] != 0 || beg[entries - 1] != nedges)
            throw IoError(bn + " does not match " + cn + " (first/last offsets)");
        for (uint64_t v = 1; v < entries; ++v)
            if (beg[v] < beg[v - 1]) throw IoError(bn + " offsets decrease at vertex " + std::to_string(v - 1));
        uint32_t mx = 0;
#pragma omp parallel for reduction(max : mx)
        for (int64_t i = 0; i < (int64_t)nedges; ++i) mx = std::max(mx, csr[i]);
        if (nedges && (uint64_t)mx >= entries - 1)
            throw IoError(cn + " holds vertex id " + std::to_string(mx) + " beyond the " +
                          std::to_string(entries - 1) + " ids of " + bn);
    }
};

inline void print_help()
{
    std::printf(
        "- dataset:       a text edge list sorted by source, converted on first use to\n"
        "                 <base>.beg/.csr/.meta (or those files, e.g. from gw_generate)\n"
        "- blocksize:     the size of each block in MB (default 1024)\n"
        "- memory:        the HBM budget in GB for graph slots and walker state (default 4)\n"
        "- walkpersource: the number of walks for each vertex (default 1)\n"
        "- numwalks:      walkers from uniformly drawn sources instead (default 0 = off)\n"
        "- length:        the number of steps for each walk (default 30)\n"
        "- p, q:          node2vec parameters (default 0.5, 2)\n"
        "- alpha:         SOPR/SOSR second-order weight (default 0.2)\n"
        "- term:          geometric termination probability per step (default 0)\n"
        "- zero:          finish with zero-copy reads below blocksize/128/24 walkers\n"
        "- GOWalker | SOWalker: block scheduler, walker-count rule or simulated annealing (one required)\n"
        "- sched:         plain | learned | staging (the learned scheduler without its network; default plain)\n"
        "- pack:          upload blocks as bit-packed gaps, decoded on the GPU (fewer PCIe bytes, same walks)\n"
        "- seed:          Philox seed (default 42)\n"
        "- threadperblock: workgroup size of the step kernel (default: the engine's)\n"
        "- source:        SOPR/SOSR sources: reference (vertex 1, or 1 and 3) | pervertex\n"
        "- visits <file>, hashes <file>: write per-vertex visit counts / per-walker path hashes\n"
        "- h:             print this message\n");
}

inline uint32_t uniform_source(uint64_t seed, uint64_t w, uint64_t nverts)
{
    // splitmix64 of (seed, w); a multiply-shift maps it onto the vertex range
    uint64_t z = seed * 0x9E3779B97F4A7C15ULL + w + 0x632BE59BD9B4E019ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (uint32_t)(((z >> 32) * nverts) >> 32);
}

inline int app_main(int argc, const char** argv, App app)
{
    Options o;
    o.set_argc(argc, argv);
    if (argc < 2 || o.get_bool("h")) {
        print_help();
        return argc < 2 ? 1 : 0;
    }
    const std::string input = argv[1];
    const std::string base = remove_extension(input);
    const uint64_t blocksize = o.get_long("blocksize", 1024) * 1024ull * 1024ull;
    const uint64_t memory = (uint64_t)o.get_int("memory", 4) * 1024ull * 1024ull * 1024ull;
    uint32_t wpv = (uint32_t)o.get_int("walkpersource", 1);
    uint64_t numwalks = (uint64_t)o.get_int("numwalks", 0);
    const uint32_t length = (uint32_t)o.get_int("length", 30);
    const double p = o.get_float("p", 0.5f), q = o.get_float("q", 2.0f), alpha = o.get_float("alpha", 0.2f);
    const double term = o.get_float("term", 0.0f);
    // threadperblock (apps/Node2vec.cu:47, 512 there) fixes the workgroup size when given;
    // otherwise the engine picks its default shape and the search order by its probe model
    const uint32_t tpb = (uint32_t)o.get_int("threadperblock", 0);
    const bool zero = o.get_bool("zero");
    const bool verbose = o.get_bool("verbose");
    const uint64_t seed = o.get_long("seed", 42);
    const std::string sched_name = o.get_string("sched", "plain");
    bool go = o.get_bool("GOWalker"), so = o.get_bool("SOWalker");
    if (app != App::Node2vec) { go = true; so = false; }   // SOPR.cu:171, SOSR.cu:171
    if (go == so) {
        std::printf("Please select one walker from GOWalker and SOWalker!\n");
        return 0;
    }
    try {
        // preprocess/graph_converter.hpp:286-320: convert the text edge list unless its
        // .beg/.csr/.meta exist (existing files are kept, as with `skip`; there is no
        // interactive re-convert prompt)
        if (!file_exists(beg_name(base)) || !file_exists(csr_name(base)) || !file_exists(meta_name(base))) {
            if (input == base || !file_exists(input)) throw IoError("no preprocessed graph and no edge list " + input);
            const Meta cm = convert_text(input);
            if (verbose) std::printf("converted %s: %u vertices, %llu edges\n", input.c_str(), cm.nvertices, (unsigned long long)cm.nedges);
        }
        Meta meta = read_meta(base);
        HostGraph g;
        g.load(base);
        if (meta.nedges != g.nedges) throw IoError("meta edge count differs from the .csr size");
        if (!file_exists(vert_blocks_name(base, blocksize)) || !file_exists(edge_blocks_name(base, blocksize)))
            write_block_files(base, g.beg, g.entries, blocksize);
        auto vb = read_vector<uint32_t>(vert_blocks_name(base, blocksize));
        auto eb = read_vector<uint64_t>(edge_blocks_name(base, blocksize));
        std::vector<gw_block_desc> blocks = make_block_table(vb, eb, blocksize);
        if (vb.back() != g.entries - 1) throw IoError("block files do not cover the graph");
        const uint64_t nids = vb.back();

        EngineConfig ec;
        ec.memory_bytes = memory;
        ec.zero_threshold = zero ? blocksize / 128 / 24 : 0;   // apps/Node2vec.cu:100-108
        if (tpb) ec.launch.block = std::max<uint32_t>(64, std::min<uint32_t>(1024, tpb / 64 * 64));
        ec.verbose = verbose;
        PackedGraph packed;
        double pack_s = 0;
        if (o.get_bool("pack")) {
            const double t0 = omp_get_wtime();
            packed.build(g.beg, g.csr, blocks);
            pack_s = omp_get_wtime() - t0;
            ec.packed = &packed.blocks;
            if (verbose)
                std::printf("packed %u of %zu blocks: %llu of %llu csr bytes in %.2f s\n", packed.npacked, blocks.size(),
                            (unsigned long long)packed.bytes, (unsigned long long)packed.raw, pack_s);
        }
        Engine engine(ec, g.beg, g.csr, blocks);

        WalkConfig wc;
        wc.params.seed = seed;
        wc.params.ntypes = 1;
        gw_walk_type& t = wc.params.type[0];
        t.algo = app == App::Node2vec ? GW_NODE2VEC : GW_SOPR;
        t.length = length;
        t.p = p;
        t.q = q;
        t.alpha = alpha;
        t.term = term;
        wc.outputs = (o.get_string("visits", "").empty() ? 0 : GW_OUT_VISITS) |
                     (o.get_string("hashes", "").empty() ? 0 : GW_OUT_HASH);
        std::vector<uint32_t> sources;
        if (app == App::Node2vec) {
            if (numwalks == 0) {
                wc.nwalkers = (uint64_t)wpv * nids;   // one batch of wpv walkers per vertex, walk.hpp:302-320
                wc.wpv = wpv;
            } else {
                sources.resize(numwalks);
                for (uint64_t w = 0; w < numwalks; ++w) sources[w] = uniform_source(seed, w, nids);
            }
        } else {
            const std::string src = o.get_string("source", "reference");
            if (src == "pervertex") {
                wc.nwalkers = (uint64_t)wpv * nids;
                wc.wpv = wpv;
            } else {
                // walk.hpp:275-292 (SOPR: every walker at vertex 1), :355-385 (SOSR: the
                // first half at 1, the rest at 3); numwalks defaults to wpv * nvertices
                uint64_t n = numwalks ? numwalks : (uint64_t)wpv * meta.nvertices;
                if (nids < 4) throw IoError("SOPR/SOSR reference sources need vertices 1 and 3");
                sources.assign(n, 1u);
                if (app == App::SOSR)
                    for (uint64_t w = n / 2; w < n; ++w) sources[w] = 3u;
            }
        }
        if (!sources.empty()) {
            wc.nwalkers = sources.size();
            wc.sources = sources.data();
        }

        // GOWalker: the walker-count rule (or `sched learned`); SOWalker: simulated annealing
        // over whole resident sets (apps/Node2vec.cu:157-163)
        std::unique_ptr<Scheduler> sched;
        if (so) sched.reset(new SOWalkerScheduler());
        else if (sched_name == "learned") sched.reset(new LearnedScheduler(engine.ctx(), LearnedConfig()));
        else if (sched_name == "staging") {
            LearnedConfig lc;
            lc.use_model = false;
            sched.reset(new LearnedScheduler(engine.ctx(), lc));
        }
        else sched.reset(new PlainScheduler());

        RunReport r = engine.run(wc, *sched);
        std::string mname = (app == App::Node2vec ? std::string("node2vec") : std::string("SOPR")) + "_numwalks_" +
                            std::to_string(wc.nwalkers) + "_steps_" + std::to_string(length) + "_dataset_" + input;
        if (app == App::Node2vec) mname += so ? "_schedule_SOWalker" : "_schedule_GOWalker";
        Metrics m(mname);
        m.set("Graph blocks", r.nblocks, true);
        m.set("Gpu memory blocks", r.nslots, true);
        m.set("0_Total_time", r.wall_s);
        m.set("1_GPUTime", (r.kernel_ms + r.filter_ms) / 1e3);
        m.set("1.1_transfer", r.h2d_ms / 1e3);
        m.set("1.3_GPU_update", r.kernel_ms / 1e3);
        m.set("walk_steps", (double)r.steps, true);
        m.set("walk_steps_per_s", r.wall_s > 0 ? r.steps / r.wall_s : 0.0);
        m.set("h2d_bytes", (double)r.h2d_bytes, true);
        m.set("h2d_gb", r.h2d_bytes / 1e9);
        m.set("h2d_raw_bytes", (double)r.h2d_raw_bytes, true);
        if (o.get_bool("pack")) m.set("pack_time", pack_s);
        m.set("iterations", (double)r.iterations, true);
        m.set("zero_copy_steps", (double)r.zero_copy_steps, true);
        m.set_string("scheduler", sched->name());
        m.report(o.get_string("metrics.reporter.filename", "randgraph_metrics.txt"));
        const std::string vf = o.get_string("visits", ""), hf = o.get_string("hashes", "");
        if (!vf.empty()) {
            std::vector<uint32_t> v(nids);
            ck(gw_read_visits(engine.ctx(), v.data(), v.size()), engine.ctx(), "gw_read_visits");
            write_all(vf, v.data(), v.size() * 4);
        }
        if (!hf.empty()) {
            std::vector<uint64_t> h(wc.nwalkers);
            ck(gw_read_hashes(engine.ctx(), h.data(), h.size()), engine.ctx(), "gw_read_hashes");
            write_all(hf, h.data(), h.size() * 8);
        }
    } catch (const std::exception& e) {
        // the reference's checkCudaError prints "cuda false:" and exits (api/types.hpp:28-35)
        std::fprintf(stderr, "cuda false: %s\n", e.what());
        return EXIT_FAILURE;
    }
    return 0;
}

}  // namespace gwh
