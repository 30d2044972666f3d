// Unit tests of the host layer that need no GPU: option parsing (api/cmdopts.hpp
// semantics), the on-disk format (names, meta, block files, truncation), the text edge
// list converter, the plain scheduler against the oracle's known answers, slot
// placement and the launch tuner's walk-by-walk choice.
#include <algorithm>
#include <cassert>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cmdopts.hpp"
#include "format.hpp"
#include "engine.hpp"
#include "schedule.hpp"
#include "tuner.hpp"

using namespace gwh;

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                    \
        }                                                                    \
    } while (0)

static void test_options(const std::string& dir)
{
    std::string cnf = dir + "/conf";
    ::mkdir(cnf.c_str(), 0755);
    FILE* f = std::fopen((cnf + "/randgraph.cnf").c_str(), "w");
    std::fprintf(f, "# comment\nlength = 77\nq=3.5\n");
    std::fclose(f);
    setenv("RANDGRAPH_ROOT", dir.c_str(), 1);
    const char* argv[] = {"Node2vec", "g.txt", "length", "40", "p", "0.5", "gpu", "length", "41", "--blocksize=256", "zero"};
    Options o;
    o.set_argc(11, argv);
    CHECK(o.get_int("length", 30) == 41);          // last pair wins (backwards scan)
    CHECK(o.get_float("p", 1.0f) == 0.5f);
    CHECK(o.get_float("q", 2.0f) == 3.5f);         // config file when argv has no pair
    CHECK(o.get_long("blocksize", 1024) == 256);   // --key=value
    CHECK(o.get_int("memory", 4) == 4);            // default
    CHECK(o.get_bool("gpu") && o.get_bool("zero") && !o.get_bool("cpu"));
    // a name as the last token has no value: not a pair (the scan starts at argc-2)
    const char* argv2[] = {"SOPR", "g.txt", "alpha"};
    Options o2;
    o2.set_argc(3, argv2);
    CHECK(o2.get_float("alpha", 0.2f) == 0.2f);
    CHECK(o2.get_bool("alpha"));
}

static void test_names()
{
    CHECK(remove_extension("/data/tw.txt") == "/data/tw");
    CHECK(remove_extension("noext") == "noext");
    CHECK(vert_blocks_name("/data/tw", 256ull << 20) == "/data/sowalker_256/tw_256MB.vert.blocks");
    CHECK(edge_blocks_name("tw", 1ull << 30) == "./sowalker_1024/tw_1024MB.edge.blocks");
    CHECK(beg_name("/d/g") == "/d/g.beg" && csr_name("/d/g") == "/d/g.csr" && meta_name("/d/g") == "/d/g.meta");
}

static void test_format(const std::string& dir)
{
    const std::string base = dir + "/g";
    // giant vertex 0 (20 edges) then two small vertices, 32-byte blocks (8 edges, 4 verts)
    std::vector<uint64_t> beg = {0, 20, 21, 22};
    Meta m{3, 22, 0};
    write_meta(base, m);
    Meta r = read_meta(base);
    CHECK(r.nvertices == 3 && r.nedges == 22 && r.min_vert == 0);
    CHECK(file_size(meta_name(base)) == 16);
    ::mkdir((dir + "/sowalker_0").c_str(), 0755);
    write_block_files(base, beg.data(), beg.size(), 32);
    auto vb = read_vector<uint32_t>(vert_blocks_name(base, 32));
    auto eb = read_vector<uint64_t>(edge_blocks_name(base, 32));
    CHECK((vb == std::vector<uint32_t>{0, 0, 1, 3}));
    CHECK((eb == std::vector<uint64_t>{0, 0, 20, 22}));
    auto t = make_block_table(vb, eb, 32);
    CHECK(t.size() == 3 && t[0].nverts == 0 && t[1].nverts == 1 && t[1].nedges == 8 && t[2].nedges == 2);
    // a truncated file is an error, not an assert
    write_all(base + ".short", "abc", 3);
    bool threw = false;
    try {
        read_vector<uint64_t>(base + ".short");
    } catch (const IoError&) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        read_meta(dir + "/missing");
    } catch (const IoError&) {
        threw = true;
    }
    CHECK(threw);
}

// preprocess/graph_converter.hpp:186-384 rules, worked by hand: comments skipped, ids
// below the first source padded, self-loops dropped (the first line still pads), gaps
// filled, lists sorted, duplicates kept, the largest target padded at the end
static void test_convert(const std::string& dir)
{
    const std::string txt = dir + "/conv.txt";
    FILE* f = std::fopen(txt.c_str(), "w");
    std::fprintf(f, "# header\n%% other\n2 2\n2 5\n2,3\n4\t1\n4 1\n4 0\n4 7\n");
    std::fclose(f);
    Meta m = convert_text(txt);
    CHECK(m.nvertices == 8 && m.nedges == 6 && m.min_vert == 0);
    auto beg = read_vector<uint64_t>(dir + "/conv.beg");
    auto csr = read_vector<uint32_t>(dir + "/conv.csr");
    CHECK((beg == std::vector<uint64_t>{0, 0, 0, 2, 2, 6, 6, 6, 6}));
    CHECK((csr == std::vector<uint32_t>{3, 5, 0, 1, 1, 7}));
    Meta r = read_meta(dir + "/conv");
    CHECK(r.nvertices == 8 && r.nedges == 6 && r.min_vert == 0);
    // the smallest id counts only edges that are kept
    f = std::fopen(txt.c_str(), "w");
    std::fprintf(f, "1 1\n3 4\n3 5\n");
    std::fclose(f);
    m = convert_text(txt);
    CHECK(m.nvertices == 3 && m.nedges == 2 && m.min_vert == 3);
    CHECK((read_vector<uint64_t>(dir + "/conv.beg") == std::vector<uint64_t>{0, 0, 0, 0, 2, 2, 2}));
    // a decreasing source is an error that names the line
    f = std::fopen(txt.c_str(), "w");
    std::fprintf(f, "3 1\n2 1\n");
    std::fclose(f);
    bool threw = false;
    try {
        convert_text(txt);
    } catch (const IoError& e) {
        threw = std::string(e.what()).find(":2:") != std::string::npos;
    }
    CHECK(threw);
    // round trip: a random graph with empty lists and duplicate targets, written as text
    // with each list reversed and self-loops added, converts back to the same .beg/.csr
    uint64_t z = 99;
    auto rnd = [&]() { z = z * 6364136223846793005ULL + 1442695040888963407ULL; return (uint32_t)(z >> 33); };
    const uint32_t n = 3000;
    std::vector<uint64_t> b{0};
    std::vector<uint32_t> c;
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t d = rnd() % 4 == 0 ? 0 : rnd() % 40;
        std::vector<uint32_t> l;
        for (uint32_t k = 0; k < d; ++k) {
            uint32_t t = rnd() % n;
            if (t == v) t = (t + 1) % n;
            l.push_back(t);
        }
        std::sort(l.begin(), l.end());
        c.insert(c.end(), l.begin(), l.end());
        b.push_back(c.size());
    }
    f = std::fopen(txt.c_str(), "w");
    for (uint32_t v = 0; v < n; ++v) {
        if (b[v + 1] > b[v]) std::fprintf(f, "%u %u\n", v, v);
        for (uint64_t e = b[v + 1]; e > b[v]; --e) std::fprintf(f, "%u\t%u\n", v, c[e - 1]);
    }
    std::fclose(f);
    m = convert_text(txt);
    auto rb = read_vector<uint64_t>(dir + "/conv.beg");
    auto rc = read_vector<uint32_t>(dir + "/conv.csr");
    CHECK(rc == c && m.nedges == c.size());
    // ids past the largest one seen (as source or target) have no entry
    uint32_t top = 0;
    for (uint32_t v = 0; v < n; ++v)
        if (b[v + 1] > b[v]) top = std::max(top, std::max(v, *std::max_element(c.begin() + b[v], c.begin() + b[v + 1])));
    CHECK(rb.size() == (size_t)top + 2 && std::equal(rb.begin(), rb.end(), b.begin()));
}

static void test_scheduler()
{
    // same known answers as tests/test_oracle_format.py
    std::vector<uint64_t> W(9, 0);
    W[1 * 3 + 0] = 5; W[0 * 3 + 1] = 2; W[2 * 3 + 2] = 6; W[2 * 3 + 0] = 1;
    PlainScheduler s;
    s.buckets = {0};
    SchedInput in{W.data(), 3, 2, 0, nullptr, 0};
    CHECK(s.schedule(in) && (s.buckets == std::vector<uint32_t>{0, 1}));
    s.buckets = {0};
    SchedInput in7{W.data(), 3, 2, 7, nullptr, 0};
    CHECK(!s.schedule(in7) && (s.buckets == std::vector<uint32_t>{0}));
    std::vector<uint64_t> W2(16, 0);
    W2[2 * 4 + 0] = 3; W2[0 * 4 + 2] = 1; W2[2 * 4 + 1] = 10;
    s.buckets = {0, 1};
    SchedInput inf{W2.data(), 4, 2, 0, nullptr, 0};
    CHECK(s.schedule(inf) && (s.buckets == std::vector<uint32_t>{2, 1}));
    // slot placement: resident blocks stay, new ones take free then unwanted slots
    SlotMap sm;
    sm.reset(2, 4);
    CHECK(sm.place(0, {0, 1}) == 0 && sm.place(1, {0, 1}) == 1 && sm.place(1, {0, 1}) == -1);
    CHECK(sm.place(2, {2, 1}) == 0 && sm.blk_slot[0] == -1 && sm.blk_slot[2] == 0);
    CHECK(sm.place(3, {2, 1}) == -2);
    // staging: a prefetched block takes a free slot, is not resident until committed, and
    // its slot is reusable once the block is no longer wanted
    SlotMap st;
    st.reset(3, 6);
    CHECK(st.place(0, {0}) == 0);
    CHECK(st.stage(1, {0}) == 1 && st.stage(2, {0, 1}) == 2 && st.stage(3, {0, 1, 2}) == -1);
    CHECK(st.blk_slot[1] == -1 && st.blk_stg[1] == 1 && st.free_slots({0}) == 2 && st.free_slots({0, 1, 2}) == 0);
    CHECK(st.stage(1, {0}) == -1);   // already staged
    st.commit(1);
    CHECK(st.blk_slot[1] == 1 && st.blk_stg[1] == -1 && st.slot_stg[1] == -1);
    // wanted {0, 1, 4}: block 4 goes to the slot of the unwanted staged block 2
    CHECK(st.place(4, {0, 1, 4}) == 2 && st.blk_stg[2] == -1 && st.blk_slot[4] == 2);
    CHECK(st.free_slots({0, 1, 4}) == 0 && st.place(5, {0, 1, 4, 5}) == -2);
}

// SOWalker (engine/schedule.hpp:460-580): the starting set by hand, then properties of the
// annealed set: size, distinct valid blocks, a walker able to step, determinism per seed
static void test_sowalker()
{
    // 5 blocks; walkers per (prev, cur): column sums 0:3 1:11 2:9 3:2 4:0
    std::vector<uint64_t> W(25, 0);
    auto w = [&](uint32_t p, uint32_t c, uint64_t x) { W[p * 5 + c] = x; };
    w(1, 1, 4); w(2, 1, 7); w(1, 2, 9); w(0, 0, 3); w(3, 3, 2);
    SchedInput in{W.data(), 5, 3, 0, nullptr, 0};
    // top two by column sum: 1 (11), 2 (9); the third maximises walkers paired with {1, 2}:
    // block 0 pairs none, 3 none, 4 none -> first candidate (0, ranked third) stays
    auto idx = SOWalkerScheduler::initial_order(in);
    CHECK(idx[0] == 1 && idx[1] == 2 && idx[2] == 0);
    w(4, 2, 5);   // column 2 now holds 14 walkers, and block 4 pairs 5 of them with {2, 1}
    idx = SOWalkerScheduler::initial_order(in);
    CHECK(idx[0] == 2 && idx[1] == 1 && idx[2] == 4);
    CHECK(SOWalkerScheduler::swap_budget(5, 3) == 10 && SOWalkerScheduler::swap_budget(200, 100) == (1ull << 40));
    for (uint32_t k = 1; k <= 6; ++k) {
        SchedInput ik{W.data(), 5, k, 0, nullptr, 0};
        SOWalkerScheduler a(7), b(7);
        CHECK(a.schedule(ik) && b.schedule(ik) && a.buckets == b.buckets);
        std::vector<uint32_t> s = a.buckets;
        std::sort(s.begin(), s.end());
        CHECK(s.size() == std::min<uint32_t>(k, 5) && std::unique(s.begin(), s.end()) == s.end() && s.back() < 5);
        uint64_t pairs = 0;
        for (uint32_t p : s)
            for (uint32_t c : s) pairs += W[p * 5 + c];
        CHECK(pairs > 0);
    }
    // one slot and only cross-block walkers: no set can make progress
    std::vector<uint64_t> X(9, 0);
    X[0 * 3 + 2] = 6;
    SchedInput one{X.data(), 3, 1, 0, nullptr, 0};
    SOWalkerScheduler z;
    CHECK(!z.schedule(one));
    X[1 * 3 + 1] = 1;   // a diagonal bucket: that block alone
    CHECK(z.schedule(one) && z.buckets == std::vector<uint32_t>{1});
}

// the engine's search-order model: evenly spread ids favour interpolation, lists of
// clustered small ids (an unpermuted power-law graph) do not
static void test_search_order_model()
{
    const uint64_t n = 1 << 14, d = 1024;
    std::vector<uint64_t> beg(n + 1);
    std::vector<uint32_t> even(n * d), skew(n * d);
    uint64_t z = 12345;
    auto rnd = [&]() { z = z * 6364136223846793005ULL + 1442695040888963407ULL; return (uint32_t)(z >> 33); };
    for (uint64_t v = 0; v < n; ++v) {
        beg[v] = v * d;
        for (uint64_t k = 0; k < d; ++k) {
            even[v * d + k] = rnd() % n;
            const uint32_t r = rnd() % 1000;
            skew[v * d + k] = (uint32_t)((uint64_t)r * r * r % n / 997 / 997);   // clustered near 0
        }
        std::sort(even.begin() + v * d, even.begin() + (v + 1) * d);
        std::sort(skew.begin() + v * d, skew.begin() + (v + 1) * d);
    }
    beg[n] = n * d;
    CHECK(prefer_interpolation(beg.data(), even.data(), n));
    CHECK(!prefer_interpolation(beg.data(), skew.data(), n));
}

static void test_launch_tuner()
{
    LaunchTuner t;
    const LaunchTuner::Key a{16ull << 30, 256ull << 20, 1000, 1000}, b{8ull << 30, 256ull << 20, 1000, 1000};
    const uint32_t nc = (uint32_t)LaunchTuner::candidates().size();
    CHECK(nc >= 4);
    CHECK(LaunchTuner::candidates()[0].launch.block == 256 && LaunchTuner::candidates()[0].launch.waves_per_cu == 16);
    CHECK(t.best(a) == 0);   // nothing timed: the engine's default
    // the first walk of a workload runs the default and is not timed
    CHECK(t.pick(a) == 0);
    t.record(a, 0, 1.0, 1000000000ull);
    CHECK(t.find(a) && (*t.find(a))[0].walks == 0);
    // then each walk runs the next untried candidate; candidate 2 is the fastest per step
    for (uint32_t i = 0; i < nc; ++i) {
        const uint32_t c = t.pick(a);
        CHECK(c == i);
        t.record(a, c, c == 2 ? 100.0 : 150.0 + c, 1000000000ull);
    }
    CHECK(t.pick(a) == 2 && t.best(a) == 2);
    t.record(a, 2, 300.0, 1000000000ull);   // a slower repeat moves the mean: (100 + 300) / 2
    CHECK(t.best(a) == 0);                  // 150 per 10^9 steps now beats 200
    const auto* tab = t.find(a);
    CHECK(tab && (*tab)[2].walks == 2 && (*tab)[2].ms_per_gstep == 200.0);
    // another workload keeps its own table
    CHECK(t.pick(b) == 0 && t.last_key() && *t.last_key() == b);
    CHECK(t.find(b) && (*t.find(b))[0].walks == 0);
}

int main(int argc, char** argv)
{
    std::string dir = argc > 1 ? argv[1] : "/tmp";
    test_options(dir);
    test_launch_tuner();
    test_names();
    test_format(dir);
    test_convert(dir);
    test_scheduler();
    test_sowalker();
    test_search_order_model();
    std::printf("host_unit: all passed\n");
    return 0;
}
