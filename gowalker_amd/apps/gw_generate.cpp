// gw_generate -- writes a synthetic graph in GOWalker's on-disk format (.beg/.csr/.meta,
// plus the sowalker_<MB>/ block files for a blocksize), generated on the GPU.  It stands
// in for the reference's text->CSR converter, which cannot run as shipped
// (api/constants.hpp:5 makes EDGE_SIZE 0; preprocess/graph_converter.hpp:227-231).
//   gw_generate <base> rmat scale 20 edges 16777216 undirected dedup [permute] seed 42 blocksize 1024
//   gw_generate <base> chunglu nverts 41600000 edges 1470000000 gamma 2.2 maxdeg 3000000 seed 42 [multi]
#include <cstdio>
#include <string>
#include <vector>

#include "cmdopts.hpp"
#include "format.hpp"

int main(int argc, const char** argv)
{
    using namespace gwh;
    if (argc < 3) {
        std::fprintf(stderr, "usage: gw_generate <base> rmat|chunglu [options]\n");
        return 1;
    }
    Options o;
    o.set_argc(argc, argv);
    const std::string base = argv[1], kind = argv[2];
    const uint64_t seed = o.get_long("seed", 42);
    gw_config gc{0, 0, 0};
    gw_ctx* ctx = nullptr;
    if (gw_create(&gc, &ctx) != GW_OK) {
        std::fprintf(stderr, "cuda false: %s\n", gw_last_error(nullptr));
        return 1;
    }
    gw_graph* g = nullptr;
    gw_status st;
    if (kind == "rmat") {
        uint32_t flags = (o.get_bool("undirected") ? (uint32_t)GW_GEN_UNDIRECTED : 0u) | (o.get_bool("permute") ? (uint32_t)GW_GEN_PERMUTE : 0u) |
                         (o.get_bool("dedup") ? (uint32_t)GW_GEN_DEDUP : 0u);
        st = gw_gen_rmat(ctx, (uint32_t)o.get_int("scale", 20), o.get_long("edges", 1ull << 24), o.get_float("a", 0.57f),
                         o.get_float("b", 0.19f), o.get_float("c", 0.19f), seed, flags, &g);
    } else if (kind == "chunglu") {
        st = gw_gen_chung_lu(ctx, o.get_long("nverts", 1000000), o.get_long("edges", 10000000), o.get_float("gamma", 2.2f),
                             o.get_long("maxdeg", 100000), seed, o.get_bool("multi") ? 0u : (uint32_t)GW_GEN_DEDUP, &g);
    } else {
        std::fprintf(stderr, "unknown generator %s\n", kind.c_str());
        gw_destroy(ctx);
        return 1;
    }
    if (st != GW_OK) {
        std::fprintf(stderr, "cuda false: %s\n", gw_last_error(ctx));
        gw_destroy(ctx);
        return 1;
    }
    uint64_t nv = 0, ne = 0;
    gw_graph_size(g, &nv, &ne);
    std::vector<uint64_t> beg(nv + 1);
    std::vector<uint32_t> csr(ne);
    st = gw_graph_read(ctx, g, beg.data(), csr.data());
    gw_graph_free(ctx, g);
    gw_destroy(ctx);
    if (st != GW_OK) {
        std::fprintf(stderr, "cuda false: graph readback failed\n");
        return 1;
    }
    try {
        write_all(beg_name(base), beg.data(), beg.size() * 8);
        write_all(csr_name(base), csr.data(), csr.size() * 4);
        Meta m;
        m.nvertices = (uint32_t)nv;   // every id 0..nv-1 is a vertex (min_vert 0)
        m.nedges = ne;
        m.min_vert = 0;
        write_meta(base, m);
        const uint64_t bs = o.get_long("blocksize", 0) * 1024ull * 1024ull;
        if (bs) write_block_files(base, beg.data(), beg.size(), bs);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    std::printf("wrote %s: %llu vertices, %llu arcs\n", base.c_str(), (unsigned long long)nv, (unsigned long long)ne);
    return 0;
}
