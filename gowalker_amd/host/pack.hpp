// pack.hpp -- a graph's blocks in the transfer encoding of gw_pack_block, in pinned host
// memory, for packed uploads (gw_upload_block_packed / gw_stage_block_packed).  Blocks
// without arcs, or whose encoding is not smaller than their beg and csr, stay raw
// (PackedBlock::data == nullptr).
#pragma once
#include <cstdint>
#include <vector>

#include "engine.hpp"

namespace gwh {

struct PackedGraph {
    std::vector<PackedBlock> blocks;
    std::vector<void*> pinned;
    uint64_t raw = 0, bytes = 0;   // the blocks' beg and csr bytes; what their uploads send
    uint32_t npacked = 0;
    PackedGraph() = default;
    PackedGraph(const PackedGraph&) = delete;
    PackedGraph& operator=(const PackedGraph&) = delete;
    ~PackedGraph()
    {
        for (void* p : pinned) gw_free_pinned(p);
    }

    // sizes in parallel over blocks, pinned buffers one by one, then packing in parallel
    void build(const uint64_t* beg, const uint32_t* csr, const std::vector<gw_block_desc>& desc)
    {
        const int64_t nb = (int64_t)desc.size();
        std::vector<uint64_t> sz(nb, 0);
        std::vector<gw_status> st(nb, GW_OK);
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t b = 0; b < nb; ++b) {
            const gw_block_desc& d = desc[b];
            if (d.nedges) st[b] = gw_pack_block(beg + d.start_vert, csr + d.start_edge, d.nverts, d.nedges, nullptr, 0,
                                                &sz[b]);
        }
        for (int64_t b = 0; b < nb; ++b) ck(st[b], nullptr, "gw_pack_block (size)");
        blocks.assign(nb, PackedBlock{});
        for (int64_t b = 0; b < nb; ++b) {
            const uint64_t r = ((uint64_t)desc[b].nverts + 1) * 8 + desc[b].nedges * 4;
            raw += r;
            if (!desc[b].nedges || sz[b] >= r) {
                bytes += r;
                continue;
            }
            void* p = nullptr;
            ck(gw_alloc_pinned(sz[b], &p), nullptr, "gw_alloc_pinned");
            pinned.push_back(p);
            blocks[b] = PackedBlock{p, sz[b]};
            bytes += sz[b];
            npacked++;
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t b = 0; b < nb; ++b) {
            const gw_block_desc& d = desc[b];
            const PackedBlock& q = blocks[b];
            if (!q.data) continue;
            uint64_t got = 0;
            st[b] = gw_pack_block(beg + d.start_vert, csr + d.start_edge, d.nverts, d.nedges, const_cast<void*>(q.data),
                                  q.bytes, &got);
        }
        for (int64_t b = 0; b < nb; ++b) ck(st[b], nullptr, "gw_pack_block");
    }
};

}  // namespace gwh
