// sched_fuzz.cpp -- runs the product PlainScheduler (gowalker_amd/host/schedule.hpp, the
// reference's max_choose_blocks, engine/schedule.hpp:70-138) on cases read from a file and
// writes its decisions, for tests/test_sched_fuzz.py to compare with the oracle's
// orc_max_choose.  Case layout (all little-endian): u32 nblocks, u32 ncblock, u64
// zero_threshold, u32 nbuckets, u32 buckets[nbuckets], u64 W[nblocks * nblocks].  Output
// per case, one line: "<accepted> <n> <buckets...>".
#include <cstdio>
#include <cstdint>
#include <vector>

#include "schedule.hpp"

template <class T> static bool rd(FILE* f, T* v, size_t n = 1) { return fread(v, sizeof(T), n, f) == n; }

int main(int argc, char** argv)
{
    if (argc != 3) return 2;
    FILE* in = fopen(argv[1], "rb");
    FILE* out = fopen(argv[2], "w");
    if (!in || !out) return 2;
    uint32_t nb, nc, nbk;
    uint64_t zt;
    while (rd(in, &nb)) {
        if (!rd(in, &nc) || !rd(in, &zt) || !rd(in, &nbk)) return 3;
        gwh::PlainScheduler s;
        s.buckets.resize(nbk);
        if (nbk && !rd(in, s.buckets.data(), nbk)) return 3;
        std::vector<uint64_t> W((size_t)nb * nb);
        if (!rd(in, W.data(), W.size())) return 3;
        std::vector<gw_block_desc> blocks(nb);
        gwh::SchedInput si{W.data(), nb, nc, zt, blocks.data(), 0, nullptr};
        const bool r = s.schedule(si);
        fprintf(out, "%d %zu", r ? 1 : 0, s.buckets.size());
        for (uint32_t b : s.buckets) fprintf(out, " %u", b);
        fprintf(out, "\n");
    }
    fclose(in);
    fclose(out);
    return 0;
}
