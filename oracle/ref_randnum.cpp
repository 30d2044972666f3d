// Driver that links the reference's own RandNum (util/util.hpp:219-302) to print golden
// vectors.  util.hpp uses vid_t/bid_t from api/types.hpp, which cannot be included here
// (it needs the CUDA runtime); the two integer typedefs are declared in this driver with
// the values api/types.hpp:12-14 gives them.  No reference source is copied.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <vector>
typedef uint32_t vid_t;
typedef uint32_t bid_t;
#include "util/util.hpp"

int main(int argc, char** argv)
{
    // usage: randnum_ref seed n_lrand max_val bound n_pairs
    uint64_t seed = strtoull(argv[1], nullptr, 10);
    int n_lrand = atoi(argv[2]);
    float max_val = (float)atof(argv[3]);
    size_t bound = strtoull(argv[4], nullptr, 10);
    int n_pairs = atoi(argv[5]);
    RandNum a(seed);
    printf("{\"seed\": %llu, \"lrand\": [", (unsigned long long)seed);
    for (int i = 0; i < n_lrand; i++) printf("%s%llu", i ? ", " : "", (unsigned long long)a.lRand());
    RandNum b(seed);
    printf("], \"max_val\": %.9g, \"bound\": %zu, \"pairs\": [", max_val, bound);
    for (int i = 0; i < n_pairs; i++) {
        float rv; size_t pos;
        b.inlineRandPair(max_val, bound, rv, pos);
        uint32_t bits; memcpy(&bits, &rv, 4);
        printf("%s[%u, %zu]", i ? ", " : "", bits, pos);
    }
    printf("]}\n");
    return 0;
}
