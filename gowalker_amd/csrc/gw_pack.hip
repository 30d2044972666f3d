// gw_pack.hip -- the transfer encoding of a block's adjacency (a packed block) and its
// decoder.  The reference uploads each block's csr as it is on disk, 4 bytes per arc
// (gpu/memory.hpp:124-139); out of core the walk is bound by those PCIe copies.  Sorted
// adjacency lists have small gaps, so a block is sent as gaps, bit-packed per 256-arc chunk,
// and a decode kernel on the copy stream rebuilds the csr words in the HBM slot: the slot's
// contents, and so every walk, are the same as after a raw upload.
//
// Packed block of nv vertices and ne arcs (offsets relative to the block's first arc):
//   chunk headers   ceil(ne / 256) x {u32 word offset of the chunk's payload, u32 bit width
//                   w (0..32), u32 carry (the arc before the chunk when the chunk starts
//                   inside a list, else 0), u32 first vertex (block-local id of the vertex
//                   whose list holds the chunk's first arc)}
//   starts          nv x u32: the first neighbour of each vertex (0 for an empty list)
//   payload         per chunk 8w u32 words: arc j of the chunk holds, in bits [jw, jw + w),
//                   its gap to the arc before it in the same list (0 for a list's first arc)
// A block whose packed form is not smaller than its csr is sent raw by the engine.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gw_internal.h"

namespace gw {

constexpr uint32_t kPackChunk = 256;

static inline uint32_t bits_for(uint32_t x) { return x ? 32u - (uint32_t)__builtin_clz(x) : 0u; }

// One workgroup of 256 threads decodes one chunk.  Thread i: arc c * 256 + i.
// The chunks [c_begin, c_begin + gridDim.x) of a packed block whose headers and starts are
// at pk and whose payload words from `word_base` on are at pk + payload_off.  One wavefront
// per chunk (a workgroup of 64), lane l holding arcs 4l .. 4l + 3: many chunks in flight per
// CU next to the walk kernels, no workgroup barrier beyond the wave's own.
__global__ void __launch_bounds__(64) unpack_kernel(const uint32_t* __restrict__ pk, const uint64_t* __restrict__ beg,
                                                    uint64_t start_edge, uint32_t nv, uint64_t ne,
                                                    uint64_t nchunks, uint64_t starts_off, uint64_t payload_off,
                                                    uint64_t c_begin, uint32_t word_base, uint32_t* __restrict__ out)
{
    constexpr uint32_t kPer = kPackChunk / 64;
    __shared__ uint32_t mark[kPackChunk];      // block-local vertex whose list starts at arc i, or ~0
    const uint64_t c = c_begin + blockIdx.x;
    if (c >= nchunks) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t* hdr = pk + c * 4;
    const uint32_t word = hdr[0], w = hdr[1], carry = hdr[2], vfirst = hdr[3];
    const uint32_t vlast = c + 1 < nchunks ? pk[(c + 1) * 4 + 3] : nv - 1;
    const uint64_t a0 = c * kPackChunk;
    // the lane's 4w payload bits (up to 5 words: the bits may start inside a word)
    uint32_t pw[5] = {0, 0, 0, 0, 0};
    if (w) {
        const uint32_t* pl = pk + payload_off + (word - word_base);
        const uint32_t bit = lane * kPer * w, k = bit >> 5, nw = (((bit & 31) + kPer * w) + 31) >> 5;
#pragma unroll
        for (uint32_t t = 0; t < 5; ++t)
            if (t < nw) pw[t] = pl[k + t];
    }
#pragma unroll
    for (uint32_t t = 0; t < kPer; ++t) mark[lane * kPer + t] = ~0u;
    __syncthreads();
    // list starts inside the chunk: vertices from the chunk's first to the next chunk's first
    for (uint64_t v = (uint64_t)vfirst + lane; v <= vlast; v += 64) {
        const uint64_t p = beg[v] - start_edge, q = beg[v + 1] - start_edge;
        if (p >= a0 && p < a0 + kPackChunk && q > p) mark[p - a0] = (uint32_t)v;
    }
    __syncthreads();
    // the lane's 4 arcs: (flag, value) with flag = a list starts here (or the chunk's carry)
    uint32_t val[kPer], flg[kPer];
    const uint32_t s0 = (lane * kPer * w) & 31;
#pragma unroll
    for (uint32_t t = 0; t < kPer; ++t) {
        const uint64_t a = a0 + lane * kPer + t;
        val[t] = 0;
        flg[t] = 0;
        if (a >= ne) continue;
        const uint32_t m = mark[lane * kPer + t];
        if (m != ~0u) {
            val[t] = pk[starts_off + m];
            flg[t] = 1;
        } else if (w) {
            const uint32_t b = s0 + t * w, k = b >> 5, sh = b & 31;
            uint32_t x = pw[0];
#pragma unroll
            for (uint32_t u = 1; u < 5; ++u) x = (k == u) ? pw[u] : x;
            uint32_t y = 0;
#pragma unroll
            for (uint32_t u = 1; u < 5; ++u) y = (k + 1 == u) ? pw[u] : y;
            x >>= sh;
            if (sh + w > 32) x |= y << (32 - sh);
            val[t] = w == 32 ? x : (x & ((1u << w) - 1u));
        }
        if (a == a0 && !flg[t]) {   // the chunk continues a list of the previous chunk
            val[t] += carry;
            flg[t] = 1;
        }
    }
    // segmented sums: (f1, v1) . (f2, v2) = (f1 | f2, f2 ? v2 : v1 + v2); first within the
    // lane, then across the wave over the lanes' totals
#pragma unroll
    for (uint32_t t = 1; t < kPer; ++t)
        if (!flg[t]) {
            val[t] += val[t - 1];
            flg[t] = flg[t - 1];
        }
    uint32_t tv = val[kPer - 1], tf = flg[kPer - 1];   // inclusive over lanes 0..lane
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t ov = __shfl_up(tv, off), of = __shfl_up(tf, off);
        if (lane >= off && !tf) {
            tv += ov;
            tf = of;
        }
    }
    // what the lanes before this one carry into its first segment
    uint32_t in = __shfl_up(tv, 1);
    if (lane == 0) in = 0;
#pragma unroll
    for (uint32_t t = 0; t < kPer; ++t) {
        const uint64_t a = a0 + lane * kPer + t;
        const uint32_t x = flg[t] ? val[t] : val[t] + in;
        if (a < ne) out[a] = x;
    }
}

hipError_t launch_unpack(const uint32_t* staging, const uint64_t* beg, uint64_t start_edge, uint32_t nv, uint64_t ne,
                         uint64_t c_begin, uint64_t c_end, uint32_t word_base, uint32_t* out, hipStream_t s)
{
    const uint64_t nch = (ne + kPackChunk - 1) / kPackChunk;
    if (c_end > nch) c_end = nch;
    if (c_begin >= c_end) return hipSuccess;
    const uint64_t starts_off = nch * 4, payload_off = starts_off + nv;
    hipLaunchKernelGGL(unpack_kernel, dim3((uint32_t)(c_end - c_begin)), dim3(64), 0, s, staging, beg,
                       start_edge, nv, ne, nch, starts_off, payload_off, c_begin, word_base, out);
    return hipGetLastError();
}

}  // namespace gw

using namespace gw;

// Packed size of a block (host; no GPU), and the packing itself.  beg: the block's nv + 1
// global offsets; csr: its ne arcs (the lists clipped at ne for a truncated giant vertex).
static gw_status pack_impl(const uint64_t* beg, const uint32_t* csr, uint32_t nv, uint64_t ne, uint32_t* out,
                           uint64_t cap_words, uint64_t* words)
{
    if (!beg || (ne && !csr) || !words) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: null argument");
    const uint64_t se = beg[0];
    const uint64_t nch = (ne + kPackChunk - 1) / kPackChunk;
    if (nch > 0xFFFFFFFFull) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: block too large");
    std::vector<uint32_t> width(nch), vfirst(nch), carry(nch);
    std::vector<uint64_t> woff(nch + 1, 0);
    // gap of arc a (0 for a list's first arc), with lists clipped at ne
    auto is_start_of = [&](uint64_t a, uint32_t v) { return beg[v] - se == a; };
    for (int64_t c = 0; c < (int64_t)nch; ++c) {
        const uint64_t a0 = (uint64_t)c * kPackChunk, a1 = std::min<uint64_t>(a0 + kPackChunk, ne);
        // vertex of arc a0: the last v with beg[v] - se <= a0 and a non-empty list
        uint64_t lo = 0, hi = nv;   // invariant: beg[lo] - se <= a0 < beg[hi] - se (beg[nv] >= ne)
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) / 2;
            if (beg[mid] - se <= a0) lo = mid; else hi = mid;
        }
        uint32_t v = (uint32_t)lo, mx = 0;
        vfirst[c] = v;
        carry[c] = is_start_of(a0, v) ? 0u : csr[a0 - 1];
        for (uint64_t a = a0; a < a1; ++a) {
            while (v + 1 <= nv - 1 && beg[v + 1] - se <= a) ++v;
            if (!is_start_of(a, v) && a > a0) {
                const uint32_t g = csr[a] - csr[a - 1];
                mx = g > mx ? g : mx;
            } else if (!is_start_of(a, v)) {
                mx = std::max(mx, csr[a] - carry[c]);
            }
        }
        width[c] = bits_for(mx);
    }
    for (uint64_t c = 0; c < nch; ++c) woff[c + 1] = woff[c] + 8ull * width[c];
    const uint64_t total = nch * 4 + nv + woff[nch];
    *words = total;
    if (!out) return GW_OK;
    if (cap_words < total) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: output buffer too small");
    uint32_t* hdr = out;
    uint32_t* starts = out + nch * 4;
    uint32_t* pl = starts + nv;
    for (int64_t v = 0; v < (int64_t)nv; ++v) {
        const uint64_t p = beg[v] - se, q = std::min<uint64_t>(beg[v + 1] - se, ne);
        starts[v] = p < q ? csr[p] : 0u;
    }
    for (int64_t c = 0; c < (int64_t)nch; ++c) {
        const uint64_t a0 = (uint64_t)c * kPackChunk, a1 = std::min<uint64_t>(a0 + kPackChunk, ne);
        const uint32_t w = width[c];
        hdr[c * 4 + 0] = (uint32_t)woff[c];
        hdr[c * 4 + 1] = w;
        hdr[c * 4 + 2] = carry[c];
        hdr[c * 4 + 3] = vfirst[c];
        uint32_t* p = pl + woff[c];
        std::memset(p, 0, 32ull * w);
        if (!w) continue;
        uint32_t v = vfirst[c];
        for (uint64_t a = a0; a < a1; ++a) {
            while (v + 1 <= nv - 1 && beg[v + 1] - se <= a) ++v;
            uint32_t g = 0;
            if (!is_start_of(a, v)) g = a > a0 ? csr[a] - csr[a - 1] : csr[a] - carry[c];
            const uint32_t bit = (uint32_t)(a - a0) * w, k = bit >> 5, s = bit & 31;
            p[k] |= g << s;
            if (s + w > 32) p[k + 1] |= g >> (32 - s);
        }
    }
    return GW_OK;
}

extern "C" gw_status gw_pack_block(const uint64_t* beg, const uint32_t* csr, uint32_t nverts, uint64_t nedges,
                                   void* out, uint64_t cap_bytes, uint64_t* bytes)
{
    if (!bytes) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: null argument");
    if (nverts == 0 && nedges) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: arcs without vertices");
    if (nverts && beg && beg[nverts] - beg[0] < nedges)
        return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: nedges exceeds the block's offsets");
    if (out && ((uintptr_t)out & 3)) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: output not 4-byte aligned");
    if (nedges == 0) {
        *bytes = 0;
        return GW_OK;
    }
    uint64_t words = 0;
    gw_status st = pack_impl(beg, csr, nverts, nedges, (uint32_t*)out, cap_bytes / 4, &words);
    *bytes = words * 4;
    return st;
}
