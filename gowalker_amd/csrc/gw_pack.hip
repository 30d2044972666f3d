// gw_pack.hip -- the transfer encoding of a block (a packed block) and its decoders.  The
// reference uploads each block's beg and csr as they are on disk, 8 bytes per vertex and 4
// per arc (gpu/memory.hpp:124-139); out of core the walk is bound by those PCIe copies.
// Sorted adjacency lists have small gaps and offsets small differences, so a block is sent
// as Elias-Fano-coded gap sums and bit-packed degrees, and two decode kernels on the copy
// stream rebuild beg and csr in the HBM slot: the slot's contents, and so every walk, are
// the same as after a raw upload.
//
// Packed block of nv vertices and ne > 0 arcs, in u32 words (arc and entry indices local to
// the block):
//   header          {magic 0x32505747, nvc = ceil((nv + 1) / 256), vpw, 0}
//   beg chunks      nvc x {base low, base high, payload word offset (from the first beg
//                   payload word), widths w0 | w1 << 8 | w2 << 16 | w3 << 24}: chunk k covers
//                   entries [256k, 256k + 256) of beg (nv + 1 entries), base = beg[256k]
//   beg payload     vpw words: per chunk, per group g of 64 entries, 2 w_g words; entry e of
//                   the group in bits [e w_g, (e + 1) w_g): beg[i] - beg[i - 1] (0 for the
//                   chunk's first entry)
//   arc chunks      ceil(ne / 256) x {payload word offset (from the first arc payload word),
//                   L | nz << 8, carry (the arc before the chunk when the chunk starts inside
//                   a list, else 0), first vertex (the block-local vertex whose list holds
//                   the chunk's first arc)}
//   arc payload     per chunk of n arcs: with g_j the gap of arc j -- its value minus the
//                   arc before it in the same list (minus carry for a continued list's first
//                   arc in the chunk), or its value at a list start, all mod 2^32 -- and
//                   S_j = g_0 + ... + g_j (64-bit, non-decreasing): 8L words holding S_j mod
//                   2^L in bits [jL, (j + 1)L), then ceil((n + nz) / 32) words with bit
//                   (S_j >> L) + j set for every j (the Elias-Fano upper bits; nz =
//                   S_{n-1} >> L zeros among them).  L = floor(log2(S_{n-1} / n)) when
//                   S_{n-1} > n, else 0: so L <= 31 and nz < 2n <= 512.
// Arc j of a chunk is then S_j - S_{s-1} for s the last list start at or before j in the
// chunk (S_{-1} = 0), or carry + S_j when no list starts there.  The bits an arc costs
// follow log2 of its chunk's mean gap (2 bits above it at most) instead of the widest gap's.
// A block whose packed form is not smaller than its beg and csr is sent raw by the engine.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gw_internal.h"

namespace gw {

constexpr uint32_t kPackChunk = kPackChunkArcs;

static inline uint32_t bits_for(uint64_t x) { return x ? 64u - (uint32_t)__builtin_clzll(x) : 0u; }

// bits [b, b + w) of the 5-word window pw (w <= 32, b + w <= 160)
__device__ __forceinline__ uint32_t window_bits(const uint32_t (&pw)[5], uint32_t b, uint32_t w)
{
    const uint32_t k = b >> 5, sh = b & 31;
    uint32_t x = pw[0], y = 0;
#pragma unroll
    for (uint32_t u = 1; u < 5; ++u) {
        x = (k == u) ? pw[u] : x;
        y = (k + 1 == u) ? pw[u] : y;
    }
    x >>= sh;
    if (sh + w > 32) x |= y << (32 - sh);
    return w >= 32 ? x : (x & ((1u << w) - 1u));
}

// Rebuilds beg from its packed chunks: one wavefront per chunk of 256 entries, lane l holding
// entries 4l .. 4l + 3 (of group l / 16); an in-lane then a wave-wide prefix sum.
// Reads stay inside the vpw payload words whatever the headers hold.
__global__ void __launch_bounds__(64) unpack_beg_kernel(const uint32_t* __restrict__ vh,
                                                        const uint32_t* __restrict__ vp, uint64_t vpw, uint32_t nv,
                                                        uint64_t* __restrict__ beg)
{
    const uint32_t k = blockIdx.x, lane = threadIdx.x;
    const uint32_t* h = vh + 4ull * k;
    const unsigned long long base = (unsigned long long)h[0] | ((unsigned long long)h[1] << 32);
    const uint32_t ws = h[3], g = lane >> 4;
    uint64_t goff = h[2];
    for (uint32_t i = 0; i < g; ++i) goff += 2u * min((ws >> (8 * i)) & 0xFFu, 32u);
    uint32_t w = min((ws >> (8 * g)) & 0xFFu, 32u);
    if (goff + 2 * w > vpw) w = 0;
    // the lane's 4w bits of its group's 64w (up to 5 words: they may start inside a word)
    uint32_t pw[5] = {0, 0, 0, 0, 0};
    const uint32_t bit = (lane & 15) * 4 * w;
    if (w) {
        const uint32_t kw = bit >> 5, nw = ((bit & 31) + 4 * w + 31) >> 5;
#pragma unroll
        for (uint32_t t = 0; t < 5; ++t)
            if (t < nw) pw[t] = vp[goff + kw + t];
    }
    unsigned long long x[4];
    unsigned long long run = 0;
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
        run += w ? window_bits(pw, (bit & 31) + t * w, w) : 0u;
        x[t] = run;
    }
    unsigned long long tot = run;   // inclusive over lanes 0..lane
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const unsigned long long o = __shfl_up(tot, off);
        if (lane >= off) tot += o;
    }
    const unsigned long long in = tot - run;
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
        const uint64_t i = (uint64_t)k * kPackChunk + lane * 4 + t;
        if (i <= nv) beg[i] = base + in + x[t];
    }
}

// Rebuilds the csr words of arc chunks [c_begin, c_begin + gridDim.x): one wavefront per
// chunk, lane l holding arcs 4l .. 4l + 3.  ah: the block's arc chunk headers; ap: its arc
// payload words from word_base on; beg: the block's nv + 1 global offsets (decoded first).
// Reads stay inside the `avail` payload words at ap and the nv + 1 offsets whatever the
// headers hold (a chunk whose payload would not is left undecoded).
__global__ void __launch_bounds__(64) unpack_arcs_kernel(const uint32_t* __restrict__ ah,
                                                         const uint32_t* __restrict__ ap, uint64_t word_base,
                                                         uint64_t avail, const uint64_t* __restrict__ beg,
                                                         uint64_t start_edge, uint32_t nv, uint64_t ne,
                                                         uint64_t nchunks, uint64_t c_begin,
                                                         uint32_t* __restrict__ out)
{
    __shared__ uint32_t hi[kPackChunk];              // S_j >> L
    __shared__ unsigned long long sum[kPackChunk];   // S_j
    __shared__ uint32_t mark[kPackChunk];            // 1: a list starts at arc j
    const uint64_t c = c_begin + blockIdx.x;
    if (c >= nchunks) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t* hdr = ah + c * 4;
    const uint32_t L = min(hdr[1] & 0xFFu, 31u), nz = (hdr[1] >> 8) & 0xFFFFu, carry = hdr[2], vfirst = hdr[3];
    const uint32_t vlast = min(c + 1 < nchunks ? ah[(c + 1) * 4 + 3] : nv - 1, nv - 1);
    const uint64_t a0 = c * kPackChunk;
    const uint32_t n = (uint32_t)min((uint64_t)kPackChunk, ne - a0);
    // upper bits: the set bit of arc j sits at (S_j >> L) + j, j-th among the set bits
    const uint32_t hw = min((n + nz + 31) >> 5, 24u);
    if ((uint64_t)hdr[0] < word_base || (uint64_t)hdr[0] - word_base + 8 * L + hw > avail) return;
    const uint32_t* pl = ap + ((uint64_t)hdr[0] - word_base);
    uint32_t word = lane < hw ? pl[8 * L + lane] : 0u;
    const uint32_t pc = __popc(word);
    uint32_t incl = pc;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
    }
    uint32_t r = incl - pc;
    while (word) {
        const uint32_t b = __builtin_ctz(word);
        if (r < n) hi[r] = 32 * lane + b - r;
        ++r;
        word &= word - 1;
    }
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) mark[lane * 4 + t] = 0;
    __syncthreads();
    // list starts inside the chunk: vertices from the chunk's first to the next chunk's first
    for (uint64_t v = (uint64_t)vfirst + lane; v <= vlast; v += 64) {
        const uint64_t p = beg[v] - start_edge, q = beg[v + 1] - start_edge;
        if (p >= a0 && p < a0 + kPackChunk && q > p) mark[p - a0] = 1;
    }
    // lower bits: the lane's 4L bits (up to 5 words)
    uint32_t pw[5] = {0, 0, 0, 0, 0};
    const uint32_t bit = lane * 4 * L;
    if (L) {
        const uint32_t kw = bit >> 5, nw = ((bit & 31) + 4 * L + 31) >> 5;
#pragma unroll
        for (uint32_t t = 0; t < 5; ++t)
            if (t < nw) pw[t] = pl[kw + t];
    }
    __syncthreads();
    unsigned long long s[4];
    int last = -1;   // the lane's last list start so far
    int st[4];
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
        const uint32_t j = lane * 4 + t;
        s[t] = 0;
        if (j < n) {
            s[t] = ((unsigned long long)hi[j] << L) | (L ? window_bits(pw, (bit & 31) + t * L, L) : 0u);
            sum[j] = s[t];
            if (mark[j]) last = (int)j;
        }
        st[t] = last;
    }
    int mx = last;   // inclusive over lanes 0..lane
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(mx, off);
        if (lane >= off) mx = max(mx, o);
    }
    int in = __shfl_up(mx, 1);
    if (lane == 0) in = -1;
    __syncthreads();
#pragma unroll
    for (uint32_t t = 0; t < 4; ++t) {
        const uint32_t j = lane * 4 + t;
        if (j >= n) continue;
        const int sj = st[t] >= 0 ? st[t] : in;
        const unsigned long long b = sj < 0 ? 0ull - (unsigned long long)carry : (sj == 0 ? 0ull : sum[sj - 1]);
        out[a0 + j] = (uint32_t)(s[t] - b);
    }
}

const char* pack_layout(const uint32_t* pk, uint64_t words, uint32_t nv, uint64_t ne, bool full, PackLayout* lay)
{
    const uint64_t nvc = ((uint64_t)nv + 1 + kPackChunk - 1) / kPackChunk;
    const uint64_t nch = (ne + kPackChunk - 1) / kPackChunk;
    if (!ne || words < 4 || pk[0] != kPackMagic || pk[1] != nvc) return "not a packed block of this block's size";
    lay->nvc = nvc;
    lay->vpw = pk[2];
    lay->nch = nch;
    lay->vh_off = 4;
    lay->vp_off = 4 + 4 * nvc;
    lay->ah_off = lay->vp_off + lay->vpw;
    lay->head_w = lay->ah_off + 4 * nch;
    if (lay->head_w > words) return "packed block shorter than its headers";
    if (!full) return nullptr;
    const uint32_t* vh = pk + lay->vh_off;
    uint64_t off = 0;
    for (uint64_t k = 0; k < nvc; ++k) {
        if (vh[4 * k + 2] != off) return "packed block: beg chunk offsets inconsistent";
        for (uint32_t g = 0; g < 4; ++g) {
            const uint32_t w = (vh[4 * k + 3] >> (8 * g)) & 0xFFu;
            if (w > 32) return "packed block: beg width above 32";
            off += 2 * w;
        }
    }
    if (off != lay->vpw) return "packed block: beg payload size inconsistent";
    const uint32_t* ah = pk + lay->ah_off;
    off = 0;
    for (uint64_t c = 0; c < nch; ++c) {
        const uint32_t* h = ah + 4 * c;
        const uint64_t n = std::min<uint64_t>(kPackChunk, ne - c * kPackChunk);
        if (h[0] != off || (h[1] & 0xFFu) > 31 || (h[1] >> 8) >= 2 * n || h[3] >= nv || (c && h[3] < h[-1]))
            return "packed block: arc chunk header malformed";
        off += pack_chunk_words(h, n);
    }
    if (lay->head_w + off != words) return "packed block: arc payload size inconsistent";
    return nullptr;
}

hipError_t launch_unpack_beg(const uint32_t* staging, const PackLayout& lay, uint32_t nv, uint64_t* beg, hipStream_t s)
{
    hipLaunchKernelGGL(unpack_beg_kernel, dim3((uint32_t)lay.nvc), dim3(64), 0, s, staging + lay.vh_off,
                       staging + lay.vp_off, lay.vpw, nv, beg);
    return hipGetLastError();
}

hipError_t launch_unpack_arcs(const uint32_t* head, const uint32_t* payload, const PackLayout& lay, const uint64_t* beg,
                              uint64_t start_edge, uint32_t nv, uint64_t ne, uint64_t c_begin, uint64_t c_end,
                              uint64_t word_base, uint64_t avail, uint32_t* out, hipStream_t s)
{
    if (c_end > lay.nch) c_end = lay.nch;
    if (c_begin >= c_end) return hipSuccess;
    hipLaunchKernelGGL(unpack_arcs_kernel, dim3((uint32_t)(c_end - c_begin)), dim3(64), 0, s, head + lay.ah_off,
                       payload, word_base, avail, beg, start_edge, nv, ne, lay.nch, c_begin, out);
    return hipGetLastError();
}

}  // namespace gw

using namespace gw;

namespace {

struct BitWriter {
    uint32_t* p;
    void put(uint64_t bit, uint32_t w, uint32_t v) const
    {
        if (!w) return;
        const uint64_t k = bit >> 5;
        const uint32_t s = (uint32_t)(bit & 31);
        p[k] |= v << s;
        if (s + w > 32) p[k + 1] |= v >> (32 - s);
    }
};

// Packed size of a block (host; no GPU), and the packing itself: sizes in a first pass, the
// words in a second.  beg: the block's nv + 1 global offsets; csr: its ne > 0 arcs (the
// lists clipped at ne for a truncated giant vertex).
gw_status pack_impl(const uint64_t* beg, const uint32_t* csr, uint32_t nv, uint64_t ne, uint32_t* out,
                    uint64_t cap_words, uint64_t* words)
{
    if (!beg || !csr || !words) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: null argument");
    const uint64_t se = beg[0];
    const uint64_t nent = (uint64_t)nv + 1, nvc = (nent + kPackChunk - 1) / kPackChunk;
    const uint64_t nch = (ne + kPackChunk - 1) / kPackChunk;
    if (nch > 0xFFFFFFFFull) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: block too large");
    // beg: a width per group of 64 entries
    std::vector<uint32_t> vw(nvc);
    uint64_t vpw = 0;
    for (uint64_t k = 0; k < nvc; ++k) {
        uint32_t packed_w = 0;
        for (uint32_t g = 0; g < 4; ++g) {
            uint64_t mx = 0;
            const uint64_t i0 = k * kPackChunk + 64 * g, i1 = std::min<uint64_t>(i0 + 64, nent);
            for (uint64_t i = std::max<uint64_t>(i0, k * kPackChunk + 1); i < i1; ++i) {
                if (beg[i] < beg[i - 1]) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: beg decreases");
                mx = std::max(mx, beg[i] - beg[i - 1]);
            }
            if (mx > 0xFFFFFFFFull) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: a list of 2^32 arcs or more");
            const uint32_t w = bits_for(mx);
            packed_w |= w << (8 * g);
            vpw += 2 * w;
        }
        vw[k] = packed_w;
    }
    if (vpw > 0xFFFFFFFFull) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: block too large");
    const uint64_t head = 4 + 4 * nvc + vpw + 4 * nch;
    std::vector<uint64_t> sums(kPackChunk);
    uint64_t apw = 0;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            *words = head + apw;
            if (!out) return GW_OK;
            if (cap_words < *words) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: output buffer too small");
            std::memset(out, 0, *words * 4);
            out[0] = kPackMagic;
            out[1] = (uint32_t)nvc;
            out[2] = (uint32_t)vpw;
            out[3] = 0;
            uint32_t* vh = out + 4;
            const BitWriter vb{out + 4 + 4 * nvc};
            uint64_t off = 0;
            for (uint64_t k = 0; k < nvc; ++k) {
                const uint64_t b0 = beg[k * kPackChunk];
                vh[4 * k + 0] = (uint32_t)b0;
                vh[4 * k + 1] = (uint32_t)(b0 >> 32);
                vh[4 * k + 2] = (uint32_t)off;
                vh[4 * k + 3] = vw[k];
                for (uint32_t g = 0; g < 4; ++g) {
                    const uint32_t w = (vw[k] >> (8 * g)) & 0xFFu;
                    const uint64_t i0 = k * kPackChunk + 64 * g, i1 = std::min<uint64_t>(i0 + 64, nent);
                    for (uint64_t i = i0; i < i1; ++i)
                        if (i > k * kPackChunk) vb.put(off * 32 + (i - i0) * w, w, (uint32_t)(beg[i] - beg[i - 1]));
                    off += 2 * w;
                }
            }
        }
        uint32_t* ap = out ? out + head : nullptr;
        uint64_t woff = 0, v = 0;
        for (uint64_t c = 0; c < nch; ++c) {
            const uint64_t a0 = c * kPackChunk, n = std::min<uint64_t>(kPackChunk, ne - a0);
            // vertex of arc a0: the last v with beg[v] - se <= a0 (its list holds a0)
            while (v + 1 < nv && beg[v + 1] - se <= a0) ++v;
            const uint32_t vfirst = (uint32_t)v;
            const uint32_t carry = beg[v] - se == a0 ? 0u : csr[a0 - 1];
            unsigned long long S = 0;
            uint64_t u = v;
            for (uint64_t j = 0; j < n; ++j) {
                const uint64_t a = a0 + j;
                while (u + 1 < nv && beg[u + 1] - se <= a) ++u;
                uint32_t g;
                if (beg[u] - se == a) g = csr[a];
                else if (j == 0) g = csr[a] - carry;
                else g = csr[a] - csr[a - 1];
                S += g;
                sums[j] = S;
            }
            const uint32_t L = S > n ? bits_for(S / n) - 1 : 0;
            const uint32_t nz = (uint32_t)(S >> L);
            if (pass == 1) {
                uint32_t* h = out + 4 + 4 * nvc + vpw + 4 * c;
                h[0] = (uint32_t)woff;
                h[1] = L | (nz << 8);
                h[2] = carry;
                h[3] = vfirst;
                const BitWriter lo{ap + woff}, up{ap + woff + 8ull * L};
                for (uint64_t j = 0; j < n; ++j) {
                    lo.put(j * L, L, L ? (uint32_t)(sums[j] & ((1ull << L) - 1)) : 0u);
                    up.put((sums[j] >> L) + j, 1, 1u);
                }
            }
            woff += 8ull * L + ((n + nz + 31) >> 5);
        }
        if (pass == 0) {
            apw = woff;
            if (head + apw > 0xFFFFFFFFull) return gen_fail(nullptr, GW_ERR_ARG, "gw_pack_block: block too large");
        }
    }
    return GW_OK;
}

}  // namespace

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
