// gw_gen.hip -- synthetic graph generators on the GPU, writing GOWalker's CSR layout
// (u64 offsets indexed by vertex id, u32 sorted adjacency; preprocess/graph_converter.hpp
// :186-284 semantics) straight into HBM.
//
// * RMAT (a, b, c, d): edge e descends `scale` quadrant levels; level l uses word l % 4
//   of philox(seed, (e_lo, e_hi, l / 4, 7)) against thresholds floor(a 2^32),
//   floor((a+b) 2^32), floor((a+b+c) 2^32).  Optional random id permutation (a keyed
//   Feistel bijection), symmetrisation, self-loop removal and deduplication.  Keys are
//   bucketed by source range so that each radix sort stays under 2^29 keys.
// * Chung-Lu (directed, power law): vertex i has out-degree
//   floor(E F(i+1)) - floor(E F(i)), F the normalised cumulative weight of
//   w_i = (i + i0)^(-1/(gamma-1)); targets follow the same weights, drawn by stratified
//   inverse-CDF sampling so that every adjacency list comes out sorted (multi-edges kept).
#include "gw_device.h"
#include "gw.h"

#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

namespace gw {

struct RmatParams {
    uint64_t nedges;
    uint32_t scale, perm_bits, undirected, permute;
    uint32_t ta, tab, tabc;
    uint32_t k0, k1;       // Philox key
    uint32_t fk0, fk1;     // Feistel key
    uint32_t bucket_shift; // src >> bucket_shift = histogram bucket
};

__device__ __forceinline__ void rmat_edge(const RmatParams& p, uint64_t e, uint32_t& s, uint32_t& d)
{
    s = 0;
    d = 0;
    for (uint32_t l = 0; l < p.scale; l += 4) {
        const uint4 r = philox((uint32_t)e, (uint32_t)(e >> 32), l >> 2, 7u, p.k0, p.k1);
        const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            if (l + j < p.scale) {
                const uint32_t x = rr[j];
                const uint32_t row = x >= p.tab;
                const uint32_t col = (x >= p.ta && x < p.tab) || x >= p.tabc;
                s = (s << 1) | row;
                d = (d << 1) | col;
            }
        }
    }
}

__device__ __forceinline__ uint32_t feistel_round(uint32_t r, uint32_t k, uint32_t round)
{
    uint32_t h = r * 0x9E3779B1u + k + round * 0x85EBCA77u;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return h;
}

// keyed bijection of [0, 2^scale): 4-round balanced Feistel over perm_bits (even) bits,
// cycle-walked back into range when perm_bits > scale
__device__ __forceinline__ uint32_t permute_id(const RmatParams& p, uint32_t x)
{
    const uint32_t h = p.perm_bits / 2, mask = (1u << h) - 1u;
    do {
        uint32_t L = x >> h, R = x & mask;
        for (uint32_t r = 0; r < 4; ++r) {
            const uint32_t t = L ^ (feistel_round(R, (r & 1) ? p.fk1 : p.fk0, r) & mask);
            L = R;
            R = t;
        }
        x = (L << h) | R;
    } while (x >> p.scale);
    return x;
}

__device__ __forceinline__ void rmat_final(const RmatParams& p, uint64_t e, uint32_t& s, uint32_t& d)
{
    rmat_edge(p, e, s, d);
    if (p.permute) {
        s = permute_id(p, s);
        d = permute_id(p, d);
    }
}

// histogram of emitted keys per source bucket
__global__ void __launch_bounds__(256) rmat_hist_kernel(RmatParams p, unsigned long long* hist)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < p.nedges; e += stride) {
        uint32_t s, d;
        rmat_final(p, e, s, d);
        if (s == d) continue;
        atomicAdd(hist + (s >> p.bucket_shift), 1ull);
        if (p.undirected) atomicAdd(hist + (d >> p.bucket_shift), 1ull);
    }
}

// keys (s << scale | d) with s in [lo, hi)
__global__ void __launch_bounds__(256) rmat_fill_kernel(RmatParams p, uint32_t lo, uint32_t hi, uint64_t* keys,
                                                         unsigned long long* n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t lane = __lane_id();
    const uint64_t npad = (p.nedges + 63) & ~63ull;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < npad; e += stride) {
        uint32_t s = 0, d = 0;
        bool k1 = false, k2 = false;
        if (e < p.nedges) {
            rmat_final(p, e, s, d);
            if (s != d) {
                k1 = s >= lo && s < hi;
                k2 = p.undirected && d >= lo && d < hi;
            }
        }
        const uint64_t m1 = __ballot(k1), m2 = __ballot(k2);
        const uint32_t c1 = __popcll(m1), c = c1 + __popcll(m2);
        if (c) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(n, (unsigned long long)c);
            base = __shfl(base, 0, 64);
            if (k1) keys[base + lane_rank(m1)] = ((uint64_t)s << p.scale) | d;
            if (k2) keys[base + c1 + lane_rank(m2)] = ((uint64_t)d << p.scale) | s;
        }
    }
}

__global__ void __launch_bounds__(256) keys_to_csr_kernel(const uint64_t* keys, uint64_t n, uint32_t scale,
                                                           uint32_t* csr, uint32_t* deg)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t mask = (1ull << scale) - 1;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t k = keys[i];
        csr[i] = (uint32_t)(k & mask);
        atomicAdd(deg + (k >> scale), 1u);
    }
}

// ---------------- Chung-Lu ----------------
struct ClParams {
    uint64_t n, nedges;
    double beta1;      // 1 - beta, beta = 1/(gamma-1)
    double i0, c0;     // offset and i0^(1-beta)
    double cn;         // C(n)
    uint32_t k0, k1;
};

__device__ __forceinline__ double cl_cum(const ClParams& p, double x)
{
    return (pow(x + p.i0, p.beta1) - p.c0) / p.beta1;
}

__global__ void __launch_bounds__(256) cl_degree_kernel(ClParams p, uint32_t* deg)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n; i += stride) {
        const double e = (double)p.nedges / p.cn;
        const double a = i == 0 ? 0.0 : floor(e * cl_cum(p, (double)i));
        double b = floor(e * cl_cum(p, (double)(i + 1)));
        if (i + 1 == p.n) b = (double)p.nedges;
        deg[i] = (uint32_t)(b - a);
    }
}

// one wavefront per vertex; lanes stride over its out-edges
__global__ void __launch_bounds__(256) cl_target_kernel(ClParams p, const uint64_t* beg, uint32_t* csr)
{
    const uint32_t lane = __lane_id();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const double inv = 1.0 / p.beta1;
    for (uint64_t i = wave; i < p.n; i += nwaves) {
        const uint64_t h = beg[i], d = beg[i + 1] - h;
        for (uint64_t k = lane; k < d; k += 64) {
            const uint4 r = philox((uint32_t)i, (uint32_t)k, (uint32_t)(k >> 32), 11u, p.k0, p.k1);
            const double u = ((double)k + (double)r.x * 2.3283064365386963e-10) / (double)d;
            double x = pow(u * p.cn * p.beta1 + p.c0, inv) - p.i0;
            uint64_t j = x <= 0.0 ? 0 : (uint64_t)x;
            if (j >= p.n) j = p.n - 1;
            csr[h + k] = (uint32_t)j;
        }
    }
}

// floating-point rounding can invert neighbouring targets in rare cases: restore order
__global__ void __launch_bounds__(256) cl_sortfix_kernel(uint64_t n, const uint64_t* beg, uint32_t* csr,
                                                          unsigned long long* fixed)
{
    const uint32_t lane = __lane_id();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t i = wave; i < n; i += nwaves) {
        const uint64_t h = beg[i], d = beg[i + 1] - h;
        bool bad = false;
        for (uint64_t k = lane + 1; k < d; k += 64) bad |= csr[h + k - 1] > csr[h + k];
        if (__any(bad) && lane == 0) {
            for (uint64_t k = 1; k < d; ++k) {   // insertion sort: the list is nearly sorted
                const uint32_t v = csr[h + k];
                uint64_t j = k;
                while (j > 0 && csr[h + j - 1] > v) { csr[h + j] = csr[h + j - 1]; --j; }
                csr[h + j] = v;
            }
            atomicAdd(fixed, 1ull);
        }
    }
}

// duplicate / self-loop removal over sorted lists, one wavefront per vertex
__device__ __forceinline__ bool cl_keep(const uint32_t* csr, uint64_t h, uint64_t k, uint64_t v)
{
    const uint32_t x = csr[h + k];
    return x != (uint32_t)v && (k == 0 || x != csr[h + k - 1]);
}

__global__ void __launch_bounds__(256) cl_uniq_count_kernel(uint64_t n, const uint64_t* beg, const uint32_t* csr,
                                                             uint32_t* deg)
{
    const uint32_t lane = __lane_id();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t i = wave; i < n; i += nwaves) {
        const uint64_t h = beg[i], d = beg[i + 1] - h;
        uint64_t cnt = 0;
        for (uint64_t k = lane; k < d; k += 64) cnt += cl_keep(csr, h, k, i);
        cnt = wave_sum(cnt);
        if (lane == 0) deg[i] = (uint32_t)cnt;
    }
}

__global__ void __launch_bounds__(256) cl_uniq_write_kernel(uint64_t n, const uint64_t* beg, const uint32_t* csr,
                                                             const uint64_t* beg2, uint32_t* csr2)
{
    const uint32_t lane = __lane_id();
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    for (uint64_t i = wave; i < n; i += nwaves) {
        const uint64_t h = beg[i], d = beg[i + 1] - h;
        uint64_t out = beg2[i];
        for (uint64_t k0 = 0; k0 < d; k0 += 64) {
            const uint64_t k = k0 + lane;
            const bool keep = k < d && cl_keep(csr, h, k, i);
            const uint64_t m = __ballot(keep);
            if (keep) csr2[out + lane_rank(m)] = csr[h + k];
            out += __popcll(m);
        }
    }
}

__global__ void u32_to_u64_kernel(const uint32_t* in, uint64_t* out, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = in[i];
}

}  // namespace gw

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------
#include <cmath>
#include <cstdio>
#include <vector>

struct gw_graph {
    uint64_t nverts = 0, nedges = 0;
    uint64_t* beg = nullptr;   // device, nverts+1
    uint32_t* csr = nullptr;   // device, nedges (+16 B pad)
};

namespace gw {
gw_status gen_fail(gw_ctx* ctx, gw_status st, const char* msg);   // gw_api.cpp
hipStream_t ctx_stream(gw_ctx* ctx);
}

using namespace gw;

extern "C" gw_status gw_graph_free(gw_ctx* ctx, gw_graph* g);

#define GCHK(call)                                                           \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) {                                              \
            char b_[256];                                                    \
            snprintf(b_, sizeof b_, "%s: %s", #call, hipGetErrorString(e_)); \
            st = gen_fail(ctx, GW_ERR_HIP, b_);                              \
            goto done;                                                       \
        }                                                                    \
    } while (0)

static uint32_t grid_for(uint64_t n, uint32_t block = 256, uint32_t cap = 65536)
{
    uint64_t g = (n + block - 1) / block;
    return (uint32_t)(g == 0 ? 1 : (g > cap ? cap : g));
}

// exclusive scan of n u32 degrees into n+1 u64 offsets
static hipError_t degrees_to_offsets(const uint32_t* deg, uint64_t* beg, uint64_t n, hipStream_t s)
{
    hipError_t e;
    uint64_t* tmp64 = nullptr;
    void* tmp = nullptr;
    size_t bytes = 0;
    if ((e = hipMalloc(&tmp64, (n + 1) * 8)) != hipSuccess) return e;
    hipLaunchKernelGGL(u32_to_u64_kernel, dim3(grid_for(n)), dim3(256), 0, s, deg, tmp64, n);
    if ((e = hipMemsetAsync(tmp64 + n, 0, 8, s)) != hipSuccess) goto out;
    if ((e = rocprim::exclusive_scan(nullptr, bytes, tmp64, beg, (uint64_t)0, (size_t)(n + 1),
                                     rocprim::plus<uint64_t>(), s)) != hipSuccess) goto out;
    if ((e = hipMalloc(&tmp, bytes)) != hipSuccess) goto out;
    e = rocprim::exclusive_scan(tmp, bytes, tmp64, beg, (uint64_t)0, (size_t)(n + 1), rocprim::plus<uint64_t>(), s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
out:
    if (tmp) (void)hipFree(tmp);
    (void)hipFree(tmp64);
    return e;
}

extern "C" gw_status gw_gen_rmat(gw_ctx* ctx, uint32_t scale, uint64_t nedges, double a, double b, double c,
                                 uint64_t seed, uint32_t flags, gw_graph** out)
{
    gw_status st = GW_OK;
    if (!ctx || !out) return gen_fail(ctx, GW_ERR_ARG, "null argument");
    *out = nullptr;
    if (scale < 1 || scale > 31) return gen_fail(ctx, GW_ERR_ARG, "scale must be in [1, 31]");
    if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return gen_fail(ctx, GW_ERR_ARG, "bad RMAT probabilities");
    hipStream_t s = ctx_stream(ctx);
    const uint64_t n = 1ull << scale;
    const bool und = flags & GW_GEN_UNDIRECTED;
    RmatParams p{};
    p.nedges = nedges;
    p.scale = scale;
    p.perm_bits = scale + (scale & 1);
    p.undirected = und;
    p.permute = (flags & GW_GEN_PERMUTE) != 0;
    auto th = [](double x) -> uint32_t { return x >= 1.0 ? 0xFFFFFFFFu : (uint32_t)std::floor(x * 4294967296.0); };
    p.ta = th(a);
    p.tab = th(a + b);
    p.tabc = th(a + b + c);
    p.k0 = (uint32_t)seed;
    p.k1 = (uint32_t)(seed >> 32);
    p.fk0 = (uint32_t)(seed * 0x9E3779B97F4A7C15ull >> 32);
    p.fk1 = (uint32_t)(seed * 0xC2B2AE3D27D4EB4Full >> 32);
    const uint32_t hb = scale > 12 ? 12 : scale;
    p.bucket_shift = scale - hb;
    const uint64_t nb = 1ull << hb;
    const uint64_t part_cap = 1ull << 29;

    gw_graph* g = new gw_graph();
    unsigned long long* d_hist = nullptr;
    unsigned long long* d_n = nullptr;
    uint32_t* deg = nullptr;
    uint64_t *keys = nullptr, *keys2 = nullptr, *ukeys = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    std::vector<unsigned long long> hist(nb);
    uint64_t total = 0, written = 0;

    GCHK(hipMalloc(&d_hist, nb * 8));
    GCHK(hipMalloc(&d_n, 16));
    GCHK(hipMemsetAsync(d_hist, 0, nb * 8, s));
    hipLaunchKernelGGL(rmat_hist_kernel, dim3(grid_for(nedges)), dim3(256), 0, s, p, d_hist);
    GCHK(hipGetLastError());
    GCHK(hipMemcpyAsync(hist.data(), d_hist, nb * 8, hipMemcpyDeviceToHost, s));
    GCHK(hipStreamSynchronize(s));
    for (auto v : hist) total += v;
    if (total >= (1ull << 32) * 4) { st = gen_fail(ctx, GW_ERR_ARG, "too many edges"); goto done; }
    for (auto v : hist)
        if (v > part_cap) { st = gen_fail(ctx, GW_ERR_ARG, "a source bucket exceeds the sort chunk"); goto done; }
    GCHK(hipMalloc(&g->csr, total * 4 + 64));
    GCHK(hipMalloc(&deg, n * 4));
    GCHK(hipMemsetAsync(deg, 0, n * 4, s));
    {
        uint64_t maxpart = 0;
        // group buckets into parts of at most part_cap keys
        std::vector<std::pair<uint64_t, uint64_t>> parts;   // bucket ranges [lo, hi)
        std::vector<uint64_t> pk;
        uint64_t lo = 0, acc = 0;
        for (uint64_t i = 0; i < nb; ++i) {
            if (acc + hist[i] > part_cap && acc > 0) {
                parts.push_back({lo, i});
                pk.push_back(acc);
                lo = i;
                acc = 0;
            }
            acc += hist[i];
        }
        parts.push_back({lo, nb});
        pk.push_back(acc);
        for (auto v : pk) maxpart = v > maxpart ? v : maxpart;
        GCHK(hipMalloc(&keys, maxpart * 8 + 8));
        GCHK(hipMalloc(&keys2, maxpart * 8 + 8));
        ukeys = keys;   // unique output reuses the first buffer
        {
            size_t b1 = 0, b2 = 0;
            GCHK(rocprim::radix_sort_keys(nullptr, b1, keys, keys2, (size_t)maxpart, 0, (int)(2 * scale), s));
            GCHK(rocprim::unique(nullptr, b2, keys2, ukeys, d_n + 1, (size_t)maxpart, rocprim::equal_to<uint64_t>(), s));
            tmp_bytes = b1 > b2 ? b1 : b2;
            GCHK(hipMalloc(&tmp, tmp_bytes));
        }
        for (size_t q = 0; q < parts.size(); ++q) {
            const uint64_t cnt = pk[q];
            if (cnt == 0) continue;
            const uint32_t slo = (uint32_t)(parts[q].first << p.bucket_shift);
            const uint64_t shi64 = parts[q].second << p.bucket_shift;
            const uint32_t shi = shi64 > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)shi64;
            GCHK(hipMemsetAsync(d_n, 0, 16, s));
            hipLaunchKernelGGL(rmat_fill_kernel, dim3(grid_for(nedges)), dim3(256), 0, s, p, slo, shi, keys, d_n);
            GCHK(hipGetLastError());
            size_t tb = tmp_bytes;
            GCHK(rocprim::radix_sort_keys(tmp, tb, keys, keys2, (size_t)cnt, 0, (int)(2 * scale), s));
            uint64_t nu = cnt;
            if (flags & GW_GEN_DEDUP) {
                tb = tmp_bytes;
                GCHK(rocprim::unique(tmp, tb, keys2, ukeys, d_n + 1, (size_t)cnt, rocprim::equal_to<uint64_t>(), s));
                unsigned long long hn = 0;
                GCHK(hipMemcpyAsync(&hn, d_n + 1, 8, hipMemcpyDeviceToHost, s));
                GCHK(hipStreamSynchronize(s));
                nu = hn;
                hipLaunchKernelGGL(keys_to_csr_kernel, dim3(grid_for(nu)), dim3(256), 0, s, ukeys, nu, scale,
                                   g->csr + written, deg);
            } else {
                hipLaunchKernelGGL(keys_to_csr_kernel, dim3(grid_for(nu)), dim3(256), 0, s, keys2, nu, scale,
                                   g->csr + written, deg);
            }
            GCHK(hipGetLastError());
            written += nu;
        }
    }
    GCHK(hipMalloc(&g->beg, (n + 1) * 8));
    GCHK(degrees_to_offsets(deg, g->beg, n, s));
    g->nverts = n;
    g->nedges = written;
    *out = g;
    g = nullptr;
done:
    if (tmp) (void)hipFree(tmp);
    if (keys) (void)hipFree(keys);
    if (keys2) (void)hipFree(keys2);
    if (deg) (void)hipFree(deg);
    if (d_hist) (void)hipFree(d_hist);
    if (d_n) (void)hipFree(d_n);
    if (g) {
        if (g->beg) (void)hipFree(g->beg);
        if (g->csr) (void)hipFree(g->csr);
        delete g;
    }
    return st;
}

static gw_status chung_lu_once(gw_ctx* ctx, uint64_t nverts, uint64_t nedges, double gamma, uint64_t max_degree,
                               uint64_t seed, gw_graph** out);

// With GW_GEN_DEDUP, duplicate targets and self-loops are removed; the edge count drawn
// is raised once by the observed duplicate rate so that about `nedges` arcs remain.
extern "C" gw_status gw_gen_chung_lu(gw_ctx* ctx, uint64_t nverts, uint64_t nedges, double gamma,
                                     uint64_t max_degree, uint64_t seed, uint32_t flags, gw_graph** out)
{
    if (!(flags & GW_GEN_DEDUP)) return chung_lu_once(ctx, nverts, nedges, gamma, max_degree, seed, out);
    if (!ctx || !out) return gen_fail(ctx, GW_ERR_ARG, "null argument");
    gw_status st = GW_OK;
    hipStream_t s = ctx_stream(ctx);
    uint64_t draw = nedges;
    gw_graph* g = nullptr;
    gw_graph* u = nullptr;
    uint32_t* deg = nullptr;
    for (int pass = 0; pass < 2; ++pass) {
        if ((st = chung_lu_once(ctx, nverts, draw, gamma, max_degree, seed, &g)) != GW_OK) goto done;
        GCHK(hipMalloc(&deg, nverts * 4));
        hipLaunchKernelGGL(cl_uniq_count_kernel, dim3(grid_for(nverts * 64, 256, 1 << 20)), dim3(256), 0, s, nverts,
                           g->beg, g->csr, deg);
        GCHK(hipGetLastError());
        u = new gw_graph();
        GCHK(hipMalloc(&u->beg, (nverts + 1) * 8));
        GCHK(degrees_to_offsets(deg, u->beg, nverts, s));
        GCHK(hipMemcpy(&u->nedges, u->beg + nverts, 8, hipMemcpyDeviceToHost));
        u->nverts = nverts;
        GCHK(hipMalloc(&u->csr, u->nedges * 4 + 64));
        hipLaunchKernelGGL(cl_uniq_write_kernel, dim3(grid_for(nverts * 64, 256, 1 << 20)), dim3(256), 0, s, nverts,
                           g->beg, g->csr, u->beg, u->csr);
        GCHK(hipGetLastError());
        GCHK(hipStreamSynchronize(s));
        (void)hipFree(deg);
        deg = nullptr;
        const uint64_t kept = u->nedges;
        gw_graph_free(ctx, g);
        g = nullptr;
        if (pass == 1 || kept == 0 || kept >= nedges) break;
        const double scale = (double)nedges / (double)kept;
        draw = (uint64_t)((double)draw * scale);
        gw_graph_free(ctx, u);
        u = nullptr;
    }
    *out = u;
    u = nullptr;
done:
    if (deg) (void)hipFree(deg);
    if (g) gw_graph_free(ctx, g);
    if (u) gw_graph_free(ctx, u);
    return st;
}

static gw_status chung_lu_once(gw_ctx* ctx, uint64_t nverts, uint64_t nedges, double gamma, uint64_t max_degree,
                               uint64_t seed, gw_graph** out)
{
    gw_status st = GW_OK;
    if (!ctx || !out) return gen_fail(ctx, GW_ERR_ARG, "null argument");
    *out = nullptr;
    if (nverts < 2 || nverts > 0xFFFFFFFFull || nedges == 0 || !(gamma > 2.0) || max_degree == 0 ||
        max_degree > nedges)
        return gen_fail(ctx, GW_ERR_ARG, "bad Chung-Lu parameters");
    hipStream_t s = ctx_stream(ctx);
    ClParams p{};
    p.n = nverts;
    p.nedges = nedges;
    const double beta = 1.0 / (gamma - 1.0);
    p.beta1 = 1.0 - beta;
    p.k0 = (uint32_t)seed;
    p.k1 = (uint32_t)(seed >> 32);
    // choose the offset i0 so that vertex 0's out-degree E*C(1)/C(n) matches max_degree
    auto d0 = [&](double i0) {
        const double c0 = std::pow(i0, p.beta1);
        const double c1 = (std::pow(1.0 + i0, p.beta1) - c0) / p.beta1;
        const double cn = (std::pow((double)nverts + i0, p.beta1) - c0) / p.beta1;
        return (double)nedges * c1 / cn;
    };
    double lo = 1e-6, hi = 1e12;
    for (int it = 0; it < 200; ++it) {
        const double mid = std::sqrt(lo * hi);
        if (d0(mid) > (double)max_degree) lo = mid; else hi = mid;
    }
    p.i0 = hi;
    p.c0 = std::pow(p.i0, p.beta1);
    p.cn = (std::pow((double)nverts + p.i0, p.beta1) - p.c0) / p.beta1;

    gw_graph* g = new gw_graph();
    uint32_t* deg = nullptr;
    unsigned long long* d_fixed = nullptr;
    GCHK(hipMalloc(&deg, nverts * 4));
    hipLaunchKernelGGL(cl_degree_kernel, dim3(grid_for(nverts)), dim3(256), 0, s, p, deg);
    GCHK(hipGetLastError());
    GCHK(hipMalloc(&g->beg, (nverts + 1) * 8));
    GCHK(degrees_to_offsets(deg, g->beg, nverts, s));
    GCHK(hipMalloc(&g->csr, nedges * 4 + 64));
    hipLaunchKernelGGL(cl_target_kernel, dim3(grid_for(nverts * 64, 256, 1 << 20)), dim3(256), 0, s, p, g->beg, g->csr);
    GCHK(hipGetLastError());
    GCHK(hipMalloc(&d_fixed, 8));
    GCHK(hipMemsetAsync(d_fixed, 0, 8, s));
    hipLaunchKernelGGL(cl_sortfix_kernel, dim3(grid_for(nverts * 64, 256, 1 << 20)), dim3(256), 0, s, nverts, g->beg,
                       g->csr, d_fixed);
    GCHK(hipGetLastError());
    GCHK(hipStreamSynchronize(s));
    g->nverts = nverts;
    g->nedges = nedges;
    *out = g;
    g = nullptr;
done:
    if (deg) (void)hipFree(deg);
    if (d_fixed) (void)hipFree(d_fixed);
    if (g) {
        if (g->beg) (void)hipFree(g->beg);
        if (g->csr) (void)hipFree(g->csr);
        delete g;
    }
    return st;
}

extern "C" gw_status gw_graph_size(const gw_graph* g, uint64_t* nverts, uint64_t* nedges)
{
    if (!g || !nverts || !nedges) return gen_fail(nullptr, GW_ERR_ARG, "null argument");
    *nverts = g->nverts;
    *nedges = g->nedges;
    return GW_OK;
}

extern "C" gw_status gw_graph_read(gw_ctx* ctx, const gw_graph* g, uint64_t* beg, uint32_t* csr)
{
    gw_status st = GW_OK;
    if (!ctx || !g || !beg || !csr) return gen_fail(ctx, GW_ERR_ARG, "null argument");
    hipStream_t s = ctx_stream(ctx);
    GCHK(hipMemcpyAsync(beg, g->beg, (g->nverts + 1) * 8, hipMemcpyDeviceToHost, s));
    GCHK(hipMemcpyAsync(csr, g->csr, g->nedges * 4, hipMemcpyDeviceToHost, s));
    GCHK(hipStreamSynchronize(s));
done:
    return st;
}

extern "C" gw_status gw_graph_free(gw_ctx* ctx, gw_graph* g)
{
    (void)ctx;
    if (!g) return GW_OK;
    if (g->beg) (void)hipFree(g->beg);
    if (g->csr) (void)hipFree(g->csr);
    delete g;
    return GW_OK;
}
