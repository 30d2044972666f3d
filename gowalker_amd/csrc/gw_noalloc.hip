// gw_noalloc.hip -- gw_walk_resident_noalloc: the walk of gw_init_walkers + gw_run over one
// resident block, for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The graph, the hashes
// and the visit counts are the caller's device buffers; everything else the kernels
// touch -- the block descriptor, the cut points, the counters, the work queue and the
// walker state (cur, prev, meta: 12 B per walker) -- lives in `scratch`, a
// fixed file-scope device array.  Each GPU holds its own copy of the array from the first
// call on it until the process ends; its device address is looked up once per device
// (hipGetSymbolAddress) and passed to the kernels like any buffer.  Calls on one GPU share
// that GPU's array, so calls that overlap on one GPU race.
#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "gw_internal.h"

namespace gw {

constexpr uint64_t kNoallocHead = 1024;   // descriptor, cut points, counters, queue
constexpr uint64_t kNoallocScratch = kNoallocHead + 12ull * GW_NOALLOC_MAX_WALKERS;
static_assert(kNoallocScratch == GW_NOALLOC_SCRATCH_BYTES, "gw.h states the scratch size");
// offsets in the head
constexpr uint64_t kOffBlock = 0, kOffVcut = 64, kOffCnt = 128, kOffMisc = 256;
constexpr uint32_t kCntWords = 32;        // 16 u64 counters, cleared before every walk
constexpr uint32_t kMiscWords = 4;        // worklist length, work queue

}  // namespace gw

static __device__ __align__(256) unsigned char scratch[gw::kNoallocScratch];

namespace gw {

// Writes the single block's descriptor and its two cut points, one 32-bit word per lane
// (plain vector stores).
__global__ void noalloc_setup_kernel(uint32_t* blk_words, uint32_t* vcut, DevBlock v, uint32_t nverts)
{
    constexpr uint32_t kWords = sizeof(DevBlock) / 4;
    static_assert(sizeof(DevBlock) % 4 == 0, "descriptor copied by words");
    const uint32_t t = threadIdx.x;
    uint32_t words[kWords];
    __builtin_memcpy(words, &v, sizeof(DevBlock));
    if (t < kWords) blk_words[t] = words[t];
    if (t == kWords) vcut[0] = 0;
    if (t == kWords + 1) vcut[1] = nverts;
}

namespace {
struct DeviceScratch {
    unsigned char* ptr = nullptr;   // the device address of `scratch` on this GPU
    uint32_t ncu = 0, blocks_per_cu = 0;
};
std::mutex g_mu;
DeviceScratch g_dev[64];

}  // namespace

static gw_status nfail(gw_status st, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static gw_status nfail(gw_status st, const char* fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return gen_fail(nullptr, st, buf);
}

// Set-up on the current device, once: the device address of its copy of `scratch`, the
// CU count and the step kernel's occupancy.
static gw_status device_scratch(DeviceScratch* out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return nfail(GW_ERR_HIP, "hipGetDevice failed: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= 64) return nfail(GW_ERR_ARG, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    DeviceScratch& d = g_dev[dev];
    if (!d.ptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return nfail(GW_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                         hipGetErrorString(e));
        int ncu = 0;
        e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess || ncu <= 0) return nfail(GW_ERR_HIP, "CU count of device %d unavailable", dev);
        d.ncu = (uint32_t)ncu;
        d.blocks_per_cu = walk_blocks_per_cu(true, 1, 0, 256);
        if (!d.blocks_per_cu) d.blocks_per_cu = 1;
        d.ptr = (unsigned char*)p;
    }
    *out = d;
    return GW_OK;
}

}  // namespace gw

using namespace gw;

#define NCHK(call)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return nfail(GW_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

extern "C" gw_status gw_walk_resident_noalloc(const uint64_t* beg, const uint32_t* csr, uint32_t nverts,
                                              uint64_t nedges, const gw_walk_params* params, uint64_t nwalkers,
                                              uint32_t wpv, uint64_t* hashes, uint32_t* visits, void* stream,
                                              gw_run_stats* stats)
{
    if (!beg || !params || (nedges && !csr)) return nfail(GW_ERR_ARG, "gw_walk_resident_noalloc: null argument");
    if (nverts == 0) return nfail(GW_ERR_ARG, "gw_walk_resident_noalloc: the graph has no vertex");
    if (nwalkers == 0 || nwalkers > GW_NOALLOC_MAX_WALKERS)
        return nfail(GW_ERR_ARG,
                     "gw_walk_resident_noalloc: %llu walkers; the allocation-free walk takes 1 to %u "
                     "(its walker state lives in a fixed %llu-byte device array)",
                     (unsigned long long)nwalkers, (unsigned)GW_NOALLOC_MAX_WALKERS,
                     (unsigned long long)GW_NOALLOC_SCRATCH_BYTES);
    if (wpv == 0) return nfail(GW_ERR_ARG, "walkers_per_vertex must be >= 1");
    if ((nwalkers - 1) / wpv >= nverts)
        return nfail(GW_ERR_ARG, "%llu walkers at %u per vertex exceed %u vertices", (unsigned long long)nwalkers, wpv,
                     nverts);
    if (params->ntypes < 1 || params->ntypes > 2) return nfail(GW_ERR_ARG, "ntypes must be 1 or 2");
    DevType dt[2];
    for (uint32_t t = 0; t < params->ntypes; ++t)
        if (const char* e = dev_type_from(params->type[t], &dt[t])) return nfail(GW_ERR_ARG, "%s", e);
    if (params->ntypes == 1) dt[1] = dt[0];

    DeviceScratch d;
    if (gw_status st = device_scratch(&d)) return st;
    hipStream_t s = (hipStream_t)stream;
    unsigned char* base = d.ptr;
    DevBlock* blk = (DevBlock*)(base + kOffBlock);
    uint32_t* vcut = (uint32_t*)(base + kOffVcut);
    unsigned long long* cnt = (unsigned long long*)(base + kOffCnt);
    uint32_t* misc = (uint32_t*)(base + kOffMisc);
    uint32_t* soa = (uint32_t*)(base + kNoallocHead);
    const uint64_t cap = GW_NOALLOC_MAX_WALKERS;

    DevBlock b{};
    b.beg = beg;
    b.csr = csr;
    b.start_edge = 0;
    b.nedges = nedges;
    b.start_vert = 0;
    b.nverts = nverts;
    b.resident = kInSlot;
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3(1), dim3(64), 0, s, (uint32_t*)blk, vcut, b, nverts);
    NCHK(hipGetLastError());
    // the counters and the queue start at zero on every walk; so do the visit counts
    NCHK(launch_fill_u32((uint32_t*)cnt, 0, kCntWords, s));
    NCHK(launch_fill_u32(misc, 0, kMiscWords, s));
    if (visits) NCHK(launch_fill_u32(visits, 0, nverts, s));

    InitArgs ia{};
    ia.ntypes = params->ntypes;
    ia.len0 = dt[0].length;
    ia.len1 = dt[1].length;
    ia.nwalkers = nwalkers;
    ia.wpv = wpv;
    ia.cur = soa;
    ia.prev = soa + cap;
    ia.meta = soa + 2 * cap;
    ia.hash = hashes;
    NCHK(launch_init(ia, s));

    WalkArgs a{};
    a.blocks = blk;
    a.vcut = vcut;
    a.nblocks = 1;
    a.ntypes = params->ntypes;
    a.seed_lo = (uint32_t)params->seed;
    a.seed_hi = (uint32_t)(params->seed >> 32);
    a.type0 = dt[0];
    a.type1 = dt[1];
    a.cur = ia.cur;
    a.prev = ia.prev;
    a.meta = ia.meta;
    a.hash = hashes;
    a.visits = visits;
    a.visits_stride = nverts;
    a.visits_parts = 1;
    a.nwork = (uint32_t)nwalkers;
    a.queue = misc + 1;
    a.counters = cnt;
    // variant 0, bisection, 256 threads: the persistent grid as run_impl sizes it
    const uint32_t block = 256;
    uint64_t grid = (uint64_t)d.ncu * d.blocks_per_cu;
    const uint64_t need = (nwalkers + block - 1) / block;
    if (grid > need) grid = need;
    if (grid == 0) grid = 1;
    uint64_t ch = nwalkers / (grid * (block / 64) * 8 + 1);
    ch = ch < 64 ? 64 : (ch > 1024 ? 1024 : ch);
    a.chunk = (uint32_t)(ch / 64 * 64);
    NCHK(launch_walk(a, true, 0, (uint32_t)grid, block, s));
    if (!stats) return GW_OK;

    unsigned long long h[8] = {0};
    NCHK(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, s));
    NCHK(hipStreamSynchronize(s));
    gw_run_stats r{};
    r.steps = h[0];
    r.trials = h[1];
    r.probes = h[2];
    r.mix_steps = h[3];
    r.walkers_ended = h[4];
    r.walkers_run = h[4] + h[5];
    r.walkers_left = nwalkers - (h[4] < nwalkers ? h[4] : nwalkers);
    *stats = r;
    return GW_OK;
}
