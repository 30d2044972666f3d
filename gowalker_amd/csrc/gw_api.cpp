// gw_api.cpp -- the C ABI of libgw.so (declared in include/gw.h): device context,
// block table, HBM slot cache, walker state and the run loop over the gfx950 kernels.
//
// Replaces gpu/memory.hpp (initwalk, initgpugraph, initgpucache, cudarand,
// copycachetogpu, pipeline_gpugraph) and the GPU half of graph_engine::pipeline
// (engine/engine.hpp:406-457).  No call exits the process; errors are returned.
#include "gw.h"
#include "gw_internal.h"

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace gw;

// Everything a run's captured graph depends on (arguments zeroed bytewise before they are
// filled in): equal signatures launch the same kernels with the same arguments.
struct RunSig {
    WalkArgs a;
    FilterArgs f;
    uint64_t grid;
    uint32_t block, variant, flags, ncu;
    const void *d_bucket, *m_cnt, *m_trace, *m_bucket, *m_blocks;
};

// One block copy (and decode) to put on the copy streams.
struct CopyJob {
    uint32_t slot = 0, blk = 0, lane = 0;
    const uint64_t* beg = nullptr;       // raw upload
    const uint32_t* csr = nullptr;
    const uint32_t* packed = nullptr;    // packed upload: its words
    uint64_t words = 0;
    PackLayout lay{};
    bool ring = false;                   // through the staging ring and s_dec
    hipEvent_t after_compute = nullptr;  // compute-stream work that may still read the slot
    uint64_t seq = 0;
};

struct gw_ctx {
    int device = 0;
    hipStream_t s_compute = nullptr, s_copy = nullptr, s_read = nullptr;
    // copy lanes: block copies go on s_copy, or alternate between s_copy and s_copy2, each
    // lane with its own part of the packed staging buffer, so a decode on one lane does not
    // hold up the link while the other copies
    hipStream_t s_copy2 = nullptr;
    uint32_t lanes = 1, next_lane = 0;
    hipEvent_t ev_f0 = nullptr, ev_f1 = nullptr, ev_k0 = nullptr, ev_k1 = nullptr;
    hipEvent_t ev_c1 = nullptr;
    uint32_t ncu = 0;
    gw_device_info info{};
    uint64_t cap = 0, held = 0;
    std::string err;

    // block table
    uint32_t nblocks = 0;
    uint64_t nverts_total = 0;
    std::vector<gw_block_desc> desc;
    std::vector<DevBlock> hblk;      // host mirror of the device table
    std::vector<int32_t> blk_slot;   // slot holding each block, -1 if none
    std::vector<const uint64_t*> map_beg;
    std::vector<const uint32_t*> map_csr;
    DevBlock* d_blocks = nullptr;
    DevBlock* h_blocks_pinned = nullptr;
    uint32_t* d_vcut = nullptr;
    std::vector<uint32_t> vcut;

    // staging buffer of packed block copies
    uint32_t* d_pack = nullptr;
    uint64_t pack_cap = 0;
    // With one copy lane, packed blocks are decoded on their own stream: s_copy only copies,
    // a block's headers into one of two head areas at the front of the staging buffer (the
    // blocks alternate) and its arc payload in pieces into a ring over the rest, and s_dec
    // decodes each piece into the slot once its copy has landed.  The link then carries the
    // next piece while a piece decodes, and the copy stream never waits for another stream:
    // the decoder, which writes the slot, waits for the compute stream, and the host enqueues
    // a copy once the decodes that still read the head area or ring region it overwrites are
    // done.  GW_PACK_INLINE_DECODE=1 decodes on the copy stream instead (the two-lane path,
    // and blocks whose headers do not fit).
    hipStream_t s_dec = nullptr;
    bool inline_decode = false;
    bool no_inplace = false;        // GW_PACK_INPLACE=0: every payload through the ring
    uint64_t inplace_blocks = 0, ring_blocks = 0, inline_blocks = 0;   // packed blocks by decode path
    struct PkRegion {
        uint64_t a, b;        // words [a, b) of the ring
        hipEvent_t freed;     // recorded on s_dec after the decode that reads them
    };
    std::vector<PkRegion> pk_live;
    uint64_t pk_wp = 0;                        // next ring position (words)
    uint32_t pk_head_next = 0;                 // head area of the next block
    hipEvent_t pk_head_ev[2] = {nullptr, nullptr};   // last decode that read each head area
    std::vector<hipEvent_t> pk_pool, pk_retired;     // free events; events a stream may still wait on

    // Upload worker (one copy lane): a thread of the context puts block copies and decodes
    // on the streams, in the order they were asked for, so that a host wait inside the HIP
    // runtime behind a long copy queue holds up neither scheduling nor kernel launches.  The
    // caller's thread validates and books each upload; anything that needs a slot's copy
    // events first waits until the worker has enqueued that slot's last job.
    // GW_UPLOAD_THREAD=0 enqueues on the caller's thread instead.
    bool sync_uploads = false;
    std::thread worker;
    std::mutex wmu;
    std::condition_variable wcv;
    std::deque<CopyJob> jobs;
    uint64_t job_seq = 0, job_done = 0;   // jobs submitted; jobs enqueued (in order)
    bool wstop = false;
    gw_status werr = GW_OK;               // the worker's first failure
    std::string werr_msg;
    std::vector<uint64_t> slot_seq;       // last job of each slot
    std::vector<hipEvent_t> cev_pool;     // events for jobs' compute-stream waits
    // per block: the packed buffer (address, size, fingerprint) whose headers were last
    // checked in full
    struct PackSeen {
        const void* p;
        uint64_t bytes, fp;
        bool operator==(const PackSeen& o) const { return p == o.p && bytes == o.bytes && fp == o.fp; }
    };
    std::vector<PackSeen> pack_checked;

    // slots
    uint32_t nslots = 0;
    uint64_t slot_beg_bytes = 0, slot_csr_bytes = 0;
    std::vector<uint64_t*> slot_beg;
    std::vector<uint32_t*> slot_csr;
    std::vector<int32_t> slot_blk;
    std::vector<int32_t> slot_staged;     // block copied by gw_stage_block, not yet committed
    std::vector<hipEvent_t> slot_ev0, slot_ev1;   // start / end of the slot's last copy
    std::vector<uint8_t> slot_wait;       // a run must wait for the slot's copy
    std::vector<uint8_t> slot_timed;      // copy time not yet added to tot.h2d_ms
    std::vector<uint8_t> slot_lane;       // lane of the slot's last copy
    std::vector<uint8_t> slot_read;       // a block in the slot has been readable by kernels

    // walk configuration
    bool walk_set = false;
    gw_walk_params wp{};
    DevType dt[2]{};

    // walkers
    uint64_t nwalkers = 0;
    uint32_t outputs = 0;
    uint32_t path_stride = 0;
    uint32_t *d_cur = nullptr, *d_prev = nullptr, *d_meta = nullptr;
    uint64_t* d_hash = nullptr;
    uint32_t* d_paths = nullptr;
    uint32_t* d_visits = nullptr;
    uint32_t visit_parts = 1;              // visit-count arrays (summed by gw_read_visits)
    uint32_t* d_work = nullptr;            // worklist of runnable walkers: work_cap entries
    uint64_t work_cap = 0, work_cap_req = 0, work_rounds = 0;
    // blocks resident in some run since gw_init_walkers: a walker's previous block was resident
    // when it stepped from it, so only a block of this set that is gone can be one that is not
    std::vector<uint8_t> ever_res;
    uint32_t* d_bucket = nullptr;
    uint32_t* d_bpart = nullptr;           // per-workgroup rows of the bucket pass (one per CU)
    unsigned long long* d_trace = nullptr; // per-block steps of the last traced run
    float* d_feat = nullptr;               // nblocks x kNumFeatures
    float* d_age = nullptr;
    uint32_t* d_misc = nullptr;            // [0] nwork, [1] queue
    // [0..5] run counters, [6] left, [7] zero-copy reads, [8] hops, [10..13] walkers and hops
    // the run's list took out of the bucket matrix and put back
    unsigned long long* d_cnt = nullptr;
    unsigned long long* h_cnt = nullptr;   // pinned readback
    unsigned long long* m_cnt = nullptr;   // its device-side (mapped) address
    uint32_t* m_bucket = nullptr;          // mapped address of h_bucket
    unsigned long long* h_trace = nullptr; // pinned copy of d_trace, written by traced runs
    unsigned long long* m_trace = nullptr; // its mapped address
    bool trace_valid = false;              // h_trace holds the last traced run's counts
    DevBlock* m_blocks = nullptr;          // mapped address of h_blocks_pinned
    uint8_t* h_ring = nullptr;             // pinned ring for small transfers
    uint8_t* m_ring = nullptr;
    uint64_t ring_cap = 0, ring_pos = 0;
    uint64_t walker_bytes = 0;

    // bucket matrix of the walkers left by the last run (GW_RUN_BUCKETS) or pass
    bool w_valid = false;
    uint32_t* h_bucket = nullptr;          // pinned nblocks^2
    uint64_t unfinished = 0;               // unfinished walkers
    uint64_t pending_hops = 0;             // their remaining hop caps (as of the last bucket pass)

    // cumulative counters
    gw_counters tot{};

    // graphs of run sequences (GW_RUN_GRAPH=1; by default every kernel is launched one by
    // one): the few sequences of a walk (fill runs with a filter pass, runs with every
    // block resident, zero-copy runs) are captured once and replayed.  Events recorded in a
    // graph cannot be timed, so graph runs time the step kernel and the bucket pass's filter
    // with device-clock stamps written into pinned memory (h_ts / m_ts: [0, 1] step kernel,
    // [2, 3] filter), ts_khz ticks per ms.
    bool run_graphs = false;
    unsigned long long *h_ts = nullptr, *m_ts = nullptr;
    double ts_khz = 0;
    struct GraphEntry {
        RunSig sig;
        hipGraphExec_t ex;
        uint64_t used;
    };
    std::vector<GraphEntry> graphs;
    uint64_t graph_clock = 0, graph_runs = 0, graph_captures = 0;

    // launch shape chosen by gw_autotune
    bool have_tuned = false;
    gw_launch tuned{256, 0, 0, 0};

    // fault injection (GW_FAULT=upload:N or alloc:N fails the N-th such call)
    int fault_kind = 0;   // 1 upload, 2 alloc
    long fault_at = 0, n_upload = 0, n_alloc = 0;
};

static std::string g_err;
static thread_local std::string* t_err = nullptr;   // the upload worker's messages go here

static gw_status fail(gw_ctx* c, gw_status st, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (t_err) *t_err = buf; else if (c) c->err = buf; else g_err = buf;
    return st;
}

#define HIPCHK(c, call)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((c), GW_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                               \
    } while (0)

static gw_status dalloc(gw_ctx* c, void** p, uint64_t bytes, bool counted)
{
    *p = nullptr;
    if (bytes == 0) return GW_OK;
    if (c->fault_kind == 2 && ++c->n_alloc == c->fault_at)
        return fail(c, GW_ERR_FAULT, "injected allocation failure (GW_FAULT=alloc:%ld)", c->fault_at);
    if (counted && c->cap && c->held + bytes > c->cap)
        return fail(c, GW_ERR_BUDGET, "HBM budget exceeded: %llu held + %llu requested > %llu cap",
                    (unsigned long long)c->held, (unsigned long long)bytes, (unsigned long long)c->cap);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, GW_ERR_NOMEM, "hipMalloc(%llu) failed: %s", (unsigned long long)bytes, hipGetErrorString(e));
    }
    if (counted) c->held += bytes;
    return GW_OK;
}

static void dfree(gw_ctx* c, void* p, uint64_t bytes, bool counted)
{
    if (!p) return;
    (void)hipFree(p);
    if (counted) c->held -= std::min(c->held, bytes);
}

static void graphs_clear(gw_ctx* c);

static void free_walkers(gw_ctx* c)
{
    // captured runs name the walker arrays
    if (c->s_compute) (void)hipStreamSynchronize(c->s_compute);
    graphs_clear(c);
    dfree(c, c->d_cur, c->nwalkers * 4, true);
    dfree(c, c->d_prev, c->nwalkers * 4, true);
    dfree(c, c->d_meta, c->nwalkers * 4, true);
    dfree(c, c->d_work, c->work_cap * 4, true);
    dfree(c, c->d_hash, c->nwalkers * 8, true);
    dfree(c, c->d_paths, c->nwalkers * (uint64_t)c->path_stride * 4, true);
    dfree(c, c->d_visits, c->nverts_total * 4 * c->visit_parts, true);
    c->d_cur = c->d_prev = c->d_meta = c->d_work = c->d_paths = c->d_visits = nullptr;
    c->d_hash = nullptr;
    c->nwalkers = 0;
    c->work_cap = 0;
    c->walker_bytes = 0;
}

static gw_status enqueue_copy(gw_ctx* c, const CopyJob& j);

static void upload_worker(gw_ctx* c)
{
    (void)hipSetDevice(c->device);
    std::string msg;
    t_err = &msg;
    std::unique_lock<std::mutex> lk(c->wmu);
    for (;;) {
        c->wcv.wait(lk, [&] { return c->wstop || !c->jobs.empty(); });
        if (c->jobs.empty()) break;   // asked to stop, nothing left
        const CopyJob j = c->jobs.front();
        c->jobs.pop_front();
        const bool skip = c->werr != GW_OK;   // after a failure the rest is dropped
        lk.unlock();
        const gw_status st = skip ? GW_OK : enqueue_copy(c, j);
        lk.lock();
        if (j.after_compute) c->cev_pool.push_back(j.after_compute);
        if (st != GW_OK && c->werr == GW_OK) {
            c->werr = st;
            c->werr_msg = msg;
        }
        c->job_done = j.seq;
        c->wcv.notify_all();
    }
}

static gw_status worker_failed(gw_ctx* c)
{
    std::lock_guard<std::mutex> lk(c->wmu);
    if (c->werr == GW_OK) return GW_OK;
    return fail(c, c->werr, "block upload: %s", c->werr_msg.c_str());
}

// Waits until the worker has enqueued the last job of `slot` (every job when slot is -1).
static void wait_enqueued(gw_ctx* c, int64_t slot)
{
    if (!c->worker.joinable()) return;
    std::unique_lock<std::mutex> lk(c->wmu);
    const uint64_t want = slot < 0 ? c->job_seq : c->slot_seq[(size_t)slot];
    c->wcv.wait(lk, [&] { return c->job_done >= want; });
}

static bool enqueued(gw_ctx* c, uint32_t slot)
{
    if (!c->worker.joinable()) return true;
    std::lock_guard<std::mutex> lk(c->wmu);
    return c->job_done >= c->slot_seq[slot];
}

static void stop_worker(gw_ctx* c)
{
    if (!c->worker.joinable()) return;
    {
        std::lock_guard<std::mutex> lk(c->wmu);
        c->wstop = true;
    }
    c->wcv.notify_all();
    c->worker.join();
    c->wstop = false;
}

// Adds the copy-engine time of a slot's last copy to tot.h2d_ms (waiting for the copy
// when `wait`; otherwise only if it has completed).
static void collect_copy_time(gw_ctx* c, uint32_t s, bool wait)
{
    if (!c->slot_timed[s]) return;
    if (wait) wait_enqueued(c, s);
    else if (!enqueued(c, s)) return;
    if (wait) (void)hipEventSynchronize(c->slot_ev1[s]);
    else if (hipEventQuery(c->slot_ev1[s]) != hipSuccess) return;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->slot_ev0[s], c->slot_ev1[s]) == hipSuccess) c->tot.h2d_ms += ms;
    c->slot_timed[s] = 0;
}

// After every copy and decode has completed: the staging ring holds nothing in use.
static void pk_reset(gw_ctx* c)
{
    for (const auto& r : c->pk_live) c->pk_pool.push_back(r.freed);
    for (hipEvent_t& e : c->pk_head_ev)
        if (e) {
            c->pk_pool.push_back(e);
            e = nullptr;
        }
    for (hipEvent_t e : c->pk_retired) c->pk_pool.push_back(e);
    c->pk_live.clear();
    c->pk_retired.clear();
    c->pk_wp = 0;
    c->pk_head_next = 0;
}

// Waits for every block copy (and decode) of every lane.
static hipError_t sync_copies(gw_ctx* c)
{
    wait_enqueued(c, -1);
    hipError_t e = c->s_copy ? hipStreamSynchronize(c->s_copy) : hipSuccess;
    const hipError_t e2 = c->s_copy2 ? hipStreamSynchronize(c->s_copy2) : hipSuccess;
    const hipError_t e3 = c->s_dec ? hipStreamSynchronize(c->s_dec) : hipSuccess;
    if (e == hipSuccess && e2 == hipSuccess && e3 == hipSuccess) pk_reset(c);
    return e != hipSuccess ? e : (e2 != hipSuccess ? e2 : e3);
}

static gw_status ensure_buckets(gw_ctx* c);
static gw_status push_block_table(gw_ctx* c, bool enqueue_copy = true);

static void free_slots(gw_ctx* c)
{
    (void)sync_copies(c);
    for (uint32_t s = 0; s < c->nslots; ++s) {
        collect_copy_time(c, s, true);
        dfree(c, c->slot_beg[s], c->slot_beg_bytes, true);
        dfree(c, c->slot_csr[s], c->slot_csr_bytes, true);
        (void)hipEventDestroy(c->slot_ev0[s]);
        (void)hipEventDestroy(c->slot_ev1[s]);
    }
    c->slot_beg.clear();
    c->slot_csr.clear();
    c->slot_blk.clear();
    c->slot_staged.clear();
    c->slot_ev0.clear();
    c->slot_ev1.clear();
    c->slot_wait.clear();
    c->slot_timed.clear();
    c->slot_lane.clear();
    c->slot_read.clear();
    c->slot_seq.clear();
    c->nslots = 0;
    for (auto& b : c->blk_slot) b = -1;
}

namespace gw {
gw_status gen_fail(gw_ctx* ctx, gw_status st, const char* msg) { return fail(ctx, st, "%s", msg); }

// A region of the pinned ring (256-byte aligned).  Regions are reused only after the
// compute stream -- the only stream that reads or writes them -- has been synchronized.
static uint64_t ring_take(gw_ctx* c, uint64_t bytes, hipError_t* e)
{
    bytes = (bytes + 255) & ~255ULL;
    *e = hipSuccess;
    if (c->ring_pos + bytes > c->ring_cap) {
        *e = hipStreamSynchronize(c->s_compute);
        c->ring_pos = 0;
    }
    const uint64_t at = c->ring_pos;
    c->ring_pos += bytes;
    return at;
}

hipError_t h2d_small(gw_ctx* c, void* dst, const void* src, uint64_t bytes)
{
    if (!bytes) return hipSuccess;
    if (bytes % 4 || bytes > c->ring_cap / 2) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->s_compute);
    hipError_t e;
    const uint64_t at = ring_take(c, bytes, &e);
    if (e != hipSuccess) return e;
    memcpy(c->h_ring + at, src, bytes);
    return launch_copy_u32((uint32_t*)dst, (const uint32_t*)(c->m_ring + at), bytes / 4, c->s_compute);
}

hipError_t d2h_small(gw_ctx* c, void* dst, const void* src, uint64_t bytes)
{
    if (!bytes) return hipSuccess;
    hipError_t e;
    if (bytes % 4 || bytes > c->ring_cap / 2) {
        e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->s_compute);
        return e == hipSuccess ? hipStreamSynchronize(c->s_compute) : e;
    }
    const uint64_t at = ring_take(c, bytes, &e);
    if (e != hipSuccess) return e;
    e = launch_copy_u32((uint32_t*)(c->m_ring + at), (const uint32_t*)src, bytes / 4, c->s_compute);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s_compute);
    if (e == hipSuccess) memcpy(dst, c->h_ring + at, bytes);
    return e;
}
hipStream_t ctx_stream(gw_ctx* ctx)
{
    (void)hipSetDevice(ctx->device);
    return ctx->s_compute;
}
}  // namespace gw

extern "C" {

const char* gw_last_error(const gw_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

gw_status gw_create(const gw_config* cfg, gw_ctx** out)
{
    if (!cfg || !out) return fail(nullptr, GW_ERR_ARG, "gw_create: null argument");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(nullptr, GW_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (cfg->device < 0 || cfg->device >= n)
        return fail(nullptr, GW_ERR_ARG, "device %d out of range [0,%d)", cfg->device, n);
    gw_ctx* c = new gw_ctx();
    c->device = cfg->device;
    c->cap = cfg->cache_bytes;
    if (const char* f = getenv("GW_FAULT")) {
        if (!strncmp(f, "upload:", 7)) { c->fault_kind = 1; c->fault_at = atol(f + 7); }
        else if (!strncmp(f, "alloc:", 6)) { c->fault_kind = 2; c->fault_at = atol(f + 6); }
    }
    auto bail = [&](gw_status st) { gw_destroy(c); return st; };
    if (hipSetDevice(c->device) != hipSuccess) return bail(fail(nullptr, GW_ERR_HIP, "hipSetDevice failed"));
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess)
        return bail(fail(nullptr, GW_ERR_HIP, "hipGetDeviceProperties failed"));
    c->ncu = (uint32_t)prop.multiProcessorCount;
    snprintf(c->info.name, sizeof c->info.name, "%s", prop.gcnArchName);
    c->info.cus = c->ncu;
    c->info.wave_size = (uint32_t)prop.warpSize;
    c->info.hbm_bytes = prop.totalGlobalMem;
    c->info.lds_per_cu = (uint32_t)prop.maxSharedMemoryPerMultiProcessor;
    if (hipStreamCreateWithFlags(&c->s_compute, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->s_copy, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->s_read, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->s_dec, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(nullptr, GW_ERR_HIP, "stream creation failed"));
    if (const char* f = getenv("GW_PACK_INLINE_DECODE")) c->inline_decode = atoi(f) != 0;
    if (const char* f = getenv("GW_PACK_INPLACE")) c->no_inplace = atoi(f) == 0;
    if (const char* f = getenv("GW_UPLOAD_THREAD")) c->sync_uploads = atoi(f) == 0;
    if (const char* f = getenv("GW_RUN_GRAPH")) c->run_graphs = atoi(f) != 0;
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0)
            c->run_graphs = false;   // no clock to time graph runs with
        c->ts_khz = khz;
    }
    hipEvent_t* evs[] = {&c->ev_f0, &c->ev_f1, &c->ev_k0, &c->ev_k1, &c->ev_c1};
    for (hipEvent_t* ev : evs)
        if (hipEventCreate(ev) != hipSuccess) return bail(fail(nullptr, GW_ERR_HIP, "event creation failed"));
    void* p = nullptr;
    if (dalloc(c, &p, 4 * sizeof(uint32_t), false) != GW_OK) return bail(fail(nullptr, GW_ERR_NOMEM, "%s", c->err.c_str()));
    c->d_misc = (uint32_t*)p;
    if (dalloc(c, &p, 16 * sizeof(unsigned long long), false) != GW_OK)
        return bail(fail(nullptr, GW_ERR_NOMEM, "%s", c->err.c_str()));
    c->d_cnt = (unsigned long long*)p;
    if (hipHostMalloc((void**)&c->h_cnt, 16 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer((void**)&c->m_cnt, c->h_cnt, 0) != hipSuccess)
        return bail(fail(nullptr, GW_ERR_NOMEM, "pinned counter buffer"));
    if (hipHostMalloc((void**)&c->h_ts, 4 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer((void**)&c->m_ts, c->h_ts, 0) != hipSuccess)
        return bail(fail(nullptr, GW_ERR_NOMEM, "pinned timestamp buffer"));
    c->ring_cap = 4ull << 20;
    if (hipHostMalloc((void**)&c->h_ring, c->ring_cap, hipHostMallocDefault) != hipSuccess ||
        hipHostGetDevicePointer((void**)&c->m_ring, c->h_ring, 0) != hipSuccess)
        return bail(fail(nullptr, GW_ERR_NOMEM, "pinned transfer ring"));
    *out = c;
    return GW_OK;
}

gw_status gw_destroy(gw_ctx* c)
{
    if (!c) return GW_OK;
    (void)hipSetDevice(c->device);
    if (c->s_compute) (void)hipStreamSynchronize(c->s_compute);
    (void)sync_copies(c);
    stop_worker(c);
    graphs_clear(c);
    free_walkers(c);
    free_slots(c);
    dfree(c, c->d_pack, c->pack_cap, true);
    dfree(c, c->d_bucket, 0, false);
    dfree(c, c->d_bpart, 0, false);
    dfree(c, c->d_trace, 0, false);
    dfree(c, c->d_feat, 0, false);
    dfree(c, c->d_age, 0, false);
    dfree(c, c->d_blocks, 0, false);
    dfree(c, c->d_vcut, 0, false);
    dfree(c, c->d_misc, 0, false);
    dfree(c, c->d_cnt, 0, false);
    if (c->h_cnt) (void)hipHostFree(c->h_cnt);
    if (c->h_blocks_pinned) (void)hipHostFree(c->h_blocks_pinned);
    if (c->h_bucket) (void)hipHostFree(c->h_bucket);
    if (c->h_trace) (void)hipHostFree(c->h_trace);
    if (c->h_ring) (void)hipHostFree(c->h_ring);
    if (c->h_ts) (void)hipHostFree(c->h_ts);
    hipEvent_t evs[] = {c->ev_f0, c->ev_f1, c->ev_k0, c->ev_k1, c->ev_c1};
    for (hipEvent_t ev : evs) if (ev) (void)hipEventDestroy(ev);
    if (c->s_compute) (void)hipStreamDestroy(c->s_compute);
    if (c->s_copy) (void)hipStreamDestroy(c->s_copy);
    if (c->s_copy2) (void)hipStreamDestroy(c->s_copy2);
    if (c->s_read) (void)hipStreamDestroy(c->s_read);
    pk_reset(c);
    for (hipEvent_t e : c->pk_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->cev_pool) (void)hipEventDestroy(e);
    if (c->s_dec) (void)hipStreamDestroy(c->s_dec);
    delete c;
    return GW_OK;
}

gw_status gw_device_info_get(gw_ctx* c, gw_device_info* out)
{
    if (!c || !out) return fail(c, GW_ERR_ARG, "null argument");
    *out = c->info;
    return GW_OK;
}

gw_status gw_alloc_pinned(size_t bytes, void** out)
{
    if (!out) return fail(nullptr, GW_ERR_ARG, "null argument");
    *out = nullptr;
    if (bytes == 0) return GW_OK;
    hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(nullptr, GW_ERR_NOMEM, "hipHostMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    return GW_OK;
}

gw_status gw_free_pinned(void* p)
{
    if (p && hipHostFree(p) != hipSuccess) return fail(nullptr, GW_ERR_HIP, "hipHostFree failed");
    return GW_OK;
}

gw_status gw_set_block_table(gw_ctx* c, const gw_block_desc* b, uint32_t nb)
{
    if (!c || !b || nb == 0) return fail(c, GW_ERR_ARG, "gw_set_block_table: empty table");
    if (nb > 65535) return fail(c, GW_ERR_ARG, "at most 65535 blocks are supported (got %u)", nb);
    for (uint32_t i = 0; i < nb; ++i) {
        if (b[i].blk != i) return fail(c, GW_ERR_ARG, "block %u has id %u (ids must be 0..n-1 in order)", i, b[i].blk);
        if (i && b[i].start_vert != b[i - 1].start_vert + b[i - 1].nverts)
            return fail(c, GW_ERR_ARG, "block %u does not start where block %u ends", i, i - 1);
    }
    if (c->nwalkers) return fail(c, GW_ERR_STATE, "gw_set_block_table after gw_init_walkers");
    (void)hipSetDevice(c->device);
    free_slots(c);
    dfree(c, c->d_blocks, 0, false);
    dfree(c, c->d_vcut, 0, false);
    dfree(c, c->d_bucket, 0, false);
    dfree(c, c->d_bpart, 0, false);
    dfree(c, c->d_trace, 0, false);
    dfree(c, c->d_feat, 0, false);
    dfree(c, c->d_age, 0, false);
    c->d_blocks = nullptr; c->d_vcut = nullptr; c->d_bucket = nullptr;
    c->d_bpart = nullptr;
    c->d_trace = nullptr; c->d_feat = nullptr; c->d_age = nullptr;
    if (c->h_blocks_pinned) { (void)hipHostFree(c->h_blocks_pinned); c->h_blocks_pinned = nullptr; }
    if (c->h_bucket) { (void)hipHostFree(c->h_bucket); c->h_bucket = nullptr; }
    if (c->h_trace) { (void)hipHostFree(c->h_trace); c->h_trace = nullptr; }
    c->w_valid = false;
    c->trace_valid = false;
    c->nblocks = nb;
    c->desc.assign(b, b + nb);
    c->pack_checked.assign(nb, gw_ctx::PackSeen{nullptr, 0, 0});
    c->hblk.assign(nb, DevBlock{});
    c->blk_slot.assign(nb, -1);
    c->map_beg.assign(nb, nullptr);
    c->map_csr.assign(nb, nullptr);
    c->vcut.resize(nb + 1);
    for (uint32_t i = 0; i < nb; ++i) {
        DevBlock& d = c->hblk[i];
        d.start_vert = b[i].start_vert;
        d.nverts = b[i].nverts;
        d.start_edge = b[i].start_edge;
        d.nedges = b[i].nedges;
        c->vcut[i] = b[i].start_vert;
    }
    c->vcut[nb] = b[nb - 1].start_vert + b[nb - 1].nverts;
    c->nverts_total = c->vcut[nb];
    void* p;
    gw_status st;
    if ((st = dalloc(c, &p, sizeof(DevBlock) * nb, false)) != GW_OK) return st;
    c->d_blocks = (DevBlock*)p;
    if ((st = dalloc(c, &p, sizeof(uint32_t) * (nb + 1), false)) != GW_OK) return st;
    c->d_vcut = (uint32_t*)p;
    if ((st = dalloc(c, &p, sizeof(uint32_t) * (uint64_t)nb * nb, false)) != GW_OK) return st;
    c->d_bucket = (uint32_t*)p;
    // the bucket passes' per-workgroup rows (two per CU for the run's worklist filter): table
    // metadata like the matrix itself, outside the budget (at most 96 KB per CU: 24 MB on 256 CUs)
    if ((uint64_t)nb * nb <= kFilterLdsBuckets) {
        if ((st = dalloc(c, &p, sizeof(uint32_t) * (uint64_t)nb * nb * c->ncu * 2, false)) != GW_OK) return st;
        c->d_bpart = (uint32_t*)p;
    }
    if ((st = dalloc(c, &p, sizeof(unsigned long long) * nb, false)) != GW_OK) return st;
    c->d_trace = (unsigned long long*)p;
    HIPCHK(c, hipMemset(c->d_trace, 0, sizeof(unsigned long long) * nb));
    // features, then the means and deviations that standardize them for the network
    if ((st = dalloc(c, &p, sizeof(float) * ((uint64_t)nb + 2) * kNumFeatures, false)) != GW_OK) return st;
    c->d_feat = (float*)p;
    if ((st = dalloc(c, &p, sizeof(float) * nb, false)) != GW_OK) return st;
    c->d_age = (float*)p;
    HIPCHK(c, hipHostMalloc((void**)&c->h_blocks_pinned, sizeof(DevBlock) * nb, hipHostMallocDefault));
    HIPCHK(c, hipHostMalloc((void**)&c->h_bucket, sizeof(uint32_t) * (uint64_t)nb * nb, hipHostMallocDefault));
    HIPCHK(c, hipHostGetDevicePointer((void**)&c->m_bucket, c->h_bucket, 0));
    HIPCHK(c, hipHostMalloc((void**)&c->h_trace, sizeof(unsigned long long) * nb, hipHostMallocDefault));
    HIPCHK(c, hipHostGetDevicePointer((void**)&c->m_trace, c->h_trace, 0));
    HIPCHK(c, hipHostGetDevicePointer((void**)&c->m_blocks, c->h_blocks_pinned, 0));
    HIPCHK(c, hipMemcpy(c->d_vcut, c->vcut.data(), sizeof(uint32_t) * (nb + 1), hipMemcpyHostToDevice));
    return GW_OK;
}

// The two buffers of one slot, sized for the largest block of the table: offsets
// (nverts + 1 u64) and edges.  A block's edge words start 4 x (start_edge mod 16) bytes
// into the slot, so every word keeps its offset modulo 64 bytes from the whole edge array
// (the window search reads aligned 64-byte pieces); 64 B of head room and 64 B of tail
// pad.  gw_slot_bytes and gw_reserve_slots both take their sizes from here, so the bytes a
// caller budgets per slot are the bytes a reservation charges.
// A slot: beg of the largest vertex count, and the edge words of the largest block with 64 B
// of head room and tail pad each, plus slack for decoding a packed block in place (its
// payload sits at the buffer's end while the decoder writes from the start; with no slack
// the last chunk's output would cover its own input): 1/64 of the edge words, between 4 KB
// and 1 MB (C3: 54 MB over 54 slots; a largest block then decodes in ~8 passes).
static void slot_sizes(const gw_ctx* c, uint64_t* beg_bytes, uint64_t* csr_bytes)
{
    uint64_t mb = 0, me = 0;
    for (auto& d : c->desc) { mb = std::max<uint64_t>(mb, (uint64_t)d.nverts + 1); me = std::max<uint64_t>(me, d.nedges); }
    const uint64_t slack = std::min<uint64_t>(1ull << 20, std::max<uint64_t>(4096, me * 4 / 64));
    *beg_bytes = (mb * 8 + 255) & ~255ULL;
    *csr_bytes = (me * 4 + 128 + slack + 255) & ~255ULL;
}

gw_status gw_slot_bytes(gw_ctx* c, uint64_t* bytes)
{
    if (!c || !bytes) return fail(c, GW_ERR_ARG, "null argument");
    if (!c->nblocks) return fail(c, GW_ERR_STATE, "no block table");
    uint64_t b, e;
    slot_sizes(c, &b, &e);
    *bytes = b + e;
    return GW_OK;
}

gw_status gw_reserve_slots(gw_ctx* c, uint32_t nslots)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (!c->nblocks) return fail(c, GW_ERR_STATE, "gw_reserve_slots before gw_set_block_table");
    (void)hipSetDevice(c->device);
    HIPCHK(c, sync_copies(c));
    HIPCHK(c, hipStreamSynchronize(c->s_compute));
    free_slots(c);
    for (auto& d : c->hblk) { d.resident = 0; d.beg = nullptr; d.csr = nullptr; }
    slot_sizes(c, &c->slot_beg_bytes, &c->slot_csr_bytes);
    uint64_t need = (uint64_t)nslots * (c->slot_beg_bytes + c->slot_csr_bytes);
    if (c->cap && c->held + need > c->cap)
        return fail(c, GW_ERR_BUDGET, "%u slots of %llu B plus %llu B held exceed the %llu B cache budget", nslots,
                    (unsigned long long)(c->slot_beg_bytes + c->slot_csr_bytes), (unsigned long long)c->held,
                    (unsigned long long)c->cap);
    c->slot_beg.assign(nslots, nullptr);
    c->slot_csr.assign(nslots, nullptr);
    c->slot_blk.assign(nslots, -1);
    c->slot_staged.assign(nslots, -1);
    c->slot_ev0.assign(nslots, nullptr);
    c->slot_ev1.assign(nslots, nullptr);
    c->slot_wait.assign(nslots, 0);
    c->slot_timed.assign(nslots, 0);
    c->slot_lane.assign(nslots, 0);
    c->slot_read.assign(nslots, 0);
    c->slot_seq.assign(nslots, 0);
    c->nslots = nslots;
    for (uint32_t s = 0; s < nslots; ++s) {
        void* p;
        gw_status st;
        if (hipEventCreate(&c->slot_ev0[s]) != hipSuccess || hipEventCreate(&c->slot_ev1[s]) != hipSuccess) {
            free_slots(c);
            return fail(c, GW_ERR_HIP, "hipEventCreate for slot %u", s);
        }
        if ((st = dalloc(c, &p, c->slot_beg_bytes, true)) != GW_OK) { free_slots(c); return st; }
        c->slot_beg[s] = (uint64_t*)p;
        if ((st = dalloc(c, &p, c->slot_csr_bytes, true)) != GW_OK) { free_slots(c); return st; }
        c->slot_csr[s] = (uint32_t*)p;
    }
    // re-expose zero-copy mappings that exist
    for (uint32_t i = 0; i < c->nblocks; ++i)
        if (c->map_beg[i]) { c->hblk[i].beg = c->map_beg[i]; c->hblk[i].csr = c->map_csr[i]; c->hblk[i].resident = kMapped; }
    return GW_OK;
}

static void block_nonresident(gw_ctx* c, uint32_t blk)
{
    DevBlock& d = c->hblk[blk];
    c->blk_slot[blk] = -1;
    if (c->map_beg[blk]) { d.beg = c->map_beg[blk]; d.csr = c->map_csr[blk]; d.resident = kMapped; }
    else { d.beg = nullptr; d.csr = nullptr; d.resident = 0; }
}

// where block blk's edge words go in a slot (see gw_reserve_slots)
static uint32_t* slot_edges(gw_ctx* c, uint32_t slot, uint32_t blk)
{
    return c->slot_csr[slot] + (c->desc[blk].start_edge & 15);
}

// Staging buffer of packed uploads (one copy lane): two head areas of a quarter each at its
// front, the arc-payload ring over the other half, which must hold at least kPkMinRing
// words (pieces of half the ring hold at least the largest chunk's 272 words).
constexpr uint64_t kPkMinRing = 1024;
static uint64_t pk_head_bytes(const gw_ctx* c) { return (c->pack_cap / 4) & ~255ULL; }

// In-place decode of a block's arc payload, copied whole to the end of the slot's edge
// buffer at byte pstart: the decoder writes the block's edge words from byte E of that buffer
// up, 1024 bytes per 256-arc chunk, while the payload of chunk c starts at byte
// pstart + 4 off(c).  A pass decodes chunks [c0, c1) whose outputs all end at or below the
// first payload byte not yet read (chunk c0's), so no pass overwrites input that it or a
// later pass still reads; passes run in order on one stream.  Outputs grow by 1024 bytes per
// chunk and inputs by ~560 (C3), so a block whose payload fits below its slot's slack
// decodes in a handful of passes.  Fills cuts with c0 = 0, ..., nch; false when a pass
// would make no progress or more than kMaxPasses are needed.
constexpr size_t kMaxPasses = 48;
static bool inplace_plan(const uint32_t* ah, uint64_t nch, uint64_t ne, uint64_t total_w, uint64_t pstart, uint64_t E,
                         std::vector<uint64_t>& cuts)
{
    auto off = [&](uint64_t k) { return k < nch ? (uint64_t)ah[k * 4] : total_w; };
    cuts.assign(1, 0);
    uint64_t c0 = 0;
    while (c0 < nch) {
        const uint64_t lim = pstart + 4 * off(c0);   // first payload byte not yet read
        if (lim < E) return false;
        // the chunks whose outputs end by lim: whole chunks of 1024 bytes, or all of them
        // once the last (possibly partial) chunk's end, E + 4 ne, fits
        const uint64_t c1 = E + 4 * ne <= lim ? nch : std::min<uint64_t>(nch, (lim - E) / 1024);
        if (c1 <= c0 || cuts.size() > kMaxPasses) return false;
        cuts.push_back(c1);
        c0 = c1;
    }
    return true;
}

// An event for the ring: from the pool, or the oldest retired one once its recorded work
// has completed (a stream wait enqueued on it has then been satisfied), or a new one.
// Retired events complete roughly in the order they were retired, so only the oldest few
// are queried.
static gw_status pk_event(gw_ctx* c, hipEvent_t* ev)
{
    if (c->pk_pool.empty()) {
        size_t done = 0;
        for (; done < c->pk_retired.size() && done < 4; ++done) {
            const hipError_t q = hipEventQuery(c->pk_retired[done]);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) return fail(c, GW_ERR_HIP, "staging event: %s", hipGetErrorString(q));
        }
        c->pk_pool.insert(c->pk_pool.end(), c->pk_retired.begin(), c->pk_retired.begin() + done);
        c->pk_retired.erase(c->pk_retired.begin(), c->pk_retired.begin() + done);
    }
    if (!c->pk_pool.empty()) {
        *ev = c->pk_pool.back();
        c->pk_pool.pop_back();
        return GW_OK;
    }
    HIPCHK(c, hipEventCreateWithFlags(ev, hipEventDisableTiming));
    return GW_OK;
}

// Takes n words (a multiple of 64) of the ring at the write position, or at its start when
// they do not fit before the end, once the decodes that still read any of them are done.
// The wait is the host's (the upload worker's thread), not the copy stream's: a copy queued
// behind a cross-stream wait runs at a fraction of the link rate (ring pieces of 60 MB took
// 1.9-4.6 ms instead of 1.03 ms), while the decode of the region taken next has normally
// finished before the copy in flight does.
static gw_status pk_take(gw_ctx* c, uint64_t n, uint64_t ring_w, uint64_t* at)
{
    if (n > ring_w) return fail(c, GW_ERR_STATE, "staging ring piece of %llu words > ring of %llu",
                                (unsigned long long)n, (unsigned long long)ring_w);
    uint64_t a = c->pk_wp;
    if (a + n > ring_w) a = 0;
    const uint64_t b = a + n;
    for (size_t i = 0; i < c->pk_live.size();) {
        const gw_ctx::PkRegion r = c->pk_live[i];
        if (r.a < b && a < r.b) {
            HIPCHK(c, hipEventSynchronize(r.freed));
            c->pk_pool.push_back(r.freed);
            c->pk_live[i] = c->pk_live.back();
            c->pk_live.pop_back();
        } else {
            ++i;
        }
    }
    c->pk_wp = b;
    *at = a;
    return GW_OK;
}

// Records an event on `from` that `to` waits for (the event is retired at once).
static gw_status pk_signal(gw_ctx* c, hipStream_t from, hipStream_t to)
{
    hipEvent_t ev;
    gw_status st;
    if ((st = pk_event(c, &ev)) != GW_OK) return st;
    HIPCHK(c, hipEventRecord(ev, from));
    HIPCHK(c, hipStreamWaitEvent(to, ev, 0));
    c->pk_retired.push_back(ev);
    return GW_OK;
}

// A packed block with one copy lane: its headers into the next head area on s_copy and beg
// decoded into the slot on s_dec; its arc payload copied whole to the end of the slot's edge
// buffer and decoded there in passes (inplace_plan), or, when that plan fails, through the
// staging ring in pieces of half the ring, each decoded as its copy lands.  The slot's copy
// events bracket the block: the start on s_copy, the end after its last decode.
static gw_status ring_copy(gw_ctx* c, uint32_t slot, uint32_t blk, const uint32_t* pw, uint64_t words,
                           const PackLayout& lay, hipEvent_t after_compute)
{
    const gw_block_desc& d = c->desc[blk];
    const uint64_t hw = pk_head_bytes(c) / 4;
    const uint64_t ring_w = c->pack_cap / 4 - 2 * hw;
    const uint64_t piece_w = (ring_w / 2) & ~63ULL;
    const uint32_t h = c->pk_head_next;
    c->pk_head_next ^= 1u;
    uint32_t* head = c->d_pack + h * hw;
    uint32_t* ring = c->d_pack + 2 * hw;
    gw_status st;
    if (c->pk_head_ev[h]) {   // the head area after the decodes of the block that used it last
        HIPCHK(c, hipEventSynchronize(c->pk_head_ev[h]));
        c->pk_pool.push_back(c->pk_head_ev[h]);
        c->pk_head_ev[h] = nullptr;
    }
    HIPCHK(c, hipEventRecord(c->slot_ev0[slot], c->s_copy));
    HIPCHK(c, hipMemcpyAsync(head, pw, lay.head_w * 4, hipMemcpyHostToDevice, c->s_copy));
    if ((st = pk_signal(c, c->s_copy, c->s_dec)) != GW_OK) return st;
    // the decoder writes the slot: after the compute stream's work that may still read it
    if (after_compute) HIPCHK(c, hipStreamWaitEvent(c->s_dec, after_compute, 0));
    HIPCHK(c, launch_unpack_beg(head, lay, d.nverts, c->slot_beg[slot], c->s_dec));
    const uint32_t* ah = pw + lay.ah_off;
    const uint64_t total_w = words - lay.head_w;
    // in place: the payload lands in the slot itself (one copy), decoded in a few passes
    std::vector<uint64_t> cuts;
    const uint64_t E = 4 * (d.start_edge & 15);
    if (!c->no_inplace && c->slot_csr_bytes >= 4 * total_w + 256) {
        const uint64_t pstart = (c->slot_csr_bytes - 4 * total_w) & ~255ULL;
        if (inplace_plan(ah, lay.nch, d.nedges, total_w, pstart, E, cuts)) {
            // the copy writes the slot: after the compute stream's work that may still read it
            if (after_compute) HIPCHK(c, hipEventSynchronize(after_compute));
            uint32_t* payload = (uint32_t*)((uint8_t*)c->slot_csr[slot] + pstart);
            HIPCHK(c, hipMemcpyAsync(payload, pw + lay.head_w, total_w * 4, hipMemcpyHostToDevice, c->s_copy));
            if ((st = pk_signal(c, c->s_copy, c->s_dec)) != GW_OK) return st;
            for (size_t i = 0; i + 1 < cuts.size(); ++i)
                HIPCHK(c, launch_unpack_arcs(head, payload, lay, c->slot_beg[slot], d.start_edge, d.nverts, d.nedges,
                                             cuts[i], cuts[i + 1], 0, total_w, slot_edges(c, slot, blk), c->s_dec));
            c->inplace_blocks++;
            hipEvent_t ev;
            if ((st = pk_event(c, &ev)) != GW_OK) return st;
            HIPCHK(c, hipEventRecord(ev, c->s_dec));
            c->pk_head_ev[h] = ev;
            HIPCHK(c, hipEventRecord(c->slot_ev1[slot], c->s_dec));
            return GW_OK;
        }
    }
    // pieces: the most chunks whose payload fits a piece (chunk offsets are contiguous, so
    // chunks [c0, c1) take words [off(c0), off(c1)), off(nch) = the payload size)
    auto off = [&](uint64_t k) { return k < lay.nch ? (uint64_t)ah[k * 4] : total_w; };
    uint64_t c0 = 0;
    while (c0 < lay.nch) {
        const uint64_t w0 = off(c0);
        uint64_t lo = c0, hi = lay.nch;
        while (lo < hi) {
            const uint64_t mid = (lo + hi + 1) / 2;
            if (off(mid) - w0 <= piece_w) lo = mid; else hi = mid - 1;
        }
        const uint64_t c1 = lo;
        if (c1 == c0) return fail(c, GW_ERR_ARG, "a chunk of block %u does not fit a staging piece", blk);
        const uint64_t w1 = off(c1);
        if (w1 > total_w) return fail(c, GW_ERR_ARG, "block %u: packed chunk offsets past its end", blk);
        const uint64_t n = w1 - w0, nr = (n + 63) & ~63ULL;
        uint64_t a = 0;
        if (n) {
            if ((st = pk_take(c, nr, ring_w, &a)) != GW_OK) return st;
            HIPCHK(c, hipMemcpyAsync(ring + a, pw + lay.head_w + w0, n * 4, hipMemcpyHostToDevice, c->s_copy));
            if ((st = pk_signal(c, c->s_copy, c->s_dec)) != GW_OK) return st;
        }
        HIPCHK(c, launch_unpack_arcs(head, ring + a, lay, c->slot_beg[slot], d.start_edge, d.nverts, d.nedges, c0,
                                     c1, w0, n, slot_edges(c, slot, blk), c->s_dec));
        if (n) {
            hipEvent_t ev;
            if ((st = pk_event(c, &ev)) != GW_OK) return st;
            HIPCHK(c, hipEventRecord(ev, c->s_dec));
            c->pk_live.push_back({a, a + nr, ev});
        }
        c0 = c1;
    }
    c->ring_blocks++;
    hipEvent_t ev;
    if ((st = pk_event(c, &ev)) != GW_OK) return st;
    HIPCHK(c, hipEventRecord(ev, c->s_dec));
    c->pk_head_ev[h] = ev;
    HIPCHK(c, hipEventRecord(c->slot_ev1[slot], c->s_dec));
    return GW_OK;
}

// Starts the copy of block `blk` into `slot` (no host wait): checks and books it here and
// hands it to the upload worker (one copy lane) or enqueues it (two lanes).  The slot's
// previous block stops being readable at once; if the slot ever held a readable block and
// the compute stream still has work queued, the write into the slot waits for that work,
// which may still read it.  A slot that only ever held staged, uncommitted copies has no
// readers.  The wait is left out whenever it is not needed: a copy queued behind a
// cross-stream wait has been measured to run at half the link rate for its whole length
// (15 of 53 block copies of a 16 GB fill, 8.8 ms instead of 4.7 ms); through the staging
// ring the wait is the decoder's, and copies never wait for the compute stream.
static gw_status start_copy(gw_ctx* c, uint32_t slot, uint32_t blk, const uint64_t* beg, const uint32_t* csr,
                            const void* packed = nullptr, uint64_t packed_bytes = 0)
{
    const gw_block_desc& d = c->desc[blk];
    if (packed ? !d.nedges : (!beg || (d.nedges && !csr)))
        return fail(c, GW_ERR_ARG, packed ? "a block without arcs is uploaded raw" : "null block buffer");
    // the lane: the one of the slot's last copy while that copy may still run (a stream keeps
    // its copies in order, so two copies never write one slot at once, and no copy waits on
    // another stream), else the next in turn
    uint32_t lane = 0;
    if (c->lanes > 1) {
        const hipError_t q = hipEventQuery(c->slot_ev1[slot]);
        if (q != hipSuccess && q != hipErrorNotReady) return fail(c, GW_ERR_HIP, "copy event: %s", hipGetErrorString(q));
        if (q == hipErrorNotReady) {
            lane = c->slot_lane[slot];
        } else {
            lane = c->next_lane;
            c->next_lane = (c->next_lane + 1) % c->lanes;
        }
    }
    const uint64_t lane_cap = (c->pack_cap / c->lanes) & ~255ULL;   // staging bytes of each lane
    // a packed block: its headers and beg payload go to the staging buffer whole, its arc
    // payload in pieces of what is left of the buffer (at least the largest chunk, 1088 B)
    PackLayout lay{};
    if (packed) {
        if (packed_bytes & 3) return fail(c, GW_ERR_ARG, "packed block size must be a multiple of 4");
        // every header is checked the first time a block is sent from a buffer; later uploads
        // of the same buffer read only its first words (the decoders keep every read inside
        // the staging buffer whatever the headers say)
        const gw_ctx::PackSeen seen{packed, packed_bytes, pack_fingerprint((const uint32_t*)packed, packed_bytes / 4)};
        const bool checked = c->pack_checked[blk] == seen;
        if (const char* e = pack_layout((const uint32_t*)packed, packed_bytes / 4, d.nverts, d.nedges, !checked, &lay))
            return fail(c, GW_ERR_ARG, "block %u: %s", blk, e);
        c->pack_checked[blk] = seen;
        if (lay.head_w * 4 + 4 * (8 * 31 + 24) > lane_cap)
            return fail(c, GW_ERR_ARG, "the %llu B staging buffer of a copy lane cannot hold the block's %llu B of "
                        "headers and offsets plus one chunk (gw_set_pack_staging)", (unsigned long long)lane_cap,
                        (unsigned long long)(lay.head_w * 4));
    }
    if (c->fault_kind == 1 && ++c->n_upload == c->fault_at)
        return fail(c, GW_ERR_FAULT, "injected upload failure (GW_FAULT=upload:%ld)", c->fault_at);
    gw_status st;
    if ((st = worker_failed(c)) != GW_OK) return st;
    (void)hipSetDevice(c->device);
    if (c->slot_blk[slot] >= 0) block_nonresident(c, (uint32_t)c->slot_blk[slot]);
    c->slot_blk[slot] = -1;
    c->slot_staged[slot] = -1;
    collect_copy_time(c, slot, true);   // the slot's events are about to be re-recorded
    CopyJob j;
    j.slot = slot;
    j.blk = blk;
    j.lane = lane;
    j.beg = beg;
    j.csr = csr;
    j.packed = (const uint32_t*)packed;
    j.words = packed_bytes / 4;
    j.lay = lay;
    j.ring = packed && c->lanes == 1 && !c->inline_decode && lay.head_w * 4 <= pk_head_bytes(c) &&
             c->pack_cap / 4 >= 2 * (pk_head_bytes(c) / 4) + kPkMinRing;
    // the compute stream's work queued now may still read the slot's previous block: the
    // copy (or the decode) that writes the slot waits for it, and only for it
    if (c->slot_read[slot] && hipStreamQuery(c->s_compute) != hipSuccess) {
        {
            std::lock_guard<std::mutex> lk(c->wmu);
            if (!c->cev_pool.empty()) {
                j.after_compute = c->cev_pool.back();
                c->cev_pool.pop_back();
            }
        }
        if (!j.after_compute) HIPCHK(c, hipEventCreateWithFlags(&j.after_compute, hipEventDisableTiming));
        HIPCHK(c, hipEventRecord(j.after_compute, c->s_compute));
    }
    const uint64_t bb = ((uint64_t)d.nverts + 1) * 8, cb = d.nedges * 4;
    c->slot_timed[slot] = 1;
    c->slot_lane[slot] = (uint8_t)lane;
    c->tot.h2d_bytes += packed ? packed_bytes : bb + cb;
    c->tot.h2d_raw_bytes += bb + cb;
    c->tot.h2d_uploads += 1;
    if (c->lanes == 1 && !c->sync_uploads) {
        if (!c->worker.joinable()) c->worker = std::thread(upload_worker, c);
        std::lock_guard<std::mutex> lk(c->wmu);
        j.seq = ++c->job_seq;
        c->slot_seq[slot] = j.seq;
        c->jobs.push_back(j);
        c->wcv.notify_all();
        return GW_OK;
    }
    // two copy lanes choose a lane from the slot's copy events: enqueued here, in order
    wait_enqueued(c, -1);
    st = enqueue_copy(c, j);
    if (j.after_compute) {
        std::lock_guard<std::mutex> lk(c->wmu);
        c->cev_pool.push_back(j.after_compute);
    }
    return st;
}

// Puts one block copy on the streams (the upload worker's thread, or the caller's): a
// packed block through the staging ring when the job says so, else copied and decoded on
// its lane's copy stream.
static gw_status enqueue_copy(gw_ctx* c, const CopyJob& j)
{
    const uint32_t slot = j.slot, blk = j.blk;
    const gw_block_desc& d = c->desc[blk];
    if (j.ring) return ring_copy(c, slot, blk, j.packed, j.words, j.lay, j.after_compute);
    hipStream_t cs = j.lane ? c->s_copy2 : c->s_copy;
    const uint64_t lane_cap = (c->pack_cap / c->lanes) & ~255ULL;   // staging bytes of each lane
    uint32_t* stage = c->d_pack + j.lane * (lane_cap / 4);
    const PackLayout& lay = j.lay;
    if (j.packed && (!c->pk_live.empty() || c->pk_head_ev[0] || c->pk_head_ev[1])) {
        // decoding on the copy stream reuses the staging buffer from its start: after every
        // decode of the ring
        HIPCHK(c, hipEventRecord(c->ev_c1, c->s_dec));
        HIPCHK(c, hipStreamWaitEvent(cs, c->ev_c1, 0));
        for (const auto& r : c->pk_live) c->pk_retired.push_back(r.freed);
        for (hipEvent_t& e : c->pk_head_ev)
            if (e) {
                c->pk_retired.push_back(e);
                e = nullptr;
            }
        c->pk_live.clear();
        c->pk_wp = 0;
    }
    if (j.after_compute) HIPCHK(c, hipStreamWaitEvent(cs, j.after_compute, 0));
    HIPCHK(c, hipEventRecord(c->slot_ev0[slot], cs));
    if (j.packed) {
        c->inline_blocks++;
        // the copies and decodes share the copy stream, so each copy into the staging buffer
        // queues behind the decode that read the previous piece
        const uint32_t* pw = j.packed;
        const uint32_t* ah = pw + lay.ah_off;
        const uint64_t piece_w = (lane_cap / 4) - lay.head_w;
        const uint64_t total_w = j.words - lay.head_w;
        HIPCHK(c, hipMemcpyAsync(stage, pw, lay.head_w * 4, hipMemcpyHostToDevice, cs));
        HIPCHK(c, launch_unpack_beg(stage, lay, d.nverts, c->slot_beg[slot], cs));
        // pieces: the most chunks whose payload fits (chunk offsets are contiguous, so chunks
        // [c0, c1) take words [off(c0), off(c1)), off(nch) = total_w; a binary search)
        auto off = [&](uint64_t k) { return k < lay.nch ? (uint64_t)ah[k * 4] : total_w; };
        uint64_t c0 = 0;
        while (c0 < lay.nch) {
            const uint64_t w0 = off(c0);
            uint64_t lo = c0, hi = lay.nch;   // off(lo) - w0 <= piece_w; find the largest such
            while (lo < hi) {
                const uint64_t mid = (lo + hi + 1) / 2;
                if (off(mid) - w0 <= piece_w) lo = mid; else hi = mid - 1;
            }
            const uint64_t c1 = lo;
            if (c1 == c0) return fail(c, GW_ERR_ARG, "a chunk of block %u does not fit the staging buffer", blk);
            const uint64_t w1 = off(c1);
            if (w1 > total_w) return fail(c, GW_ERR_ARG, "block %u: packed chunk offsets past its end", blk);
            if (w1 > w0)
                HIPCHK(c, hipMemcpyAsync(stage + lay.head_w, pw + lay.head_w + w0, (w1 - w0) * 4,
                                         hipMemcpyHostToDevice, cs));
            HIPCHK(c, launch_unpack_arcs(stage, stage + lay.head_w, lay, c->slot_beg[slot], d.start_edge, d.nverts,
                                         d.nedges, c0, c1, w0, w1 - w0, slot_edges(c, slot, blk), cs));
            c0 = c1;
        }
    } else {
        const uint64_t bb = ((uint64_t)d.nverts + 1) * 8, cb = d.nedges * 4;
        HIPCHK(c, hipMemcpyAsync(c->slot_beg[slot], j.beg, bb, hipMemcpyHostToDevice, cs));
        if (cb) HIPCHK(c, hipMemcpyAsync(slot_edges(c, slot, blk), j.csr, cb, hipMemcpyHostToDevice, cs));
    }
    HIPCHK(c, hipEventRecord(c->slot_ev1[slot], cs));
    return GW_OK;
}

gw_status gw_set_pack_staging(gw_ctx* c, uint64_t bytes)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    (void)hipSetDevice(c->device);
    HIPCHK(c, sync_copies(c));
    dfree(c, c->d_pack, c->pack_cap, true);
    c->d_pack = nullptr;
    c->pack_cap = 0;
    if (!bytes) return GW_OK;
    // each copy lane gets an equal part, rounded up to 256 B
    const uint64_t per = ((bytes + c->lanes - 1) / c->lanes + 255) & ~255ULL;
    bytes = per * c->lanes;
    void* p = nullptr;
    gw_status st;
    if ((st = dalloc(c, &p, bytes, true)) != GW_OK) return st;
    c->d_pack = (uint32_t*)p;
    c->pack_cap = bytes;
    return GW_OK;
}

gw_status gw_set_copy_lanes(gw_ctx* c, uint32_t lanes)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (lanes < 1 || lanes > 2) return fail(c, GW_ERR_ARG, "copy lanes must be 1 or 2 (got %u)", lanes);
    (void)hipSetDevice(c->device);
    HIPCHK(c, sync_copies(c));
    if (lanes == 2 && !c->s_copy2 && hipStreamCreateWithFlags(&c->s_copy2, hipStreamNonBlocking) != hipSuccess) {
        c->s_copy2 = nullptr;
        return fail(c, GW_ERR_HIP, "stream creation failed");
    }
    c->lanes = lanes;
    c->next_lane = 0;
    return GW_OK;
}

static void make_resident(gw_ctx* c, uint32_t slot, uint32_t blk)
{
    c->slot_blk[slot] = (int32_t)blk;
    c->blk_slot[blk] = (int32_t)slot;
    c->slot_wait[slot] = 1;   // the next run's kernels wait for the copy (device-side)
    c->slot_read[slot] = 1;
    DevBlock& h = c->hblk[blk];
    h.beg = c->slot_beg[slot];
    h.csr = slot_edges(c, slot, blk);
    h.resident = kInSlot;
}

static gw_status check_slot_blk(gw_ctx* c, uint32_t slot, uint32_t blk)
{
    if (slot >= c->nslots) return fail(c, GW_ERR_ARG, "slot %u out of range (%u slots)", slot, c->nslots);
    if (blk >= c->nblocks) return fail(c, GW_ERR_ARG, "block %u out of range (%u blocks)", blk, c->nblocks);
    return GW_OK;
}

gw_status gw_upload_block(gw_ctx* c, uint32_t slot, uint32_t blk, const uint64_t* beg, const uint32_t* csr)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    gw_status st;
    if ((st = check_slot_blk(c, slot, blk)) != GW_OK) return st;
    if (c->blk_slot[blk] >= 0) {   // moving a resident block: its old slot becomes free
        c->slot_blk[c->blk_slot[blk]] = -1;
        block_nonresident(c, blk);
    }
    if ((st = start_copy(c, slot, blk, beg, csr)) != GW_OK) return st;
    make_resident(c, slot, blk);
    return GW_OK;
}

gw_status gw_upload_block_packed(gw_ctx* c, uint32_t slot, uint32_t blk, const void* packed, uint64_t packed_bytes)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (!packed) return fail(c, GW_ERR_ARG, "null packed block");
    gw_status st;
    if ((st = check_slot_blk(c, slot, blk)) != GW_OK) return st;
    if (c->blk_slot[blk] >= 0) {
        c->slot_blk[c->blk_slot[blk]] = -1;
        block_nonresident(c, blk);
    }
    if ((st = start_copy(c, slot, blk, nullptr, nullptr, packed, packed_bytes)) != GW_OK) return st;
    make_resident(c, slot, blk);
    return GW_OK;
}

gw_status gw_stage_block_packed(gw_ctx* c, uint32_t slot, uint32_t blk, const void* packed, uint64_t packed_bytes)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (!packed) return fail(c, GW_ERR_ARG, "null packed block");
    gw_status st;
    if ((st = check_slot_blk(c, slot, blk)) != GW_OK) return st;
    if (c->blk_slot[blk] >= 0) return fail(c, GW_ERR_ARG, "block %u is already resident (slot %d)", blk, c->blk_slot[blk]);
    for (uint32_t s = 0; s < c->nslots; ++s)
        if (s != slot && c->slot_staged[s] == (int32_t)blk) c->slot_staged[s] = -1;
    if ((st = start_copy(c, slot, blk, nullptr, nullptr, packed, packed_bytes)) != GW_OK) return st;
    c->slot_staged[slot] = (int32_t)blk;
    return GW_OK;
}

gw_status gw_stage_block(gw_ctx* c, uint32_t slot, uint32_t blk, const uint64_t* beg, const uint32_t* csr)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    gw_status st;
    if ((st = check_slot_blk(c, slot, blk)) != GW_OK) return st;
    if (c->blk_slot[blk] >= 0) return fail(c, GW_ERR_ARG, "block %u is already resident (slot %d)", blk, c->blk_slot[blk]);
    for (uint32_t s = 0; s < c->nslots; ++s)
        if (s != slot && c->slot_staged[s] == (int32_t)blk) c->slot_staged[s] = -1;   // newest staging wins
    if ((st = start_copy(c, slot, blk, beg, csr)) != GW_OK) return st;
    c->slot_staged[slot] = (int32_t)blk;
    return GW_OK;
}

gw_status gw_commit_block(gw_ctx* c, uint32_t slot)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (slot >= c->nslots) return fail(c, GW_ERR_ARG, "slot %u out of range (%u slots)", slot, c->nslots);
    const int32_t blk = c->slot_staged[slot];
    if (blk < 0) return fail(c, GW_ERR_STATE, "slot %u holds no staged block", slot);
    if (c->blk_slot[blk] >= 0) return fail(c, GW_ERR_STATE, "block %d became resident in slot %d meanwhile", blk, c->blk_slot[blk]);
    c->slot_staged[slot] = -1;
    make_resident(c, slot, (uint32_t)blk);
    return GW_OK;
}

gw_status gw_staged_block(gw_ctx* c, uint32_t slot, int32_t* blk, uint32_t* ready)
{
    if (!c || !blk) return fail(c, GW_ERR_ARG, "null argument");
    if (slot >= c->nslots) return fail(c, GW_ERR_ARG, "slot %u out of range (%u slots)", slot, c->nslots);
    *blk = c->slot_staged[slot];
    if (ready) {
        if (*blk >= 0 && !enqueued(c, slot)) {   // the worker has not even enqueued it
            *ready = 0;
            return worker_failed(c);
        }
        const hipError_t e = *blk >= 0 ? hipEventQuery(c->slot_ev1[slot]) : hipSuccess;
        if (e != hipSuccess && e != hipErrorNotReady) return fail(c, GW_ERR_HIP, "copy event: %s", hipGetErrorString(e));
        *ready = e == hipSuccess ? 1u : 0u;
    }
    return GW_OK;
}

gw_status gw_read_slot(gw_ctx* c, uint32_t slot, uint64_t* beg, uint32_t* csr)
{
    if (!c || !beg) return fail(c, GW_ERR_ARG, "null argument");
    if (slot >= c->nslots) return fail(c, GW_ERR_ARG, "slot %u out of range (%u slots)", slot, c->nslots);
    const int32_t blk = c->slot_blk[slot] >= 0 ? c->slot_blk[slot] : c->slot_staged[slot];
    if (blk < 0) return fail(c, GW_ERR_ARG, "slot %u holds no block", slot);
    const gw_block_desc& d = c->desc[blk];
    if (d.nedges && !csr) return fail(c, GW_ERR_ARG, "null argument");
    (void)hipSetDevice(c->device);
    HIPCHK(c, sync_copies(c));
    HIPCHK(c, hipMemcpy(beg, c->slot_beg[slot], ((uint64_t)d.nverts + 1) * 8, hipMemcpyDeviceToHost));
    if (d.nedges)
        HIPCHK(c, hipMemcpy(csr, slot_edges(c, slot, (uint32_t)blk), d.nedges * 4, hipMemcpyDeviceToHost));
    return GW_OK;
}

gw_status gw_evict(gw_ctx* c, uint32_t slot)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (slot >= c->nslots) return fail(c, GW_ERR_ARG, "slot %u out of range", slot);
    if (c->slot_blk[slot] >= 0) block_nonresident(c, (uint32_t)c->slot_blk[slot]);
    c->slot_blk[slot] = -1;
    c->slot_staged[slot] = -1;
    return GW_OK;
}

gw_status gw_wait_uploads(gw_ctx* c)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    HIPCHK(c, sync_copies(c));
    for (uint32_t s = 0; s < c->nslots; ++s) collect_copy_time(c, s, true);
    return worker_failed(c);
}

gw_status gw_map_host_block(gw_ctx* c, uint32_t blk, const uint64_t* beg, const uint32_t* csr)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (blk >= c->nblocks) return fail(c, GW_ERR_ARG, "block %u out of range", blk);
    void* db = nullptr;
    void* dc = nullptr;
    HIPCHK(c, hipHostGetDevicePointer(&db, (void*)beg, 0));
    if (c->desc[blk].nedges) HIPCHK(c, hipHostGetDevicePointer(&dc, (void*)csr, 0));
    c->map_beg[blk] = (const uint64_t*)db;
    c->map_csr[blk] = (const uint32_t*)dc;
    if (c->blk_slot[blk] < 0) block_nonresident(c, blk);
    return GW_OK;
}

gw_status gw_unmap_host_blocks(gw_ctx* c)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    HIPCHK(c, hipStreamSynchronize(c->s_compute));
    for (uint32_t i = 0; i < c->nblocks; ++i) {
        c->map_beg[i] = nullptr;
        c->map_csr[i] = nullptr;
        if (c->blk_slot[i] < 0) block_nonresident(c, i);
    }
    return GW_OK;
}

static uint64_t thr(double w, double m)
{
    if (w >= m) return 1ULL << 32;
    return (uint64_t)std::floor(w / m * 4294967296.0);
}

}  // extern "C"

namespace gw {
// The integer form of one walker type (acceptance thresholds, fixed-point alpha,
// termination threshold); nullptr on success, else what is wrong with it.
const char* dev_type_from(const gw_walk_type& ty, DevType* out)
{
    static thread_local char msg[160];
    DevType d{};
    if (ty.algo > 1) { snprintf(msg, sizeof msg, "unknown algorithm %u", ty.algo); return msg; }
    if (ty.length > kMaxLength) { snprintf(msg, sizeof msg, "length %u exceeds %u", ty.length, kMaxLength); return msg; }
    if (!(ty.term >= 0.0 && ty.term < 1.0)) return "term must be in [0, 1)";
    d.algo = ty.algo;
    d.length = ty.length;
    d.t_term = ty.term > 0.0 ? (uint64_t)std::floor(ty.term * 4294967296.0) : 0;
    if (ty.algo == GW_NODE2VEC) {
        if (!(ty.p >= 1.0 / 65536 && ty.p <= 65536 && ty.q >= 1.0 / 65536 && ty.q <= 65536))
            return "p and q must lie in [2^-16, 2^16]";
        const double ip = 1.0 / ty.p, iq = 1.0 / ty.q;
        const double m = std::max(std::max(ip, 1.0), iq);
        d.t_p = thr(ip, m);
        d.t_one = thr(1.0, m);
        d.t_q = thr(iq, m);
        d.t_min = std::min(std::min(d.t_p, d.t_one), d.t_q);
        if (d.t_min == 0) return "a Node2vec acceptance threshold rounds to zero";
        d.m_p = (uint32_t)(d.t_p - 1);
        d.m_one = (uint32_t)(d.t_one - 1);
        d.m_q = (uint32_t)(d.t_q - 1);
        d.m_min = (uint32_t)(d.t_min - 1);
        d.m_lo = (uint32_t)(std::min(d.t_one, d.t_q) - 1);
        d.m_hi = (uint32_t)(std::max(d.t_one, d.t_q) - 1);
    } else {
        if (!(ty.alpha >= 0.0 && ty.alpha < 1.0)) return "alpha must be in [0, 1)";
        const long long a = std::llround(ty.alpha * 65536.0);
        if (a >= 65536) return "alpha rounds to 1 in 16-bit fixed point";
        d.alpha_a = (uint32_t)a;
        d.alpha_b = 65536u - (uint32_t)a;
        d.m_prop = (uint32_t)(((uint64_t)d.alpha_b << 16) - 1);
        if (ty.sopr_sampler > 1) { snprintf(msg, sizeof msg, "unknown SOPR sampler %u", ty.sopr_sampler); return msg; }
        d.sampler = ty.sopr_sampler;
    }
    *out = d;
    return nullptr;
}
}  // namespace gw

extern "C" {

gw_status gw_set_walk(gw_ctx* c, const gw_walk_params* wp)
{
    if (!c || !wp) return fail(c, GW_ERR_ARG, "null argument");
    if (wp->ntypes < 1 || wp->ntypes > 2) return fail(c, GW_ERR_ARG, "ntypes must be 1 or 2");
    for (uint32_t t = 0; t < wp->ntypes; ++t)
        if (const char* e = dev_type_from(wp->type[t], &c->dt[t])) return fail(c, GW_ERR_ARG, "%s", e);
    if (wp->ntypes == 1) c->dt[1] = c->dt[0];
    c->wp = *wp;
    c->walk_set = true;
    return GW_OK;
}

// Worklist entries for n walkers: the caller's gw_set_worklist_cap, else n / 8 (at least
// min(n, 2^20)).  A run never needs more than its runnable walkers, and one whose runnable
// walkers exceed the worklist takes several filter + step rounds, so the 4 bytes per walker
// a full-size list would hold go to block slots instead (C5: 0.94 GB, 2-3 slots).
static uint64_t worklist_entries(const gw_ctx* c, uint64_t n)
{
    uint64_t w = c->work_cap_req ? c->work_cap_req : std::max<uint64_t>(n / 8, std::min<uint64_t>(n, 1u << 20));
    return std::min<uint64_t>(std::max<uint64_t>(w, 64), std::max<uint64_t>(n, 1));
}

gw_status gw_set_worklist_cap(gw_ctx* c, uint64_t entries)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (entries > 0xFFFFFFFFull) return fail(c, GW_ERR_ARG, "worklist of %llu entries exceeds 2^32 - 1",
                                             (unsigned long long)entries);
    c->work_cap_req = entries;
    return GW_OK;
}

gw_status gw_worklist_stats(gw_ctx* c, uint64_t* capacity, uint64_t* extra_rounds)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (capacity) *capacity = c->work_cap;
    if (extra_rounds) *extra_rounds = c->work_rounds;
    return GW_OK;
}

gw_status gw_init_walkers(gw_ctx* c, uint64_t n, uint32_t wpv, const uint32_t* sources, uint32_t outputs)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (!c->nblocks) return fail(c, GW_ERR_STATE, "gw_init_walkers before gw_set_block_table");
    if (!c->walk_set) return fail(c, GW_ERR_STATE, "gw_init_walkers before gw_set_walk");
    if (n == 0 || n > 0xFFFFFFFFull) return fail(c, GW_ERR_ARG, "walker count must be in [1, 2^32)");
    if (!sources) {
        if (wpv == 0) return fail(c, GW_ERR_ARG, "walkers_per_vertex must be >= 1");
        if ((n - 1) / wpv >= c->nverts_total)
            return fail(c, GW_ERR_ARG, "%llu walkers at %u per vertex exceed %llu vertices", (unsigned long long)n, wpv,
                        (unsigned long long)c->nverts_total);
    } else {
        for (uint64_t i = 0; i < n; ++i)
            if (sources[i] >= c->nverts_total)
                return fail(c, GW_ERR_ARG, "source %u of walker %llu is not a vertex", sources[i], (unsigned long long)i);
    }
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->s_compute));
    const uint32_t maxlen = std::max(c->dt[0].length, c->wp.ntypes == 2 ? c->dt[1].length : 0u);
    const uint32_t stride = (outputs & GW_OUT_PATHS) ? maxlen + 1 : 0;
    const uint64_t wcap = worklist_entries(c, n);
    // a re-initialisation with the same shape reuses the device buffers
    if (!(c->nwalkers == n && c->outputs == outputs && c->path_stride == stride && c->work_cap == wcap)) {
        free_walkers(c);
        c->path_stride = stride;
        void* p;
        gw_status st;
        c->nwalkers = n;
        uint32_t** u32s[] = {&c->d_cur, &c->d_prev, &c->d_meta};
        for (uint32_t** q : u32s) {
            if ((st = dalloc(c, &p, n * 4, true)) != GW_OK) { free_walkers(c); return st; }
            *q = (uint32_t*)p;
        }
        if ((st = dalloc(c, &p, wcap * 4, true)) != GW_OK) { free_walkers(c); return st; }
        c->d_work = (uint32_t*)p;
        c->work_cap = wcap;
        if (outputs & GW_OUT_HASH) {
            if ((st = dalloc(c, &p, n * 8, true)) != GW_OK) { free_walkers(c); return st; }
            c->d_hash = (uint64_t*)p;
        }
        if (outputs & GW_OUT_PATHS) {
            if ((st = dalloc(c, &p, n * (uint64_t)stride * 4, true)) != GW_OK) { free_walkers(c); return st; }
            c->d_paths = (uint32_t*)p;
        }
        if (outputs & GW_OUT_VISITS) {
            // one count array per XCD when that costs at most 1/16 of the budget (or there is
            // no budget): the step kernels' arrival counts then stay in one XCD's L2
            const uint64_t one = c->nverts_total * 4;
            c->visit_parts = (!c->cap || kVisitParts * one <= c->cap / 16) ? kVisitParts : 1u;
            if ((st = dalloc(c, &p, one * c->visit_parts, true)) != GW_OK) { free_walkers(c); return st; }
            c->d_visits = (uint32_t*)p;
        }
    }
    if (c->d_visits) HIPCHK(c, launch_fill_u32(c->d_visits, 0, c->nverts_total * c->visit_parts, c->s_compute));
    c->outputs = outputs;
    uint32_t* d_src = nullptr;
    if (sources) {
        HIPCHK(c, hipMalloc((void**)&d_src, n * 4));
        HIPCHK(c, hipMemcpy(d_src, sources, n * 4, hipMemcpyHostToDevice));
    }
    InitArgs a{};
    a.ntypes = c->wp.ntypes;
    a.len0 = c->dt[0].length;
    a.len1 = c->dt[1].length;
    a.nwalkers = n;
    a.wpv = wpv ? wpv : 1;
    a.sources = d_src;
    a.cur = c->d_cur; a.prev = c->d_prev; a.meta = c->d_meta;
    a.hash = c->d_hash; a.paths = c->d_paths; a.path_stride = c->path_stride;
    hipError_t e = launch_init(a, c->s_compute);
    if (e == hipSuccess) e = hipStreamSynchronize(c->s_compute);
    if (d_src) (void)hipFree(d_src);
    if (e != hipSuccess) return fail(c, GW_ERR_HIP, "walker init kernel failed: %s", hipGetErrorString(e));
    // walkers with a zero hop cap are done at once; the bucket matrix is computed on demand
    const uint64_t n1 = c->wp.ntypes == 2 ? n / 2 : 0, n0 = n - n1;
    const uint64_t l0 = c->dt[0].length, l1 = c->wp.ntypes == 2 ? c->dt[1].length : 0;
    c->unfinished = (l0 ? n0 : 0) + (l1 ? n1 : 0);
    c->ever_res.assign(c->nblocks, 0);
    c->pending_hops = n0 * l0 + n1 * l1;
    c->w_valid = false;
    return GW_OK;
}

// Worklist + (optional) bucket pass on the compute stream; returns nwork and left.
// Enqueues the bucket pass over all walkers on the compute stream: the matrix of
// (prev_blk, cur_blk) counts of unfinished walkers into d_bucket and, on the host, into
// the pinned h_bucket (valid after the stream is synchronized).
static gw_status enqueue_bucket_pass(gw_ctx* c, bool stamps = false)
{
    const uint64_t nn = (uint64_t)c->nblocks * c->nblocks;
    HIPCHK(c, launch_fill_u32((uint32_t*)(c->d_cnt + 6), 0, 2, c->s_compute));
    HIPCHK(c, launch_fill_u32((uint32_t*)(c->d_cnt + 8), 0, 2, c->s_compute));
    HIPCHK(c, launch_fill_u32(c->d_bucket, 0, nn, c->s_compute));
    FilterArgs f{};
    f.blocks = c->d_blocks;
    f.nblocks = c->nblocks;
    f.nwalkers = c->nwalkers;
    f.meta = c->d_meta;
    f.cur = c->d_cur;
    f.prev = c->d_prev;
    f.vcut = c->d_vcut;
    f.bucket = c->d_bucket;
    f.left = c->d_cnt + 6;
    f.hops = c->d_cnt + 8;
    f.len0 = c->dt[0].length;
    f.len1 = c->dt[1].length;
    f.bucket_part = c->d_bpart;
    f.part_rows = c->d_bpart ? c->ncu : 0;
    HIPCHK(c, stamps ? launch_timestamp(c->m_ts + 2, c->s_compute) : hipEventRecord(c->ev_f0, c->s_compute));
    HIPCHK(c, launch_filter(f, c->ncu, c->s_compute));
    HIPCHK(c, stamps ? launch_timestamp(c->m_ts + 3, c->s_compute) : hipEventRecord(c->ev_f1, c->s_compute));
    HIPCHK(c, launch_copy_u32(c->m_bucket, c->d_bucket, nn, c->s_compute));
    HIPCHK(c, launch_copy_u32((uint32_t*)(c->m_cnt + 8), (const uint32_t*)(c->d_cnt + 6), 2, c->s_compute));
    HIPCHK(c, launch_copy_u32((uint32_t*)(c->m_cnt + 9), (const uint32_t*)(c->d_cnt + 8), 2, c->s_compute));
    return GW_OK;
}

// Makes the cached bucket matrix current (one pass unless the last run left it valid).
static gw_status ensure_buckets(gw_ctx* c)
{
    if (c->w_valid) return GW_OK;
    gw_status st;
    if ((st = push_block_table(c)) != GW_OK) return st;
    if ((st = enqueue_bucket_pass(c)) != GW_OK) return st;
    HIPCHK(c, hipStreamSynchronize(c->s_compute));
    c->unfinished = c->h_cnt[8];
    c->pending_hops = c->h_cnt[9];
    c->w_valid = true;
    return GW_OK;
}

// The compute stream's waits for the copies of newly resident slots, and the block table
// written into pinned memory; the copy kernel that pushes it to the device is enqueued
// unless `enqueue_copy` is false (a run captures it into its graph).
static gw_status push_block_table(gw_ctx* c, bool enqueue_copy)
{
    // kernels read resident slots only after their copies (staged copies are not waited)
    gw_status st;
    for (uint32_t s = 0; s < c->nslots; ++s)
        if (c->slot_wait[s]) {
            wait_enqueued(c, s);
            if ((st = worker_failed(c)) != GW_OK) return st;
            HIPCHK(c, hipStreamWaitEvent(c->s_compute, c->slot_ev1[s], 0));
            c->slot_wait[s] = 0;
        }
    // the previous push has been consumed: every caller synchronizes the compute stream
    // before it returns
    memcpy(c->h_blocks_pinned, c->hblk.data(), sizeof(DevBlock) * c->nblocks);
    static_assert(sizeof(DevBlock) % 4 == 0, "block descriptors are copied as words");
    if (enqueue_copy)
        HIPCHK(c, launch_copy_u32((uint32_t*)c->d_blocks, (const uint32_t*)c->m_blocks,
                                  sizeof(DevBlock) / 4 * c->nblocks, c->s_compute));
    return GW_OK;
}

static gw_status run_impl(gw_ctx* c, gw_launch l, gw_run_stats* stats, uint64_t cap);

constexpr size_t kMaxGraphs = 8;

gw_status gw_upload_stats(gw_ctx* c, uint64_t* inplace, uint64_t* ring, uint64_t* copy_stream)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    wait_enqueued(c, -1);
    std::lock_guard<std::mutex> lk(c->wmu);
    if (inplace) *inplace = c->inplace_blocks;
    if (ring) *ring = c->ring_blocks;
    if (copy_stream) *copy_stream = c->inline_blocks;
    return GW_OK;
}

gw_status gw_graph_stats(gw_ctx* c, uint64_t* replays, uint64_t* captures)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (replays) *replays = c->graph_runs;
    if (captures) *captures = c->graph_captures;
    return GW_OK;
}

static hipGraphExec_t graph_lookup(gw_ctx* c, const RunSig& sig)
{
    for (auto& g : c->graphs)
        if (!memcmp(&g.sig, &sig, sizeof sig)) {
            g.used = ++c->graph_clock;
            return g.ex;
        }
    return nullptr;
}

// keeps the kMaxGraphs most recently used graphs
static void graph_store(gw_ctx* c, const RunSig& sig, hipGraphExec_t ex)
{
    c->graph_captures++;
    if (c->graphs.size() >= kMaxGraphs) {
        auto old = std::min_element(c->graphs.begin(), c->graphs.end(),
                                    [](const gw_ctx::GraphEntry& x, const gw_ctx::GraphEntry& y) { return x.used < y.used; });
        (void)hipGraphExecDestroy(old->ex);
        c->graphs.erase(old);
    }
    c->graphs.push_back({sig, ex, ++c->graph_clock});
}

static void graphs_clear(gw_ctx* c)
{
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.ex);
    c->graphs.clear();
}

gw_status gw_run(gw_ctx* c, const gw_launch* L, gw_run_stats* stats)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (!c->nwalkers) return fail(c, GW_ERR_STATE, "gw_run before gw_init_walkers");
    (void)hipSetDevice(c->device);
    gw_launch l = L ? *L : gw_launch{};
    if (l.block == 0) {   // shape and search order chosen by gw_autotune, else the default
        const uint32_t flags = l.flags;
        l = c->have_tuned ? c->tuned : gw_launch{256, 0, 0, 0};
        l.flags = flags | (l.flags & (GW_RUN_INTERP | GW_RUN_WINDOW));
    }
    return run_impl(c, l, stats, 0);
}

// cap > 0 limits the run to the first `cap` work items (used by the autotuner)
static gw_status run_impl(gw_ctx* c, gw_launch l, gw_run_stats* stats, uint64_t cap)
{
    if (l.block % 64 || l.block > 1024) return fail(c, GW_ERR_ARG, "block must be a multiple of 64 up to 1024");
    if (l.variant > 2) return fail(c, GW_ERR_ARG, "unknown kernel variant %u", l.variant);
    if (l.variant == 1 && l.block > 256) return fail(c, GW_ERR_ARG, "the flattened kernel runs at most 256 threads per workgroup");

    gw_status st;
    if ((st = push_block_table(c, false)) != GW_OK) return st;
    const bool w_was_valid = c->w_valid;
    c->w_valid = false;
    // The whole walker array is the work list: the kernel skips finished walkers and,
    // unless every block is readable, walkers whose (prev, cur) blocks are not both
    // resident -- a coalesced 8-byte read per walker instead of a compaction pass and a
    // host round trip.
    bool all_res = true, mapped = false;
    for (uint32_t b = 0; b < c->nblocks; ++b) {
        all_res = all_res && c->hblk[b].resident != kNotResident;
        mapped = mapped || c->hblk[b].resident == kMapped;
    }
    uint64_t nwork = c->nwalkers;
    if (cap && nwork > cap) nwork = cap;
    if (!all_res && nwork > c->work_cap) nwork = c->work_cap;   // the step kernel reads at most this many
    float fms = 0.f, kms = 0.f;
    // compact the runnable walkers on the device; the step kernel reads the count there,
    // so no host round trip separates the two kernels
    FilterArgs f;
    memset(&f, 0, sizeof f);   // the arguments are compared bytewise (graph cache)
    f.blocks = c->d_blocks;
    f.nblocks = c->nblocks;
    f.nwalkers = cap && cap < c->nwalkers ? cap : c->nwalkers;
    f.meta = c->d_meta;
    f.cur = c->d_cur;
    f.prev = c->d_prev;
    f.vcut = c->d_vcut;
    f.worklist = c->d_work;
    f.work_cap = (uint32_t)c->work_cap;
    f.nwork = c->d_misc;
    f.left = c->d_cnt + 6;
    f.len0 = c->dt[0].length;
    f.len1 = c->dt[1].length;
    WalkArgs a;
    memset(&a, 0, sizeof a);
    a.blocks = c->d_blocks;
    a.vcut = c->d_vcut;
    a.nblocks = c->nblocks;
    a.ntypes = c->wp.ntypes;
    a.seed_lo = (uint32_t)c->wp.seed;
    a.seed_hi = (uint32_t)(c->wp.seed >> 32);
    a.type0 = c->dt[0];
    a.type1 = c->dt[1];
    a.cur = c->d_cur; a.prev = c->d_prev; a.meta = c->d_meta;
    a.hash = c->d_hash; a.paths = c->d_paths; a.path_stride = c->path_stride; a.visits = c->d_visits;
    a.visits_stride = c->nverts_total;
    a.visits_parts = c->d_visits ? c->visit_parts : 1u;
    a.worklist = all_res ? nullptr : c->d_work;
    a.nwork_dev = all_res ? nullptr : c->d_misc;
    a.zc_count = mapped ? 1u : 0u;
    a.search_end = (l.flags & GW_RUN_INTERP) ? (uint32_t)c->nverts_total : 0u;
    a.search_win = (l.flags & GW_RUN_WINDOW) ? 1u : 0u;
    a.nwork = (uint32_t)nwork;
    a.queue = c->d_misc + 1;
    a.counters = c->d_cnt;
    if (l.flags & GW_RUN_TRACE) a.block_steps = c->d_trace;
    // the flattened variant reads u32s through aligned 8-byte words, which is only
    // safe inside padded HBM slots: zero-copy (host-mapped) blocks use variant 0
    if (l.variant == 1 && (l.flags & (GW_RUN_INTERP | GW_RUN_WINDOW))) l.variant = 0;   // bisection only in variant 1
    if (l.variant == 1)
        for (uint32_t b = 0; b < c->nblocks; ++b)
            if (c->hblk[b].resident == kMapped) { l.variant = 0; break; }
    // the lane kernel covers Node2vec and the SOPR proposal sampler on 32-bit degrees
    if (l.variant == 2) {
        for (uint32_t t = 0; t < 2; ++t)
            if (c->dt[t].algo == 1 && c->dt[t].sampler != 0) l.variant = 0;
        for (uint32_t b = 0; b < c->nblocks; ++b)
            if (c->hblk[b].nedges >> 32) l.variant = 0;
    }
    // the lane kernel steps walkers whose previous block is not resident until a step needs
    // that block's list (it stops there and waits for it); the other variants read both rows
    // at pickup, so their worklist keeps walkers with both blocks resident.  Such walkers
    // exist only once a block the walk has had resident is gone (walkers start with prev =
    // cur; a step leaves prev in a resident block): before that, the lane kernel instance
    // without those checks runs (C3 16 GB: its step kernel is ~4% faster)
    bool gone = false;
    if (c->ever_res.size() != c->nblocks) c->ever_res.assign(c->nblocks, 0);
    for (uint32_t b = 0; b < c->nblocks; ++b) {
        if (c->hblk[b].resident != kNotResident) c->ever_res[b] = 1;
        else if (c->ever_res[b]) gone = true;
    }
    f.relax_prev = a.relax = (l.variant == 2 && gone) ? 1u : 0u;
    // persistent grid: as many workgroups as the CUs hold at once (occupancy), or the
    // caller's waves per CU
    const bool one_blk = c->nblocks == 1;
    uint64_t grid = l.waves_per_cu ? (uint64_t)c->ncu * l.waves_per_cu * 64 / l.block
                                   : (uint64_t)c->ncu * walk_blocks_per_cu(one_blk, c->nblocks, l.variant, l.block);
    const uint64_t need = (nwork + l.block - 1) / l.block;   // an upper bound with a device-side count
    if (grid > need) grid = need;
    if (grid == 0) grid = 1;
    {
        const uint64_t waves = grid * (l.block / 64);
        uint64_t ch = nwork / (waves * 8 + 1);
        ch = ch < 64 ? 64 : (ch > 1024 ? 1024 : ch);
        a.chunk = (uint32_t)(ch / 64 * 64);
    }
    // the trace comes back with the counters, so gw_trace_read needs no device round trip
    const bool traced = a.block_steps != nullptr;
    const bool buckets = (l.flags & GW_RUN_BUCKETS) != 0;
    // With a worklist and the matrix current when the run starts, the matrix is kept current
    // instead of recounted over every walker: the filter takes the walkers it lists out of
    // it, and a pass over that list after the step kernel puts the unfinished ones back
    // with their new (prev, cur) blocks (C5, 268M walkers: the recount took ~1 ms a run,
    // on the compute stream between the step kernel and the next schedule)
    const bool delta = buckets && w_was_valid && !all_res && !cap;
    FilterArgs g;
    memset(&g, 0, sizeof g);
    if (delta) {
        f.bucket_sub = c->d_bucket;
        f.taken = c->d_cnt + 10;
        f.bucket_part = c->d_bpart;
        f.part_rows = c->d_bpart ? 2 * c->ncu : 0;
        g.blocks = c->d_blocks;
        g.nblocks = c->nblocks;
        g.nwalkers = nwork;   // the step kernel's bound on the list
        g.meta = c->d_meta;
        g.cur = c->d_cur;
        g.prev = c->d_prev;
        g.vcut = c->d_vcut;
        g.bucket = c->d_bucket;
        g.left = c->d_cnt + 12;
        g.hops = c->d_cnt + 13;
        g.len0 = c->dt[0].length;
        g.len1 = c->dt[1].length;
        g.bucket_part = c->d_bpart;
        g.part_rows = c->d_bpart ? c->ncu : 0;
        g.gather = c->d_work;
        g.gather_n = c->d_misc;
    }
    c->trace_valid = false;
    // the run's device work: the block table push, counter clears, filter, step kernel,
    // readbacks and the bucket pass -- launched one by one, or as a graph captured the
    // first time this exact sequence (every kernel argument equal) runs and replayed after
    const bool graph = c->run_graphs && !cap;
    uint32_t round = 0;   // > 0: a further round for walkers the worklist could not hold
    auto enqueue = [&]() -> gw_status {
        HIPCHK(c, launch_copy_u32((uint32_t*)c->d_blocks, (const uint32_t*)c->m_blocks,
                                  sizeof(DevBlock) / 4 * c->nblocks, c->s_compute));
        HIPCHK(c, launch_fill_u32((uint32_t*)c->d_cnt, 0, 16, c->s_compute));
        HIPCHK(c, launch_fill_u32(c->d_misc, 0, 4, c->s_compute));
        if (delta) HIPCHK(c, launch_fill_u32((uint32_t*)(c->d_cnt + 10), 0, 8, c->s_compute));
        if (!all_res) HIPCHK(c, launch_filter(f, c->ncu * 2, c->s_compute));
        if (traced && round == 0) HIPCHK(c, launch_fill_u32((uint32_t*)c->d_trace, 0, 2ull * c->nblocks, c->s_compute));
        HIPCHK(c, graph ? launch_timestamp(c->m_ts, c->s_compute) : hipEventRecord(c->ev_k0, c->s_compute));
        HIPCHK(c, launch_walk(a, one_blk, l.variant, (uint32_t)grid, l.block, c->s_compute));
        HIPCHK(c, graph ? launch_timestamp(c->m_ts + 1, c->s_compute) : hipEventRecord(c->ev_k1, c->s_compute));
        HIPCHK(c, launch_copy_u32((uint32_t*)c->m_cnt, (const uint32_t*)c->d_cnt, 16, c->s_compute));
        // the runnable count (beyond the worklist when it overflowed)
        if (!all_res) HIPCHK(c, launch_copy_u32((uint32_t*)(c->m_cnt + 10), c->d_misc, 1, c->s_compute));
        if (traced)
            HIPCHK(c, launch_copy_u32((uint32_t*)c->m_trace, (const uint32_t*)c->d_trace, 2ull * c->nblocks, c->s_compute));
        // bucket matrix of the walkers left, read back with the counters (one host wait)
        if (delta) {
            HIPCHK(c, graph ? launch_timestamp(c->m_ts + 2, c->s_compute) : hipEventRecord(c->ev_f0, c->s_compute));
            HIPCHK(c, launch_filter(g, c->ncu, c->s_compute));
            HIPCHK(c, graph ? launch_timestamp(c->m_ts + 3, c->s_compute) : hipEventRecord(c->ev_f1, c->s_compute));
            HIPCHK(c, launch_copy_u32(c->m_bucket, c->d_bucket, (uint64_t)c->nblocks * c->nblocks, c->s_compute));
            HIPCHK(c, launch_copy_u32((uint32_t*)(c->m_cnt + 11), (const uint32_t*)(c->d_cnt + 10), 8, c->s_compute));
            return GW_OK;
        }
        if (buckets) return enqueue_bucket_pass(c, graph);
        return GW_OK;
    };
    gw_run_stats s{};
    for (;; ++round) {
        if (graph) {
            RunSig sig;
            memset(&sig, 0, sizeof sig);
            sig.a = a;
            sig.f = all_res ? FilterArgs{} : f;
            sig.grid = grid;
            sig.block = l.block;
            sig.variant = l.variant;
            sig.flags = (all_res ? 1u : 0u) | (traced ? 2u : 0u) | (buckets ? 4u : 0u) | (one_blk ? 8u : 0u) |
                        (round ? 16u : 0u) | (delta ? 32u : 0u);
            sig.ncu = c->ncu;
            sig.d_bucket = c->d_bucket;
            sig.m_cnt = c->m_cnt;
            sig.m_trace = c->m_trace;
            sig.m_bucket = c->m_bucket;
            sig.m_blocks = c->m_blocks;
            hipGraphExec_t ex = graph_lookup(c, sig);
            if (!ex) {
                HIPCHK(c, hipStreamBeginCapture(c->s_compute, hipStreamCaptureModeThreadLocal));
                const gw_status est = enqueue();
                hipGraph_t g = nullptr;
                const hipError_t ce = hipStreamEndCapture(c->s_compute, &g);
                if (est != GW_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return est;
                }
                if (ce != hipSuccess) return fail(c, GW_ERR_HIP, "run graph capture: %s", hipGetErrorString(ce));
                const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ie != hipSuccess) return fail(c, GW_ERR_HIP, "run graph instantiation: %s", hipGetErrorString(ie));
                graph_store(c, sig, ex);
            }
            HIPCHK(c, hipGraphLaunch(ex, c->s_compute));
            c->graph_runs++;
        } else {
            if ((st = enqueue()) != GW_OK) return st;
        }
        HIPCHK(c, hipStreamSynchronize(c->s_compute));
        if (graph) {
            kms = (float)((double)(c->h_ts[1] - c->h_ts[0]) / c->ts_khz);
            if (buckets) fms = (float)((double)(c->h_ts[3] - c->h_ts[2]) / c->ts_khz);
        } else {
            HIPCHK(c, hipEventElapsedTime(&kms, c->ev_k0, c->ev_k1));
            if (buckets) HIPCHK(c, hipEventElapsedTime(&fms, c->ev_f0, c->ev_f1));
        }
        s.steps += c->h_cnt[0];
        s.trials += c->h_cnt[1];
        s.probes += c->h_cnt[2];
        s.mix_steps += c->h_cnt[3];
        s.walkers_ended += c->h_cnt[4];
        s.walkers_run += c->h_cnt[4] + c->h_cnt[5];
        s.kernel_ms += kms;
        s.filter_ms += fms;
        s.zc_touches += c->h_cnt[7];
        if (!buckets) c->unfinished -= std::min<uint64_t>(c->unfinished, c->h_cnt[4]);
        if (delta) {   // listed walkers out ([11] count, [12] hops), unfinished ones back ([13], [14])
            c->unfinished = c->unfinished - c->h_cnt[11] + c->h_cnt[13];
            c->pending_hops = c->pending_hops - c->h_cnt[12] + c->h_cnt[14];
        }
        // every walker of the list has run until it finished or left the resident blocks,
        // so the runnable count falls by the list's length each round
        if (all_res || (uint32_t)c->h_cnt[10] <= c->work_cap) break;
        c->work_rounds++;
    }
    c->trace_valid = traced;
    if (buckets) c->w_valid = true;
    for (uint32_t sl = 0; sl < c->nslots; ++sl) collect_copy_time(c, sl, false);
    if (buckets && !delta) {
        c->unfinished = c->h_cnt[8];
        c->pending_hops = c->h_cnt[9];
    }
    s.walkers_left = c->unfinished;
    c->tot.zc_touches += s.zc_touches;
    c->tot.steps += s.steps;
    c->tot.trials += s.trials;
    c->tot.probes += s.probes;
    c->tot.mix_steps += s.mix_steps;
    c->tot.kernel_ms += s.kernel_ms;
    if (stats) *stats = s;
    return GW_OK;
}

// Launch-shape autotuner: times the step kernel on the first `sample` work items for
// every candidate (kernel variant x workgroup size), restoring walker state, visit counts
// and counters after each trial, and keeps the fastest for later gw_run calls that pass
// block == 0.  Walks are bit-identical for every shape; only time changes.
gw_status gw_autotune(gw_ctx* c, uint64_t sample, uint32_t what, gw_launch* best, double* best_ms)
{
    if (!c) return fail(c, GW_ERR_ARG, "null context");
    if (!c->nwalkers) return fail(c, GW_ERR_STATE, "gw_autotune before gw_init_walkers");
    if (c->d_paths) return fail(c, GW_ERR_STATE, "gw_autotune cannot restore path output");
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->s_compute));
    const uint64_t n = c->nwalkers;
    struct Snap { void* dev; void* save; uint64_t bytes; };
    std::vector<Snap> snaps = {{c->d_cur, nullptr, n * 4}, {c->d_prev, nullptr, n * 4}, {c->d_meta, nullptr, n * 4}};
    if (c->d_hash) snaps.push_back({c->d_hash, nullptr, n * 8});
    if (c->d_visits) snaps.push_back({c->d_visits, nullptr, c->nverts_total * 4 * c->visit_parts});
    gw_status st = GW_OK;
    auto cleanup = [&]() { for (auto& x : snaps) if (x.save) (void)hipFree(x.save); };
    for (auto& x : snaps) {
        if (hipMalloc(&x.save, x.bytes) != hipSuccess) { cleanup(); return fail(c, GW_ERR_NOMEM, "autotune snapshot"); }
        if (hipMemcpyAsync(x.save, x.dev, x.bytes, hipMemcpyDeviceToDevice, c->s_compute) != hipSuccess) {
            cleanup();
            return fail(c, GW_ERR_HIP, "autotune snapshot copy");
        }
    }
    auto restore = [&]() -> bool {
        for (auto& x : snaps)
            if (hipMemcpyAsync(x.dev, x.save, x.bytes, hipMemcpyDeviceToDevice, c->s_compute) != hipSuccess) return false;
        return hipStreamSynchronize(c->s_compute) == hipSuccess;
    };
    const gw_counters tot0 = c->tot;
    const uint64_t unfinished0 = c->unfinished;
    const uint32_t I = GW_RUN_INTERP, W = GW_RUN_WINDOW;
    static const gw_launch cand[] = {{64, 0, 0, 0},   {128, 0, 0, 0},  {256, 0, 0, 0},  {512, 0, 0, 0},
                                     {1024, 0, 0, 0}, {64, 0, 1, 0},   {128, 0, 1, 0},  {256, 0, 1, 0},
                                     {64, 0, 0, I},   {128, 0, 0, I},  {256, 0, 0, I},  {512, 0, 0, I},
                                     {1024, 0, 0, I}, {64, 0, 2, 0},   {256, 0, 2, 0},  {1024, 0, 2, 0},
                                     {64, 0, 2, I},   {256, 0, 2, I},  {1024, 0, 2, I}, {256, 0, 0, W},
                                     {1024, 0, 0, W}, {256, 0, 2, W},  {1024, 0, 2, W}, {256, 0, 0, I | W},
                                     {1024, 0, 0, I | W}, {64, 0, 2, I | W}, {256, 0, 2, I | W}, {1024, 0, 2, I | W}};
    double bms = 1e300;
    gw_launch bl = cand[2];
    for (const gw_launch& L : cand) {
        // candidates outside `what` are skipped; 256 threads of variant 0 always run
        if (L.variant != 0 && !(what & GW_TUNE_VARIANT)) continue;
        if (L.variant == 0 && L.block != 256 && !(what & GW_TUNE_BLOCK)) continue;
        if ((L.flags & (GW_RUN_INTERP | GW_RUN_WINDOW)) && !(what & GW_TUNE_SEARCH)) continue;
        double t = 1e300;
        for (int rep = 0; rep < 2 && st == GW_OK; ++rep) {
            gw_run_stats rs{};
            if ((st = run_impl(c, L, &rs, sample)) != GW_OK) break;
            t = std::min(t, rs.kernel_ms);
            c->unfinished = unfinished0;
            if (!restore()) st = fail(c, GW_ERR_HIP, "autotune restore");
        }
        if (st != GW_OK) break;
        if (t < bms) { bms = t; bl = L; }
    }
    c->tot = tot0;
    c->unfinished = unfinished0;
    c->w_valid = false;
    cleanup();
    if (st != GW_OK) return st;
    c->tuned = bl;
    c->have_tuned = true;
    if (best) *best = bl;
    if (best_ms) *best_ms = bms;
    return GW_OK;
}

gw_status gw_bucket_counts(gw_ctx* c, uint64_t* W, uint64_t n)
{
    if (!c || !W) return fail(c, GW_ERR_ARG, "null argument");
    if (!c->nwalkers) return fail(c, GW_ERR_STATE, "no walkers");
    const uint64_t nn = (uint64_t)c->nblocks * c->nblocks;
    if (n < nn) return fail(c, GW_ERR_ARG, "bucket buffer holds %llu < %llu entries", (unsigned long long)n,
                            (unsigned long long)nn);
    (void)hipSetDevice(c->device);
    gw_status st;
    if ((st = ensure_buckets(c)) != GW_OK) return st;
    for (uint64_t i = 0; i < nn; ++i) W[i] = c->h_bucket[i];
    return GW_OK;
}

gw_status gw_trace_read(gw_ctx* c, uint64_t* block_steps, uint64_t n)
{
    if (!c || !block_steps) return fail(c, GW_ERR_ARG, "null argument");
    if (!c->nblocks) return fail(c, GW_ERR_STATE, "no block table");
    if (n < c->nblocks) return fail(c, GW_ERR_ARG, "trace buffer too small");
    if (c->trace_valid) {   // brought back with the traced run's counters
        std::memcpy(block_steps, c->h_trace, c->nblocks * 8ull);
        return GW_OK;
    }
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->s_compute));
    HIPCHK(c, d2h_small(c, block_steps, c->d_trace, c->nblocks * 8ull));
    return GW_OK;
}

gw_status gw_block_features(gw_ctx* c, const float* age, float* feats, uint64_t n)
{
    if (!c || !feats) return fail(c, GW_ERR_ARG, "null argument");
    if (!c->nwalkers) return fail(c, GW_ERR_STATE, "no walkers");
    if (n < (uint64_t)c->nblocks * kNumFeatures) return fail(c, GW_ERR_ARG, "feature buffer too small");
    (void)hipSetDevice(c->device);
    gw_status st;
    if ((st = ensure_buckets(c)) != GW_OK) return st;
    if ((st = push_block_table(c)) != GW_OK) return st;
    if (age) HIPCHK(c, h2d_small(c, c->d_age, age, c->nblocks * 4ull));
    const uint32_t maxlen = std::max(c->dt[0].length, c->wp.ntypes == 2 ? c->dt[1].length : 0u);
    const float ahead = c->unfinished && maxlen ? (float)((double)c->pending_hops / ((double)c->unfinished * maxlen)) : 0.f;
    HIPCHK(c, launch_features(c->d_bucket, c->nblocks, c->d_blocks, age ? c->d_age : nullptr, ahead, c->d_feat,
                              c->s_compute));
    HIPCHK(c, d2h_small(c, feats, c->d_feat, (uint64_t)c->nblocks * kNumFeatures * 4));
    return GW_OK;
}

gw_status gw_block_scores(gw_ctx* c, gw_mlp* m, const float* age, const float* x_mu, const float* x_sigma,
                          float* feats, float* scores, uint64_t n)
{
    if (!c || !m || !x_mu || !x_sigma || !feats || !scores) return fail(c, GW_ERR_ARG, "null argument");
    if (!c->nwalkers) return fail(c, GW_ERR_STATE, "no walkers");
    if (n < c->nblocks) return fail(c, GW_ERR_ARG, "buffers hold %llu blocks < %u", (unsigned long long)n, c->nblocks);
    if (mlp_num_inputs(m) != kNumFeatures) return fail(c, GW_ERR_ARG, "the network takes %u inputs, not %u",
                                                       mlp_num_inputs(m), kNumFeatures);
    (void)hipSetDevice(c->device);
    gw_status st;
    if ((st = ensure_buckets(c)) != GW_OK) return st;
    if ((st = push_block_table(c)) != GW_OK) return st;
    if (age) HIPCHK(c, h2d_small(c, c->d_age, age, c->nblocks * 4ull));
    float* musig = c->d_feat + (uint64_t)c->nblocks * kNumFeatures;
    HIPCHK(c, h2d_small(c, musig, x_mu, kNumFeatures * 4ull));
    HIPCHK(c, h2d_small(c, musig + kNumFeatures, x_sigma, kNumFeatures * 4ull));
    const uint32_t maxlen = std::max(c->dt[0].length, c->wp.ntypes == 2 ? c->dt[1].length : 0u);
    const float ahead = c->unfinished && maxlen ? (float)((double)c->pending_hops / ((double)c->unfinished * maxlen)) : 0.f;
    HIPCHK(c, launch_features(c->d_bucket, c->nblocks, c->d_blocks, age ? c->d_age : nullptr, ahead, c->d_feat,
                              c->s_compute));
    const float* dy = nullptr;
    if ((st = mlp_forward_device(c, m, c->d_feat, musig, c->nblocks, &dy)) != GW_OK) return st;
    // both results back through the pinned ring with one host wait
    const uint64_t fb = (uint64_t)c->nblocks * kNumFeatures * 4, sb = c->nblocks * 4ull;
    if (fb + sb + 512 > c->ring_cap / 2) {
        HIPCHK(c, d2h_small(c, feats, c->d_feat, fb));
        HIPCHK(c, d2h_small(c, scores, dy, sb));
        return GW_OK;
    }
    hipError_t e;
    const uint64_t at = ring_take(c, fb, &e);
    HIPCHK(c, e);
    const uint64_t at2 = ring_take(c, sb, &e);
    HIPCHK(c, e);
    HIPCHK(c, launch_copy_u32((uint32_t*)(c->m_ring + at), (const uint32_t*)c->d_feat, fb / 4, c->s_compute));
    HIPCHK(c, launch_copy_u32((uint32_t*)(c->m_ring + at2), (const uint32_t*)dy, sb / 4, c->s_compute));
    HIPCHK(c, hipStreamSynchronize(c->s_compute));
    memcpy(feats, c->h_ring + at, fb);
    memcpy(scores, c->h_ring + at2, sb);
    return GW_OK;
}

gw_status gw_pending_work(gw_ctx* c, uint64_t* walkers, uint64_t* hops)
{
    if (!c || !walkers || !hops) return fail(c, GW_ERR_ARG, "null argument");
    if (!c->nwalkers) return fail(c, GW_ERR_STATE, "no walkers");
    (void)hipSetDevice(c->device);
    gw_status st;
    if ((st = ensure_buckets(c)) != GW_OK) return st;
    *walkers = c->unfinished;
    *hops = c->pending_hops;
    return GW_OK;
}

gw_status gw_walkers_left(gw_ctx* c, uint64_t* left)
{
    if (!c || !left) return fail(c, GW_ERR_ARG, "null argument");
    if (!c->nwalkers) return fail(c, GW_ERR_STATE, "no walkers");
    (void)hipSetDevice(c->device);
    gw_status st;
    if ((st = ensure_buckets(c)) != GW_OK) return st;
    *left = c->unfinished;
    return GW_OK;
}

static gw_status readback(gw_ctx* c, void* dst, const void* src, uint64_t bytes, const char* what)
{
    if (!src) return fail(c, GW_ERR_STATE, "%s were not requested in gw_init_walkers", what);
    (void)hipSetDevice(c->device);
    HIPCHK(c, hipStreamSynchronize(c->s_compute));
    HIPCHK(c, hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GW_OK;
}

gw_status gw_read_visits(gw_ctx* c, uint32_t* out, uint64_t n)
{
    if (!c || !out) return fail(c, GW_ERR_ARG, "null argument");
    if (n < c->nverts_total) return fail(c, GW_ERR_ARG, "visit buffer too small");
    if (c->d_visits && c->visit_parts > 1) {
        (void)hipSetDevice(c->device);
        HIPCHK(c, launch_fold_visits(c->d_visits, c->nverts_total, c->visit_parts, c->s_compute));
    }
    return readback(c, out, c->d_visits, c->nverts_total * 4, "visits");
}

gw_status gw_read_hashes(gw_ctx* c, uint64_t* out, uint64_t n)
{
    if (!c || !out) return fail(c, GW_ERR_ARG, "null argument");
    if (n < c->nwalkers) return fail(c, GW_ERR_ARG, "hash buffer too small");
    return readback(c, out, c->d_hash, c->nwalkers * 8, "hashes");
}

gw_status gw_read_paths(gw_ctx* c, uint32_t* out, uint64_t n)
{
    if (!c || !out) return fail(c, GW_ERR_ARG, "null argument");
    if (n < c->nwalkers * (uint64_t)c->path_stride) return fail(c, GW_ERR_ARG, "path buffer too small");
    return readback(c, out, c->d_paths, c->nwalkers * (uint64_t)c->path_stride * 4, "paths");
}

gw_status gw_read_state(gw_ctx* c, uint32_t* out4, uint64_t n)
{
    if (!c || !out4) return fail(c, GW_ERR_ARG, "null argument");
    if (n < c->nwalkers) return fail(c, GW_ERR_ARG, "state buffer too small");
    std::vector<uint32_t> cur(c->nwalkers), prev(c->nwalkers), meta(c->nwalkers);
    gw_status st;
    if ((st = readback(c, cur.data(), c->d_cur, c->nwalkers * 4, "state")) != GW_OK) return st;
    if ((st = readback(c, prev.data(), c->d_prev, c->nwalkers * 4, "state")) != GW_OK) return st;
    if ((st = readback(c, meta.data(), c->d_meta, c->nwalkers * 4, "state")) != GW_OK) return st;
    for (uint64_t i = 0; i < c->nwalkers; ++i) {
        out4[4 * i + 0] = cur[i];
        out4[4 * i + 1] = prev[i];
        out4[4 * i + 2] = meta[i] & kHopMask;
        out4[4 * i + 3] = (meta[i] >> kStatusShift) & 0xFu;
    }
    return GW_OK;
}

gw_status gw_counters_get(gw_ctx* c, gw_counters* out)
{
    if (!c || !out) return fail(c, GW_ERR_ARG, "null argument");
    *out = c->tot;
    out->hbm_bytes_held = c->held;
    return GW_OK;
}

// split_blocks, preprocess/split.hpp:15-87: walk the .beg offsets once; cut before a
// vertex whose edges would overflow blocksize/4 edges (at v-1), or when the block
// reaches blocksize/8 vertices (at v); the last cut point is entries-1.
gw_status gw_split_blocks(const uint64_t* beg, uint64_t entries, uint64_t block_size, uint32_t* vblocks,
                          uint64_t* eblocks, uint64_t cap, uint64_t* ncuts)
{
    if (!beg || !vblocks || !eblocks || !ncuts || entries < 1 || block_size < 8)
        return fail(nullptr, GW_ERR_ARG, "gw_split_blocks: bad arguments");
    if (entries - 1 > 0xFFFFFFFFull) return fail(nullptr, GW_ERR_ARG, "too many vertices for u32 ids");
    const uint64_t max_e = block_size / 4, max_v = block_size / 8;
    uint64_t n = 0;
    uint32_t cur = 0;
    uint64_t rd = 0;
    auto push = [&](uint32_t v, uint64_t e) -> bool {
        if (n >= cap) return false;
        vblocks[n] = v; eblocks[n] = e; ++n;
        return true;
    };
    if (!push(0, 0)) return fail(nullptr, GW_ERR_ARG, "cut buffer too small");
    for (uint64_t v = 0; v < entries; ++v) {
        if (beg[v] - rd > max_e) {
            cur = (uint32_t)(v - 1);
            rd = beg[v - 1];
            if (!push(cur, rd)) return fail(nullptr, GW_ERR_ARG, "cut buffer too small");
        } else if ((uint32_t)v - cur >= max_v) {
            cur = (uint32_t)v;
            rd = beg[v];
            if (!push(cur, rd)) return fail(nullptr, GW_ERR_ARG, "cut buffer too small");
        }
    }
    if (!push((uint32_t)(entries - 1), beg[entries - 1])) return fail(nullptr, GW_ERR_ARG, "cut buffer too small");
    *ncuts = n;
    return GW_OK;
}

}  // extern "C"
