"""Python view of libgw.so for tests and benchmarks.

`Context` wraps one gw_ctx.  `Blocks` cuts an in-memory CSR graph into the reference's
blocks (split_blocks of preprocess/split.hpp:15-87 through gw_split_blocks, with the
giant single-vertex truncation of engine/cache.hpp:170-172).  Everything computes in
libgw.so; this module only moves numpy arrays across the C ABI.
"""
import ctypes as C

import numpy as np

from ._abi import (GW_NODE2VEC, GW_OUT_HASH, GW_OUT_PATHS, GW_OUT_VISITS, GW_RUN_BUCKETS, GW_RUN_INTERP, GW_RUN_WINDOW, GW_SOPR, GwBlockDesc,
                   GwConfig, GwCounters, GwDeviceInfo, GwLaunch, GwRunStats, GwWalkParams, check, lib)

NODE2VEC, SOPR = GW_NODE2VEC, GW_SOPR
OUT_VISITS, OUT_HASH, OUT_PATHS = GW_OUT_VISITS, GW_OUT_HASH, GW_OUT_PATHS
RUN_BUCKETS = GW_RUN_BUCKETS


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Blocks:
    """Block table of a CSR graph for a given block size in bytes."""

    def __init__(self, beg, csr, block_bytes):
        self.beg = np.ascontiguousarray(beg, dtype=np.uint64)
        self.csr = np.ascontiguousarray(csr, dtype=np.uint32)
        entries = len(self.beg)
        cap = entries + 2
        vb = np.zeros(cap, dtype=np.uint32)
        eb = np.zeros(cap, dtype=np.uint64)
        n = C.c_uint64(0)
        check(lib().gw_split_blocks(_p(self.beg), entries, block_bytes, _p(vb), _p(eb), cap, C.byref(n)))
        self.vcut = vb[:n.value].copy()
        self.ecut = eb[:n.value].copy()
        self.nblocks = n.value - 1
        maxedges = block_bytes // 4
        self.desc = (GwBlockDesc * self.nblocks)()
        for b in range(self.nblocks):
            d = self.desc[b]
            d.blk = b
            d.start_vert = int(self.vcut[b])
            d.nverts = int(self.vcut[b + 1] - self.vcut[b])
            d.start_edge = int(self.ecut[b])
            ne = int(self.ecut[b + 1] - self.ecut[b])
            if d.nverts == 1 and ne > maxedges:
                ne = maxedges
            d.nedges = ne

    def arrays(self, b):
        d = self.desc[b]
        beg = self.beg[d.start_vert:d.start_vert + d.nverts + 1]
        csr = self.csr[d.start_edge:d.start_edge + d.nedges]
        return np.ascontiguousarray(beg), np.ascontiguousarray(csr)

    def block_of(self, v):
        return int(np.searchsorted(self.vcut[:-1], v, side="right") - 1)


def pack_block(beg, csr):
    """gw_pack_block: the transfer encoding of one block (beg: nverts + 1 global offsets,
    csr: the block's arcs), as a u32 array (empty for a block without arcs)."""
    beg = np.ascontiguousarray(beg, dtype=np.uint64)
    csr = np.ascontiguousarray(csr, dtype=np.uint32)
    nb = C.c_uint64(0)
    check(lib().gw_pack_block(_p(beg), _p(csr) if len(csr) else None, len(beg) - 1, len(csr), None, 0, C.byref(nb)))
    out = np.zeros(max(1, nb.value // 4), dtype=np.uint32)
    check(lib().gw_pack_block(_p(beg), _p(csr) if len(csr) else None, len(beg) - 1, len(csr), _p(out), nb.value,
                              C.byref(nb)))
    return out[:nb.value // 4]


def walk_params(types, seed=42):
    """types: list of dicts {algo, length, p, q, alpha, term} (1 or 2 entries)."""
    wp = GwWalkParams()
    wp.seed = seed
    wp.ntypes = len(types)
    for i, t in enumerate(types):
        wt = wp.type[i]
        wt.algo = t.get("algo", NODE2VEC)
        wt.length = t["length"]
        wt.p = t.get("p", 1.0)
        wt.q = t.get("q", 1.0)
        wt.alpha = t.get("alpha", 0.2)
        wt.term = t.get("term", 0.0)
        wt.sopr_sampler = t.get("sampler", 0)
    return wp


class Context:
    def __init__(self, device=0, cache_bytes=0):
        cfg = GwConfig(device, 0, cache_bytes)
        h = C.c_void_p()
        check(lib().gw_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.blocks = None
        self._keep = []

    def close(self):
        if self.h:
            lib().gw_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st):
        return check(st, self.h)

    def device_info(self):
        i = GwDeviceInfo()
        self._chk(lib().gw_device_info_get(self.h, C.byref(i)))
        return dict(name=i.name.decode(), cus=i.cus, wave_size=i.wave_size, hbm_bytes=i.hbm_bytes,
                    lds_per_cu=i.lds_per_cu)

    def set_blocks(self, blocks):
        self.blocks = blocks
        self._chk(lib().gw_set_block_table(self.h, blocks.desc, blocks.nblocks))

    def slot_bytes(self):
        v = C.c_uint64(0)
        self._chk(lib().gw_slot_bytes(self.h, C.byref(v)))
        return v.value

    def reserve_slots(self, n):
        self._chk(lib().gw_reserve_slots(self.h, n))

    def upload(self, slot, blk):
        beg, csr = self.blocks.arrays(blk)
        self._keep.append((beg, csr))
        self._chk(lib().gw_upload_block(self.h, slot, blk, _p(beg), _p(csr)))

    def stage(self, slot, blk):
        """Prefetch blk into slot; it becomes resident only at commit(slot)."""
        beg, csr = self.blocks.arrays(blk)
        self._keep.append((beg, csr))
        self._chk(lib().gw_stage_block(self.h, slot, blk, _p(beg), _p(csr)))

    def set_pack_staging(self, nbytes):
        self._chk(lib().gw_set_pack_staging(self.h, nbytes))

    def set_copy_lanes(self, lanes):
        self._chk(lib().gw_set_copy_lanes(self.h, lanes))

    def upload_packed(self, slot, blk):
        """Upload blk into slot in the transfer encoding (packed on the host here); the
        staging buffer must hold it (set_pack_staging)."""
        beg, csr = self.blocks.arrays(blk)
        pk = pack_block(beg, csr)
        self._keep.append(pk)
        self._chk(lib().gw_upload_block_packed(self.h, slot, blk, _p(pk), 4 * len(pk)))
        return 4 * len(pk)

    def stage_packed(self, slot, blk):
        beg, csr = self.blocks.arrays(blk)
        pk = pack_block(beg, csr)
        self._keep.append(pk)
        self._chk(lib().gw_stage_block_packed(self.h, slot, blk, _p(pk), 4 * len(pk)))
        return 4 * len(pk)

    def read_slot(self, slot, blk):
        """(beg, csr) as slot holds them for blk (resident or staged), after its copy."""
        d = self.blocks.desc[blk]
        beg = np.zeros(d.nverts + 1, dtype=np.uint64)
        csr = np.zeros(max(1, d.nedges), dtype=np.uint32)
        self._chk(lib().gw_read_slot(self.h, slot, _p(beg), _p(csr)))
        return beg, csr[:d.nedges]

    def commit(self, slot):
        self._chk(lib().gw_commit_block(self.h, slot))

    def staged(self, slot):
        b, r = C.c_int32(-1), C.c_uint32(0)
        self._chk(lib().gw_staged_block(self.h, slot, C.byref(b), C.byref(r)))
        return b.value, bool(r.value)

    def evict(self, slot):
        self._chk(lib().gw_evict(self.h, slot))

    def wait_uploads(self):
        self._chk(lib().gw_wait_uploads(self.h))
        self._keep = []

    def load_all(self):
        """Reserve one slot per block and upload every block (fully resident graph)."""
        nb = self.blocks.nblocks
        self.reserve_slots(nb)
        for b in range(nb):
            self.upload(b, b)
        self.wait_uploads()

    def set_walk(self, types, seed=42):
        wp = walk_params(types, seed)
        self._chk(lib().gw_set_walk(self.h, C.byref(wp)))
        self.types = types

    def init_walkers(self, n, wpv=1, sources=None, outputs=OUT_HASH):
        src = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
        self._chk(lib().gw_init_walkers(self.h, n, wpv, _p(src), outputs))
        self.nwalkers = n
        self.outputs = outputs

    def run(self, block=0, waves_per_cu=0, variant=None, buckets=False, interp=False, window=False):
        """One gw_run.  block 0 takes the whole launch from the autotuner (or 256 threads of
        variant 0); naming a variant without a block runs it at 256 threads.  interp /
        window: membership search order (interpolated guesses / 64-byte window reads)."""
        if variant is not None and block == 0:
            block = 256
        L = GwLaunch(block, waves_per_cu, variant or 0,
                     (RUN_BUCKETS if buckets else 0) | (GW_RUN_INTERP if interp else 0) |
                     (GW_RUN_WINDOW if window else 0))
        s = GwRunStats()
        self._chk(lib().gw_run(self.h, C.byref(L), C.byref(s)))
        return {f: getattr(s, f) for f, _ in GwRunStats._fields_}

    def autotune(self, sample, what=7):
        """Pick the fastest launch shape / search order on the first `sample` walkers (state
        restored); `what`: 1 workgroup sizes, 2 kernel variants 1 and 2, 4 search order."""
        best = GwLaunch()
        ms = C.c_double()
        self._chk(lib().gw_autotune(self.h, sample, what, C.byref(best), C.byref(ms)))
        return dict(block=best.block, waves_per_cu=best.waves_per_cu, variant=best.variant,
                    interp=bool(best.flags & GW_RUN_INTERP), window=bool(best.flags & GW_RUN_WINDOW)), ms.value

    def bucket_counts(self):
        nb = self.blocks.nblocks
        W = np.zeros(nb * nb, dtype=np.uint64)
        self._chk(lib().gw_bucket_counts(self.h, _p(W), nb * nb))
        return W.reshape(nb, nb)

    def pending_work(self):
        """(unfinished walkers, sum of their remaining hop caps)"""
        w, h = C.c_uint64(0), C.c_uint64(0)
        self._chk(lib().gw_pending_work(self.h, C.byref(w), C.byref(h)))
        return w.value, h.value

    def block_features(self, age=None):
        nb = self.blocks.nblocks
        f = np.zeros(nb * 12, dtype=np.float32)
        a = None if age is None else np.ascontiguousarray(age, dtype=np.float32)
        self._chk(lib().gw_block_features(self.h, _p(a), _p(f), f.size))
        return f.reshape(nb, 12)

    def walkers_left(self):
        v = C.c_uint64(0)
        self._chk(lib().gw_walkers_left(self.h, C.byref(v)))
        return v.value

    def visits(self):
        n = int(self.blocks.vcut[-1])
        out = np.zeros(n, dtype=np.uint32)
        self._chk(lib().gw_read_visits(self.h, _p(out), n))
        return out

    def hashes(self):
        out = np.zeros(self.nwalkers, dtype=np.uint64)
        self._chk(lib().gw_read_hashes(self.h, _p(out), self.nwalkers))
        return out

    def state(self):
        out = np.zeros(self.nwalkers * 4, dtype=np.uint32)
        self._chk(lib().gw_read_state(self.h, _p(out), self.nwalkers))
        return out.reshape(self.nwalkers, 4)

    def paths(self):
        L = max(t["length"] for t in self.types)
        out = np.zeros(self.nwalkers * (L + 1), dtype=np.uint32)
        self._chk(lib().gw_read_paths(self.h, _p(out), out.size))
        return out.reshape(self.nwalkers, L + 1)

    def upload_stats(self):
        """packed blocks decoded (in place, through the ring, on the copy stream)"""
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._chk(lib().gw_upload_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_worklist_cap(self, entries):
        """worklist entries from the next init_walkers (0: nwalkers / 8, at least min(nwalkers, 2^20))"""
        self._chk(lib().gw_set_worklist_cap(self.h, entries))

    def worklist_stats(self):
        """(worklist capacity, further filter + step rounds taken since the context was made)"""
        cap, rounds = C.c_uint64(), C.c_uint64()
        self._chk(lib().gw_worklist_stats(self.h, C.byref(cap), C.byref(rounds)))
        return cap.value, rounds.value

    def graph_stats(self):
        """(runs replayed from captured graphs, graphs captured)"""
        r, c = C.c_uint64(), C.c_uint64()
        self._chk(lib().gw_graph_stats(self.h, C.byref(r), C.byref(c)))
        return r.value, c.value

    def counters(self):
        c = GwCounters()
        self._chk(lib().gw_counters_get(self.h, C.byref(c)))
        return {f: getattr(c, f) for f, _ in GwCounters._fields_}


class PinnedArray:
    """numpy view of page-locked host memory from gw_alloc_pinned (freed on close)."""

    def __init__(self, n, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        p = C.c_void_p()
        check(lib().gw_alloc_pinned(max(1, self.n * self.dtype.itemsize), C.byref(p)))
        self.ptr = p
        buf = (C.c_char * max(1, self.n * self.dtype.itemsize)).from_address(p.value)
        self.a = np.frombuffer(buf, dtype=self.dtype, count=self.n)

    def close(self):
        if self.ptr:
            self.a = None
            lib().gw_free_pinned(self.ptr)
            self.ptr = None


def _read_graph(ctx, g, pinned):
    nv, ne = C.c_uint64(), C.c_uint64()
    check(lib().gw_graph_size(g, C.byref(nv), C.byref(ne)))
    if pinned:
        beg, csr = PinnedArray(nv.value + 1, np.uint64), PinnedArray(ne.value, np.uint32)
        check(lib().gw_graph_read(ctx.h, g, _p(beg.a), _p(csr.a)), ctx.h)
        return beg, csr
    beg = np.zeros(nv.value + 1, dtype=np.uint64)
    csr = np.zeros(ne.value, dtype=np.uint32)
    check(lib().gw_graph_read(ctx.h, g, _p(beg), _p(csr)), ctx.h)
    return beg, csr


def gen_rmat(ctx, scale, nedges, a=0.57, b=0.19, c=0.19, seed=42, undirected=True, permute=False, dedup=True,
             pinned=False):
    """RMAT graph generated on the GPU, returned as host CSR arrays (numpy or PinnedArray)."""
    from ._abi import GW_GEN_DEDUP, GW_GEN_PERMUTE, GW_GEN_UNDIRECTED
    flags = (GW_GEN_UNDIRECTED if undirected else 0) | (GW_GEN_PERMUTE if permute else 0) | \
        (GW_GEN_DEDUP if dedup else 0)
    g = C.c_void_p()
    check(lib().gw_gen_rmat(ctx.h, scale, nedges, a, b, c, seed, flags, C.byref(g)), ctx.h)
    try:
        return _read_graph(ctx, g, pinned)
    finally:
        lib().gw_graph_free(ctx.h, g)


def gen_chung_lu(ctx, nverts, nedges, gamma=2.2, max_degree=3000000, seed=42, dedup=True, pinned=False):
    """Directed Chung-Lu power-law graph generated on the GPU, as host CSR arrays."""
    from ._abi import GW_GEN_DEDUP
    g = C.c_void_p()
    check(lib().gw_gen_chung_lu(ctx.h, nverts, nedges, gamma, max_degree, seed, GW_GEN_DEDUP if dedup else 0,
                                C.byref(g)), ctx.h)
    try:
        return _read_graph(ctx, g, pinned)
    finally:
        lib().gw_graph_free(ctx.h, g)


SCHED_PLAIN, SCHED_LEARNED, SCHED_SOWALKER, SCHED_STAGING = 0, 1, 2, 3
# gwh_engine_opts.flags
NO_SWAP_PREFETCH, MYOPIC_LABEL, MODEL_SWAP, VALUE_VICTIM, NO_STICKY, NO_TUNE, TUNE_FREEZE, TWO_LANES = \
    1, 2, 4, 8, 16, 32, 64, 128


class Model:
    """The learned scheduler's state across walks (gwh_model): hand the same Model to
    successive run_out_of_core calls and each walk starts from the network the earlier ones
    trained."""

    def __init__(self):
        from ._abi import host_lib
        h = C.c_void_p()
        check(host_lib().gwh_model_create(C.byref(h)))
        self.h = h

    def info(self):
        from ._abi import GwhModelInfo, host_lib
        i = GwhModelInfo()
        check(host_lib().gwh_model_info_get(self.h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in GwhModelInfo._fields_}

    def tuning(self):
        """The launch tuner's table for the workload of the last walk: a list of dicts
        (block, waves_per_cu, variant, ms_per_gstep, walks) and the index of the fastest."""
        from ._abi import GwhTuneEntry, host_lib
        out = (GwhTuneEntry * 16)()
        n, ch = C.c_uint32(0), C.c_uint32(0)
        check(host_lib().gwh_model_tuning(self.h, out, 16, C.byref(n), C.byref(ch)))
        rows = [dict(block=e.launch.block, waves_per_cu=e.launch.waves_per_cu, variant=e.launch.variant,
                     ms_per_gstep=e.ms_per_gstep, walks=e.walks) for e in out[:n.value]]
        return rows, ch.value

    def close(self):
        if self.h:
            from ._abi import host_lib
            host_lib().gwh_model_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PackedGraph:
    """gwh_pack_graph: the blocks of (beg, csr) at block_bytes in the transfer encoding, in
    pinned memory; pass it to run_out_of_core over the same arrays and block size."""

    def __init__(self, beg, csr, block_bytes):
        from ._abi import host_lib
        h = C.c_void_p()
        st = host_lib().gwh_pack_graph(_p(beg), _p(csr), len(beg), block_bytes, C.byref(h))
        if st != 0:
            from ._abi import GwError
            raise GwError(st, host_lib().gwh_last_error().decode())
        self.h = h
        self._keep = (beg, csr)

    def info(self):
        from ._abi import host_lib
        r, p, n, k = C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_uint32()
        check(host_lib().gwh_packed_info(self.h, C.byref(r), C.byref(p), C.byref(n), C.byref(k)))
        return dict(raw_bytes=r.value, packed_bytes=p.value, nblocks=n.value, blocks_packed=k.value)

    def close(self):
        if self.h:
            from ._abi import host_lib
            host_lib().gwh_packed_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_out_of_core(beg, csr, types, seed, nwalkers, wpv=1, sources=None, memory_bytes=0, block_bytes=256 << 20,
                    zero_threshold=0, scheduler=SCHED_PLAIN, hashes=False, visits=False, verbose=False, launch=None,
                    flags=0, prefetch_depth=0, model=None, packed=None):
    """Runs the C++ out-of-core engine (libgwhost.so) over a host CSR graph; beg/csr should
    be pinned (PinnedArray.a) for the zero-copy tail.  model: a Model carried across walks
    (learned scheduler).  Returns (report dict, hashes, visits)."""
    from ._abi import GwhEngineOpts, GwhReport, host_lib
    L = host_lib()
    o = GwhEngineOpts()
    o.memory_bytes = memory_bytes
    o.block_bytes = block_bytes
    o.zero_threshold = zero_threshold
    o.scheduler = scheduler
    o.verbose = 1 if verbose else 0
    o.flags = flags
    o.prefetch_depth = prefetch_depth
    o.model = model.h if model is not None else None
    o.packed = packed.h if packed is not None else None
    launch = launch or {}
    o.launch = GwLaunch(launch.get("block", 0), launch.get("waves_per_cu", 0), launch.get("variant", 2),
                        (GW_RUN_INTERP if launch.get("interp") else 0) | (GW_RUN_WINDOW if launch.get("window") else 0))
    wp = walk_params(types, seed)
    rep = GwhReport()
    h = np.zeros(nwalkers, dtype=np.uint64) if hashes else None
    v = np.zeros(len(beg) - 1, dtype=np.uint32) if visits else None
    src = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
    st = L.gwh_run(_p(beg), _p(csr), len(beg), C.byref(o), C.byref(wp), nwalkers, wpv, _p(src), C.byref(rep),
                   _p(h), _p(v))
    if st != 0:
        from ._abi import GwError
        raise GwError(st, L.gwh_last_error().decode())
    out = {f: getattr(rep, f) for f, _ in GwhReport._fields_ if f not in ("pad", "pad2", "tuned")}
    t = rep.tuned
    out["tuned"] = dict(block=t.block, waves_per_cu=t.waves_per_cu, variant=t.variant,
                        interp=bool(t.flags & GW_RUN_INTERP), window=bool(t.flags & GW_RUN_WINDOW)) if t.block else None
    return out, h, v
