"""ctypes declarations of libgw.so's C ABI (include/gw.h).

The library is loaded from gowalker_amd/_build/libgw.so; there is no fallback: if it is
missing, `lib()` raises so that no caller silently runs something else.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libgw.so")

GW_OK = 0
GW_ERR = {-1: "ARG", -2: "HIP", -3: "NOMEM", -4: "STATE", -5: "BUDGET", -6: "IO", -7: "FAULT"}
GW_NODE2VEC, GW_SOPR = 0, 1
GW_OUT_VISITS, GW_OUT_HASH, GW_OUT_PATHS = 1, 2, 4
GW_RUN_BUCKETS, GW_RUN_TRACE, GW_RUN_INTERP, GW_RUN_WINDOW = 1, 2, 4, 8
GW_GEN_UNDIRECTED, GW_GEN_PERMUTE, GW_GEN_DEDUP = 1, 2, 4


class GwConfig(C.Structure):
    _fields_ = [("device", C.c_int32), ("reserved", C.c_uint32), ("cache_bytes", C.c_uint64)]


class GwBlockDesc(C.Structure):
    _fields_ = [("blk", C.c_uint32), ("start_vert", C.c_uint32), ("nverts", C.c_uint32), ("pad", C.c_uint32),
                ("start_edge", C.c_uint64), ("nedges", C.c_uint64)]


class GwWalkType(C.Structure):
    _fields_ = [("algo", C.c_uint32), ("length", C.c_uint32), ("p", C.c_double), ("q", C.c_double),
                ("alpha", C.c_double), ("term", C.c_double), ("sopr_sampler", C.c_uint32),
                ("reserved", C.c_uint32)]


class GwWalkParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("ntypes", C.c_uint32), ("reserved", C.c_uint32), ("type", GwWalkType * 2)]


class GwLaunch(C.Structure):
    _fields_ = [("block", C.c_uint32), ("waves_per_cu", C.c_uint32), ("variant", C.c_uint32), ("flags", C.c_uint32)]


class GwRunStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("trials", C.c_uint64), ("probes", C.c_uint64), ("mix_steps", C.c_uint64),
                ("walkers_run", C.c_uint64), ("walkers_ended", C.c_uint64), ("walkers_left", C.c_uint64),
                ("kernel_ms", C.c_double), ("filter_ms", C.c_double), ("zc_touches", C.c_uint64)]


class GwCounters(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("trials", C.c_uint64), ("probes", C.c_uint64), ("mix_steps", C.c_uint64),
                ("h2d_bytes", C.c_uint64), ("h2d_uploads", C.c_uint64), ("h2d_ms", C.c_double),
                ("kernel_ms", C.c_double), ("hbm_bytes_held", C.c_uint64), ("zc_touches", C.c_uint64),
                ("h2d_raw_bytes", C.c_uint64)]


class GwDeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("cus", C.c_uint32), ("wave_size", C.c_uint32), ("hbm_bytes", C.c_uint64),
                ("lds_per_cu", C.c_uint32), ("pad", C.c_uint32)]


P = C.c_void_p
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)

# name -> (argtypes); every entry returns gw_status unless listed in _RESTYPE
SIGNATURES = {
    "gw_create": [C.POINTER(GwConfig), C.POINTER(P)],
    "gw_destroy": [P],
    "gw_last_error": [P],
    "gw_device_info_get": [P, C.POINTER(GwDeviceInfo)],
    "gw_alloc_pinned": [C.c_size_t, C.POINTER(P)],
    "gw_free_pinned": [P],
    "gw_set_block_table": [P, C.POINTER(GwBlockDesc), C.c_uint32],
    "gw_reserve_slots": [P, C.c_uint32],
    "gw_slot_bytes": [P, u64p],
    "gw_upload_block": [P, C.c_uint32, C.c_uint32, P, P],
    "gw_evict": [P, C.c_uint32],
    "gw_wait_uploads": [P],
    "gw_map_host_block": [P, C.c_uint32, P, P],
    "gw_unmap_host_blocks": [P],
    "gw_set_walk": [P, C.POINTER(GwWalkParams)],
    "gw_init_walkers": [P, C.c_uint64, C.c_uint32, P, C.c_uint32],
    "gw_stage_block": [P, C.c_uint32, C.c_uint32, P, P],
    "gw_pending_work": [P, u64p, u64p],
    "gw_commit_block": [P, C.c_uint32],
    "gw_staged_block": [P, C.c_uint32, C.POINTER(C.c_int32), u32p],
    "gw_run": [P, C.POINTER(GwLaunch), C.POINTER(GwRunStats)],
    "gw_autotune": [P, C.c_uint64, C.c_uint32, C.POINTER(GwLaunch), C.POINTER(C.c_double)],
    "gw_bucket_counts": [P, P, C.c_uint64],
    "gw_walkers_left": [P, u64p],
    "gw_read_visits": [P, P, C.c_uint64],
    "gw_read_hashes": [P, P, C.c_uint64],
    "gw_read_state": [P, P, C.c_uint64],
    "gw_read_paths": [P, P, C.c_uint64],
    "gw_counters_get": [P, C.POINTER(GwCounters)],
    "gw_graph_stats": [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
    "gw_upload_stats": [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
    "gw_set_worklist_cap": [P, C.c_uint64],
    "gw_worklist_stats": [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)],
    "gw_split_blocks": [P, C.c_uint64, C.c_uint64, P, P, C.c_uint64, u64p],
    "gw_gen_rmat": [P, C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_uint32,
                    C.POINTER(P)],
    "gw_gen_chung_lu": [P, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64, C.c_uint64, C.c_uint32, C.POINTER(P)],
    "gw_graph_size": [P, u64p, u64p],
    "gw_graph_read": [P, P, P, P],
    "gw_graph_free": [P, P],
    "gw_trace_read": [P, P, C.c_uint64],
    "gw_block_features": [P, P, P, C.c_uint64],
    "gw_block_scores": [P, P, P, P, P, P, P, C.c_uint64],
    "gw_mlp_create": [P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(P)],
    "gw_mlp_destroy": [P, P],
    "gw_mlp_num_params": [P, u32p],
    "gw_mlp_set_params": [P, P, P],
    "gw_mlp_get_params": [P, P, P],
    "gw_mlp_forward": [P, P, P, C.c_uint32, P],
    "gw_mlp_backward": [P, P, P, P, C.c_uint32, P, P],
    "gw_mlp_step": [P, P, P, P, C.c_uint32, C.c_float, P],
    "gw_set_pack_staging": [P, C.c_uint64],
    "gw_set_copy_lanes": [P, C.c_uint32],
    "gw_read_slot": [P, C.c_uint32, P, P],
    "gw_upload_block_packed": [P, C.c_uint32, C.c_uint32, P, C.c_uint64],
    "gw_stage_block_packed": [P, C.c_uint32, C.c_uint32, P, C.c_uint64],
    "gw_pack_block": [P, P, C.c_uint32, C.c_uint64, P, C.c_uint64, u64p],
    "gw_walk_resident_noalloc": [P, P, C.c_uint32, C.c_uint64, C.POINTER(GwWalkParams), C.c_uint64, C.c_uint32, P, P,
                                 P, C.POINTER(GwRunStats)],
}
GW_NOALLOC_MAX_WALKERS = 1 << 20
GW_NOALLOC_SCRATCH_BYTES = 1024 + 12 * GW_NOALLOC_MAX_WALKERS
_RESTYPE = {"gw_last_error": C.c_char_p}

_lib = None


def lib():
    """Load libgw.so (built by `make -C gowalker_amd`).  Raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgw.so not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                               "or `make -C gowalker_amd` (expected %s)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int32)
        _lib = L
    return _lib


class GwError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("gw status %d (%s): %s" % (status, GW_ERR.get(status, "?"), msg))
        self.status = status


def check(status, ctx=None):
    if status != GW_OK:
        msg = lib().gw_last_error(ctx)
        raise GwError(status, msg.decode() if msg else "")
    return status


# ---- libgwhost.so (include/gwh.h): the C++ out-of-core engine ----
HOST_LIB_PATH = os.path.join(_HERE, "_build", "libgwhost.so")


class GwhEngineOpts(C.Structure):
    _fields_ = [("memory_bytes", C.c_uint64), ("block_bytes", C.c_uint64), ("zero_threshold", C.c_uint64),
                ("scheduler", C.c_uint32), ("verbose", C.c_uint32), ("launch", GwLaunch), ("device", C.c_int32),
                ("flags", C.c_uint32), ("prefetch_depth", C.c_uint32), ("reserved", C.c_uint32),
                ("model", C.c_void_p), ("packed", C.c_void_p)]


class GwhReport(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("trials", C.c_uint64), ("probes", C.c_uint64), ("mix_steps", C.c_uint64),
                ("h2d_bytes", C.c_uint64), ("uploads", C.c_uint64), ("iterations", C.c_uint64),
                ("zero_copy_runs", C.c_uint64), ("zero_copy_steps", C.c_uint64), ("wall_s", C.c_double),
                ("kernel_ms", C.c_double), ("filter_ms", C.c_double), ("h2d_ms", C.c_double),
                ("nslots", C.c_uint32), ("nblocks", C.c_uint32), ("learned_decisions", C.c_uint64),
                ("fallbacks", C.c_uint64), ("final_loss", C.c_float), ("pad", C.c_uint32),
                ("staged", C.c_uint64), ("zc_touches", C.c_uint64), ("tail_decisions", C.c_uint64),
                ("sched_ms", C.c_double), ("tune_ms", C.c_double), ("tuned", GwLaunch),
                ("tail_model", C.c_uint64), ("tail_threshold", C.c_uint64), ("swap_model", C.c_uint64),
                ("swap_differs", C.c_uint64), ("fill_model", C.c_uint64), ("model_samples", C.c_uint64),
                ("sticky_swaps", C.c_uint64), ("tune_candidate", C.c_uint32), ("copy_lanes", C.c_uint32),
                ("h2d_raw_bytes", C.c_uint64), ("inplace_blocks", C.c_uint64), ("ring_blocks", C.c_uint64),
                ("inline_blocks", C.c_uint64), ("kernel_launches", C.c_uint64)]


class GwhTuneEntry(C.Structure):
    _fields_ = [("launch", GwLaunch), ("ms_per_gstep", C.c_double), ("walks", C.c_uint32), ("pad", C.c_uint32)]


class GwhModelInfo(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("walks", C.c_uint64), ("adam_steps", C.c_uint64), ("loss", C.c_float),
                ("has_params", C.c_uint32), ("scored", C.c_uint64), ("net_abs_err", C.c_double),
                ("mean_abs_err", C.c_double)]


HOST_SIGNATURES = {
    "gwh_run": [P, P, C.c_uint64, C.POINTER(GwhEngineOpts), C.POINTER(GwWalkParams), C.c_uint64, C.c_uint32, P,
                C.POINTER(GwhReport), P, P],
    "gwh_last_error": [],
    "gwh_model_create": [C.POINTER(P)],
    "gwh_model_destroy": [P],
    "gwh_model_info_get": [P, C.POINTER(GwhModelInfo)],
    "gwh_model_tuning": [P, P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)],
    "gwh_pack_graph": [P, P, C.c_uint64, C.c_uint64, C.POINTER(P)],
    "gwh_packed_info": [P, u64p, u64p, u32p, u32p],
    "gwh_packed_destroy": [P],
}

_hlib = None


def host_lib():
    global _hlib
    if _hlib is None:
        lib()
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError("libgwhost.so not built (expected %s)" % HOST_LIB_PATH)
        L = C.CDLL(HOST_LIB_PATH)
        for name, args in HOST_SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_char_p if name == "gwh_last_error" else C.c_int32
        _hlib = L
    return _hlib
