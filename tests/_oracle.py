"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).  Test infrastructure:
imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg."""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

N2V, SOPR = 0, 1


class Graph(C.Structure):
    _fields_ = [("nverts", C.c_uint64), ("beg", C.c_void_p), ("csr", C.c_void_p)]


class TypeCfg(C.Structure):
    _fields_ = [("algo", C.c_uint32), ("length", C.c_uint32), ("t_p", C.c_uint64), ("t_one", C.c_uint64),
                ("t_q", C.c_uint64), ("t_min", C.c_uint64), ("alpha_a", C.c_uint32), ("alpha_b", C.c_uint32),
                ("t_term", C.c_uint64), ("sopr_sampler", C.c_uint32), ("search", C.c_uint32)]


class WalkCfg(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("type", TypeCfg * 2)]


class WalkResult(C.Structure):
    _fields_ = [("cur", C.c_uint32), ("prev", C.c_uint32), ("hop", C.c_uint32), ("status", C.c_uint32),
                ("hash", C.c_uint64), ("steps", C.c_uint64), ("trials", C.c_uint64), ("probes", C.c_uint64),
                ("mix_steps", C.c_uint64)]


class Totals(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("trials", C.c_uint64), ("probes", C.c_uint64), ("mix_steps", C.c_uint64),
                ("walkers", C.c_uint64)]


class RefCfg(C.Structure):
    _fields_ = [("algo", C.c_uint32), ("p", C.c_float), ("q", C.c_float), ("alpha", C.c_float),
                ("length", C.c_uint32), ("term", C.c_double), ("literal", C.c_int), ("seed", C.c_uint64)]


class RandNum(C.Structure):
    _fields_ = [("s0", C.c_uint64), ("s1", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.orc_philox4x32_10.argtypes = [P, P, P]
        L.orc_randnum_init.argtypes = [C.POINTER(RandNum), C.c_uint64]
        L.orc_randnum_lrand.argtypes = [C.POINTER(RandNum)]
        L.orc_randnum_lrand.restype = C.c_uint64
        L.orc_randnum_pair.argtypes = [C.POINTER(RandNum), C.c_float, C.c_uint64, C.c_int,
                                       C.POINTER(C.c_float), C.POINTER(C.c_uint64)]
        L.orc_n2v_thresholds.argtypes = [C.c_double, C.c_double, C.POINTER(TypeCfg)]
        L.orc_alpha_fx.argtypes = [C.c_double]
        L.orc_alpha_fx.restype = C.c_uint32
        L.orc_term_threshold.argtypes = [C.c_double]
        L.orc_term_threshold.restype = C.c_uint64
        L.orc_walk.argtypes = [C.POINTER(Graph), C.POINTER(WalkCfg), C.c_uint32, C.c_uint32, C.c_uint32, P,
                               C.POINTER(WalkResult)]
        L.orc_run.argtypes = [C.POINTER(Graph), C.POINTER(WalkCfg), C.c_uint64, P, C.c_uint32, P, C.c_uint32,
                              P, P, P, P, C.POINTER(Totals), C.c_int]
        L.orc_sopr_use_mixture.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_sopr_use_mixture.restype = C.c_int
        L.orc_ref_run.argtypes = [C.POINTER(Graph), C.POINTER(RefCfg), C.c_uint64, C.c_uint64, P, C.c_uint32, P,
                                  C.POINTER(C.c_uint64), C.c_double, C.POINTER(C.c_uint64), C.c_int]
        L.orc_ref_run.restype = C.c_uint64
        L.orc_split_blocks.argtypes = [P, C.c_uint64, C.c_uint64, P, P, C.c_size_t]
        L.orc_split_blocks.restype = C.c_size_t
        L.orc_max_choose.argtypes = [P, C.c_uint32, P, P, C.c_uint32, C.c_uint64]
        L.orc_max_choose.restype = C.c_int
        L.orc_member.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64)]
        L.orc_member.restype = C.c_int
        _lib = L
    return _lib


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(ptr(c), ptr(k), ptr(out))
    return out


def make_graph(beg, csr):
    beg = np.ascontiguousarray(beg, dtype=np.uint64)
    csr = np.ascontiguousarray(csr, dtype=np.uint32)
    g = Graph(len(beg) - 1, beg.ctypes.data, csr.ctypes.data if len(csr) else beg.ctypes.data)
    g._keep = (beg, csr)
    return g


def type_cfg(algo, length, p=1.0, q=1.0, alpha=0.2, term=0.0, sampler=0, search=0):
    t = TypeCfg()
    t.algo = algo
    t.length = length
    t.sopr_sampler = sampler
    t.search = search
    if algo == N2V:
        lib().orc_n2v_thresholds(p, q, C.byref(t))
    else:
        a = lib().orc_alpha_fx(alpha)
        t.alpha_a = a
        t.alpha_b = 65536 - a
    t.t_term = lib().orc_term_threshold(term)
    return t


def walk_cfg(seed, t0, t1=None):
    w = WalkCfg()
    w.seed = seed
    w.type[0] = t0
    w.type[1] = t1 if t1 is not None else t0
    return w


def run(beg, csr, cfg, nwalkers, wpv=1, sources=None, ntypes=1, paths=False, hashes=True, state=True,
        visits=False, nthreads=0):
    """Replays every walker; returns a dict of numpy outputs and totals."""
    g = make_graph(beg, csr)
    src = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
    L = max(cfg.type[0].length, cfg.type[1].length if ntypes == 2 else 0)
    out = {}
    pth = np.zeros(nwalkers * (L + 1), dtype=np.uint32) if paths else None
    hs = np.zeros(nwalkers, dtype=np.uint64) if hashes else None
    st = np.zeros(nwalkers * 4, dtype=np.uint32) if state else None
    vs = np.zeros(g.nverts, dtype=np.uint32) if visits else None
    tot = Totals()
    lib().orc_run(C.byref(g), C.byref(cfg), nwalkers, ptr(src), wpv, None, 2 if ntypes == 2 else 0,
                  ptr(pth), ptr(hs), ptr(st), ptr(vs), C.byref(tot), nthreads)
    if paths:
        out["paths"] = pth.reshape(nwalkers, L + 1)
    if hashes:
        out["hashes"] = hs
    if state:
        out["state"] = st.reshape(nwalkers, 4)
    if visits:
        out["visits"] = vs
    out["totals"] = dict(steps=tot.steps, trials=tot.trials, probes=tot.probes, mix_steps=tot.mix_steps)
    return out


def walk_one(beg, csr, cfg, wid, source, wtype=0):
    g = make_graph(beg, csr)
    L = cfg.type[wtype].length
    path = np.zeros(L + 1, dtype=np.uint32)
    r = WalkResult()
    lib().orc_walk(C.byref(g), C.byref(cfg), wid, wtype, source, ptr(path), C.byref(r))
    return path, r


def ref_run(beg, csr, algo, length, p=1.0, q=1.0, alpha=0.2, term=0.0, literal=False, seed=9898676785859,
            nwalkers=None, wpv=1, w0=0, sources=None, visits=False, max_seconds=0.0, nthreads=0):
    """Transcription of the reference's CPU step (engine/engine.hpp:158-404) to completion."""
    g = make_graph(beg, csr)
    cfg = RefCfg(algo, p, q, alpha, length, term, 1 if literal else 0, seed)
    if nwalkers is None:
        nwalkers = g.nverts * wpv
    src = None if sources is None else np.ascontiguousarray(sources, dtype=np.uint32)
    vs = np.zeros(g.nverts, dtype=np.uint32) if visits else None
    tr = C.c_uint64(0)
    done = C.c_uint64(0)
    moves = lib().orc_ref_run(C.byref(g), C.byref(cfg), w0, nwalkers, ptr(src), wpv, ptr(vs), C.byref(tr),
                              max_seconds, C.byref(done), nthreads)
    return dict(moves=moves, trials=tr.value, done=done.value, visits=vs)


def split_blocks(beg, block_size, cap=1 << 16):
    beg = np.ascontiguousarray(beg, dtype=np.uint64)
    vb = np.zeros(cap, dtype=np.uint32)
    eb = np.zeros(cap, dtype=np.uint64)
    n = lib().orc_split_blocks(ptr(beg), len(beg), block_size, ptr(vb), ptr(eb), cap)
    return vb[:n].copy(), eb[:n].copy()


def max_choose(W, buckets, ncblock, zero_threshold=0):
    W = np.ascontiguousarray(W, dtype=np.uint64)
    nb = W.shape[0]
    b = np.zeros(max(ncblock, 1) + 1, dtype=np.uint32)
    b[:len(buckets)] = buckets
    n = C.c_uint32(len(buckets))
    r = lib().orc_max_choose(ptr(W), nb, ptr(b), C.byref(n), ncblock, zero_threshold)
    return r, list(b[:n.value])


def member(a, x, search=0, id_end=None, head=0):
    """(found, probes) of the oracle's membership test on the sorted u32 list a, whose
    first entry is entry `head` of an edge array (places the window orders' reads)."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    pr = C.c_uint64(0)
    end = int(id_end if id_end is not None else (int(a.max()) + 1 if len(a) else 1))
    r = lib().orc_member(a.ctypes.data, len(a), int(x), search, end, int(head), C.byref(pr))
    return bool(r), pr.value
