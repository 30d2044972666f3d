"""The allocation-free walk (gw_walk_resident_noalloc) against the context path it varies
(gw_set_block_table + gw_upload_block + gw_set_walk + gw_init_walkers + gw_run over one
resident block): per-walker path hashes, visit counts and the step / trial / probe
counters must be identical, on two calls in a row -- the second with a smaller graph and
fewer walkers, so that state the first call left in the library's fixed device array
would show -- for Node2vec and for a mixed Node2vec + SOSR batch with termination.  The
test's own device buffers (graph, outputs) come from the HIP runtime through ctypes; the
entry point allocates nothing."""
import ctypes as C

import numpy as np
import pytest

import graphs as G
from gowalker_amd import engine as E
from gowalker_amd._abi import GwRunStats, lib

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        lib()   # the HIP runtime libgw.so links, loaded once
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class Dev:
    """A device buffer holding (or sized for) a numpy array."""

    def __init__(self, arr=None, n=0, dtype=np.uint32):
        self.dtype = np.dtype(arr.dtype if arr is not None else dtype)
        self.n = len(arr) if arr is not None else n
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(1, self.n * self.dtype.itemsize)) == 0
        if arr is not None and self.n:
            a = np.ascontiguousarray(arr)
            assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def get(self):
        out = np.zeros(self.n, dtype=self.dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        hip().hipFree(self.p)


def context_walk(beg, csr, types, seed, n, wpv):
    with E.Context() as ctx:
        ctx.set_blocks(E.Blocks(beg, csr, 1 << 34))
        ctx.load_all()
        ctx.set_walk(types, seed)
        ctx.init_walkers(n, wpv, outputs=E.OUT_HASH | E.OUT_VISITS)
        st = ctx.run()
        return ctx.hashes(), ctx.visits(), st


def noalloc_walk(beg, csr, types, seed, n, wpv):
    d_beg, d_csr = Dev(beg), Dev(csr)
    d_hash, d_vis = Dev(n=n, dtype=np.uint64), Dev(n=len(beg) - 1, dtype=np.uint32)
    wp = E.walk_params(types, seed)
    st = GwRunStats()
    rc = lib().gw_walk_resident_noalloc(d_beg.p, d_csr.p, len(beg) - 1, len(csr), C.byref(wp), n, wpv, d_hash.p,
                                        d_vis.p, None, C.byref(st))
    assert rc == 0, lib().gw_last_error(None)
    h, v = d_hash.get(), d_vis.get()
    for d in (d_beg, d_csr, d_hash, d_vis):
        d.free()
    return h, v, {f: getattr(st, f) for f, _ in GwRunStats._fields_}


CASES = [
    ("node2vec", [dict(algo=E.NODE2VEC, length=40, p=0.5, q=2.0)], 1),
    ("mixed", [dict(algo=E.NODE2VEC, length=30, p=2.0, q=0.5), dict(algo=E.SOPR, length=30, alpha=0.2, term=0.15)],
     2),
]


@pytest.mark.parametrize("name,types,wpv", CASES)
def test_noalloc_walk_matches_context_walk(gpu, name, types, wpv):
    graphs = [G.rmat(14, 150000, seed=3, permute=True), G.rmat(11, 12000, seed=4, undirected=False)]
    for call, (beg, csr) in enumerate(graphs):
        n = (len(beg) - 1) * wpv if call == 0 else (len(beg) - 1) * wpv // 3
        print(name, "call", call, "nverts", len(beg) - 1, "narcs", len(csr), "walkers", n, flush=True)
        h0, v0, s0 = context_walk(beg, csr, types, 77 + call, n, wpv)
        h1, v1, s1 = noalloc_walk(beg, csr, types, 77 + call, n, wpv)
        np.testing.assert_array_equal(h1, h0)
        np.testing.assert_array_equal(v1, v0)
        for k in ("steps", "trials", "probes", "mix_steps", "walkers_ended"):
            assert s1[k] == s0[k], (k, s1[k], s0[k])
        assert s1["walkers_left"] == 0 and s1["steps"] > 0
