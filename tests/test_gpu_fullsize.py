"""Parity at the full out-of-core size (Node2vec on RMAT scale 26, 3.6B arcs, 256 MB
blocks; the mixed batch on RMAT scale 27, 5.5B arcs), where the oracle cannot follow, through properties that do not depend on size:

* on a symmetrized graph without termination every walker that starts on a vertex with
  neighbours takes exactly `length` steps (a vertex a walk reaches has its predecessor as
  a neighbour, so no later dead end), and the others none: steps == 40 x #non-isolated;
* walks depend only on (seed, walker, hop, trial), so the per-walker path hashes are
  identical across block schedulers, HBM caps (16 GB: every block resident once; 8 GB:
  evictions and the zero-copy tail) and step kernel variants."""
import time

import numpy as np
import pytest

from gowalker_amd import engine as E

pytestmark = pytest.mark.gpu


def test_node2vec_rmat26_out_of_core_invariants(gpu):
    ctx = E.Context(0)
    # 1.856B edges drawn leave the stated >= 3.6B arcs after symmetrizing and dedup
    beg, csr = E.gen_rmat(ctx, 26, 1_856_000_000, seed=42, undirected=True, permute=True, pinned=True)
    ctx.close()
    b, c = beg.a, csr.a
    n = len(b) - 1
    active = int(np.count_nonzero(b[1:] != b[:-1]))
    print("nverts", n, "narcs", len(c), "non-isolated", active, flush=True)
    assert len(c) >= 3_600_000_000
    types = [dict(algo=0, length=40, p=0.5, q=2.0)]
    ref = None
    for name, sched, mem_gb, zt, launch in (("plain v0 16GB", E.SCHED_PLAIN, 16, 0, dict(variant=0)),
                                           ("learned v2 16GB", E.SCHED_LEARNED, 16, 0, None),
                                           ("learned v2 8GB", E.SCHED_LEARNED, 8, 0, None),
                                           ("plain+zero v2 8GB", E.SCHED_PLAIN, 8, (256 << 20) // 3072, None)):
        rep, h, _ = E.run_out_of_core(b, c, types, 42, n, memory_bytes=mem_gb << 30, block_bytes=256 << 20,
                                      zero_threshold=zt, scheduler=sched, hashes=True, launch=launch)
        print(name, "steps", rep["steps"], "uploads", rep["uploads"], "staged", rep["staged"], "zero-copy reads",
              rep["zc_touches"], "wall %.2f s" % rep["wall_s"], flush=True)
        assert rep["steps"] == 40 * active
        if ref is None:
            ref = h.copy()
        else:
            np.testing.assert_array_equal(h, ref)
    assert rep["zc_touches"] > 0   # the 8 GB plain+zero walk ended with zero-copy reads


def test_sopr_chung_lu_resident_invariants(gpu):
    """SOPR (proposal sampler, termination 0.15) on the directed Chung-Lu graph of 41.6M
    vertices and ~1.47B arcs, 4 walkers per vertex, fully resident: the visit counts sum to
    the steps, and paths, visits and counters are identical whether the graph is one block
    (descriptor in registers) or ~25 blocks of 256 MB (table in LDS), and for step kernel
    variants 0 and 2 -- and they are the oracle's, replayed on the host cores."""
    import _oracle as O

    ctx = E.Context(0)
    beg, csr = E.gen_chung_lu(ctx, 41_600_000, 1_470_000_000, 2.2, 3_000_000, 42, pinned=True)
    ctx.close()
    b, c = beg.a, csr.a
    n = len(b) - 1
    types = [dict(algo=E.SOPR, length=30, alpha=0.2, term=0.15)]
    ref = None
    for block_bytes, variant in ((1 << 34, 0), (1 << 34, 2), (256 << 20, 0), (256 << 20, 2)):
        bl = E.Blocks(b, c, block_bytes)
        with E.Context() as cx:
            cx.set_blocks(bl)
            cx.load_all()
            cx.set_walk(types, 42)
            cx.init_walkers(4 * n, 4, outputs=E.OUT_HASH | E.OUT_VISITS)
            st = cx.run(variant=variant)
            h, v = cx.hashes(), cx.visits()
        print("blocks", bl.nblocks, "variant", variant, "steps", st["steps"], "trials", st["trials"], "probes",
              st["probes"], "kernel %.1f ms" % st["kernel_ms"], flush=True)
        assert st["walkers_left"] == 0
        assert int(v.sum(dtype=np.uint64)) == st["steps"]
        tot = (st["steps"], st["trials"], st["probes"])
        if ref is None:
            o = O.run(b, c, O.walk_cfg(42, O.type_cfg(O.SOPR, 30, alpha=0.2, term=0.15)), 4 * n, wpv=4, visits=True,
                      state=False, nthreads=16)
            print("oracle totals", o["totals"], flush=True)
            np.testing.assert_array_equal(h, o["hashes"])
            np.testing.assert_array_equal(v, o["visits"])
            assert tot == (o["totals"]["steps"], o["totals"]["trials"], o["totals"]["probes"])
            del o
            ref = (h.copy(), v.copy(), tot)
        else:
            np.testing.assert_array_equal(h, ref[0])
            np.testing.assert_array_equal(v, ref[1])
            assert tot == ref[2]


def test_node2vec_rmat20_full_size_matches_oracle(gpu):
    """The resident Node2vec workload at its full size (RMAT scale 20, ~31M arcs, 2^20
    walkers x 80 hops, p 2, q 0.5), bit for bit against the oracle's replay: path hashes,
    final states, and step / trial / probe counts, for the autotuned launch and for every
    search order of both resident kernel variants."""
    import _oracle as O

    ctx = E.Context(0)
    beg, csr = E.gen_rmat(ctx, 20, 16 * 1024 * 1024, seed=42, undirected=True, permute=False)
    ctx.close()
    n = len(beg) - 1
    types = [dict(algo=E.NODE2VEC, length=80, p=2.0, q=0.5)]
    want = {}
    for search in (0, 1, 2, 3):
        cfg = O.walk_cfg(42, O.type_cfg(O.N2V, 80, p=2.0, q=0.5, search=search))
        want[search] = O.run(beg, csr, cfg, n, nthreads=16)
    print("oracle totals", [want[s]["totals"] for s in want], flush=True)
    bl = E.Blocks(beg, csr, 1 << 34)
    with E.Context() as cx:
        cx.set_blocks(bl)
        cx.load_all()
        cx.set_walk(types, 42)
        cx.init_walkers(n, 1, outputs=E.OUT_HASH)
        cx.autotune(1 << 18)
        for launch in (dict(block=0), dict(variant=0), dict(variant=2), dict(variant=0, interp=True),
                       dict(variant=2, interp=True), dict(variant=0, window=True), dict(variant=2, window=True),
                       dict(variant=2, interp=True, window=True)):
            cx.init_walkers(n, 1, outputs=E.OUT_HASH)
            st = cx.run(**launch)
            o = want[(1 if launch.get("interp") else 0) | (2 if launch.get("window") else 0)]
            print(launch, {k: st[k] for k in ("steps", "trials", "probes")}, "kernel %.2f ms" % st["kernel_ms"],
                  flush=True)
            if launch != dict(block=0):
                assert (st["steps"], st["trials"], st["probes"]) == (o["totals"]["steps"], o["totals"]["trials"],
                                                                     o["totals"]["probes"])
            else:
                assert st["steps"] == o["totals"]["steps"] and st["trials"] == o["totals"]["trials"]
            np.testing.assert_array_equal(cx.hashes(), o["hashes"])
            np.testing.assert_array_equal(cx.state(), o["state"])


@pytest.mark.timeout(1200)
def test_mixed_rmat27_out_of_core_invariants(gpu):
    """The mixed batch at its stated size: RMAT scale 27 (134.2M ids), directed edges drawn
    until >= 5.5B arcs remain after deduplication, ids in generation order (locality kept),
    256 MB blocks; 2 walkers per vertex, Node2vec (p 1, q 0.25, length 80) and SOSR (alpha
    0.2, termination 0.2, cap 80) co-scheduled in one kernel.  Size-independent properties:
    per-walker path hashes identical for the learned scheduler at 16 GB, the plain scheduler
    with the reference's zero-copy tail at 16 GB and the learned scheduler at 24 GB (three
    schedules, two caps); visit counts sum to the steps (each step is one arrival); no walker
    exceeds its cap.  The first 131,072 walkers (ids 0..131071, both types) are replayed by
    the oracle on the host cores over the same full graph and their hashes and arrival
    counts must match the GPU's bit for bit."""
    import _oracle as O

    t0 = time.time()
    ctx = E.Context(0)
    draw = 5_630_000_000
    beg, csr = E.gen_rmat(ctx, 27, draw, seed=42, undirected=False, permute=False, pinned=True)
    ctx.close()
    b, c = beg.a, csr.a
    n = len(b) - 1
    print("nverts", n, "narcs", len(c), "drawn", draw, "gen %.1f s" % (time.time() - t0), flush=True)
    assert len(c) >= 5_500_000_000
    types = [dict(algo=E.NODE2VEC, length=80, p=1.0, q=0.25), dict(algo=E.SOPR, length=80, alpha=0.2, term=0.2)]
    nw = 2 * n
    zt = (256 << 20) // 128 // 24
    ref = None
    for name, sched, mem_gb, z in (("learned 16GB", E.SCHED_LEARNED, 16, 0), ("plain+zero 16GB", E.SCHED_PLAIN, 16, zt),
                                   ("learned 24GB", E.SCHED_LEARNED, 24, 0)):
        rep, h, v = E.run_out_of_core(b, c, types, 42, nw, wpv=2, memory_bytes=mem_gb << 30, block_bytes=256 << 20,
                                      zero_threshold=z, scheduler=sched, hashes=True, visits=(ref is None))
        print(name, "steps", rep["steps"], "uploads", rep["uploads"], "staged", rep["staged"], "slots", rep["nslots"],
              "blocks", rep["nblocks"], "zero-copy reads", rep["zc_touches"], "wall %.2f s" % rep["wall_s"], flush=True)
        assert rep["nslots"] < rep["nblocks"]
        assert rep["steps"] <= 80 * nw
        if ref is None:
            ref, steps0 = h, rep["steps"]
            assert int(v.sum(dtype=np.uint64)) == rep["steps"]
            visits = v
        else:
            assert rep["steps"] == steps0
            np.testing.assert_array_equal(h, ref)
    k = 131072
    tc0 = O.type_cfg(O.N2V, 80, p=1.0, q=0.25)
    tc1 = O.type_cfg(O.SOPR, 80, alpha=0.2, term=0.2)
    t1 = time.time()
    o = O.run(b, c, O.walk_cfg(42, tc0, tc1), k, wpv=2, ntypes=2, state=False, visits=True)
    print("oracle replay of %d walkers: %.1f s, %d steps" % (k, time.time() - t1, o["totals"]["steps"]), flush=True)
    np.testing.assert_array_equal(ref[:k], o["hashes"])
    # arrivals of the replayed walkers are a part of the GPU's counts
    assert np.all(o["visits"] <= visits)
