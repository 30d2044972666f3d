"""Bit-exact parity of the gfx950 walk kernels (through libgw.so's C ABI) with the CPU
oracle: same graph, same seed => identical paths, path hashes, final states, per-vertex
visit counts and step/trial/probe counters."""
import sys

import numpy as np
import pytest

import _oracle as O
import graphs as G
from gowalker_amd.engine import NODE2VEC, OUT_HASH, OUT_PATHS, OUT_VISITS, SOPR, Blocks, Context


def oracle_types(types, search=0):
    ts = []
    for t in types:
        if t.get("algo", NODE2VEC) == NODE2VEC:
            ts.append(O.type_cfg(O.N2V, t["length"], p=t["p"], q=t["q"], term=t.get("term", 0.0), search=search))
        else:
            ts.append(O.type_cfg(O.SOPR, t["length"], alpha=t.get("alpha", 0.2), term=t.get("term", 0.0),
                                 sampler=t.get("sampler", 0), search=search))
    return ts


def gpu_run(beg, csr, types, seed, n, wpv=1, block_bytes=1 << 30, nslots=None, sources=None, launch=None):
    bl = Blocks(beg, csr, block_bytes)
    with Context() as ctx:
        ctx.set_blocks(bl)
        if nslots is None:
            ctx.load_all()
        ctx.set_walk(types, seed)
        ctx.init_walkers(n, wpv, sources=sources, outputs=OUT_HASH | OUT_PATHS | OUT_VISITS)
        tot = dict(steps=0, trials=0, probes=0, mix_steps=0)
        runs = 0
        if nslots is None:
            st = ctx.run(**(launch or {}))
            runs = 1
            for k in tot:
                tot[k] += st[k]
            assert st["walkers_left"] == 0
        else:
            # out-of-core rehearsal: resident window of nslots blocks slides over the
            # bucket matrix until every walker is done (walks are schedule-independent)
            ctx.reserve_slots(nslots)
            while True:
                W = ctx.bucket_counts()
                if W.sum() == 0:
                    break
                a, b = np.unravel_index(np.argmax(W), W.shape)
                want = [int(a)] if a == b else [int(a), int(b)]
                score = W.sum(axis=0) + W.sum(axis=1)
                for x in np.argsort(-score, kind="stable"):
                    if len(want) >= nslots:
                        break
                    if int(x) not in want:
                        want.append(int(x))
                for s, blk in enumerate(want):
                    ctx.upload(s, blk)
                ctx.wait_uploads()
                st = ctx.run(**(launch or {}))
                runs += 1
                for k in tot:
                    tot[k] += st[k]
                assert st["walkers_run"] > 0
        out = dict(paths=ctx.paths(), hashes=ctx.hashes(), state=ctx.state(), visits=ctx.visits(), totals=tot,
                   runs=runs, counters=ctx.counters())
    return out


def check_case(beg, csr, types, seed, n, wpv=1, **kw):
    g = gpu_run(beg, csr, types, seed, n, wpv, **kw)
    L = kw.get("launch") or {}
    ts = oracle_types(types, search=(1 if L.get("interp") else 0) | (2 if L.get("window") else 0))
    cfg = O.walk_cfg(seed, *ts)
    o = O.run(beg, csr, cfg, n, wpv=wpv, sources=kw.get("sources"), ntypes=len(types), paths=True, visits=True)
    print("case", types, "n", n, "gpu totals", g["totals"], "oracle totals", o["totals"], "runs", g["runs"])
    sys.stdout.flush()
    np.testing.assert_array_equal(g["state"], o["state"])
    np.testing.assert_array_equal(g["paths"], o["paths"])
    np.testing.assert_array_equal(g["hashes"], o["hashes"])
    np.testing.assert_array_equal(g["visits"], o["visits"])
    assert g["totals"] == o["totals"]
    return g, o


pytestmark = pytest.mark.gpu


def test_n2v_single_block(gpu):
    beg, csr = G.random_undirected(2000, 12000, seed=3)
    check_case(beg, csr, [dict(algo=NODE2VEC, length=20, p=2.0, q=0.5)], 42, 2000)


def test_n2v_multi_block_resident(gpu):
    beg, csr = G.rmat(10, 8000, seed=5, permute=True)
    g, o = check_case(beg, csr, [dict(algo=NODE2VEC, length=25, p=0.5, q=2.0)], 7, 1024, block_bytes=4096)
    assert o["totals"]["probes"] > 0


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_sopr_envelope_sampler(gpu, variant):
    # multigraph lists; the envelope sampler switches between rejection and the mixture
    beg, csr = G.power_law_directed(3000, 40000, seed=2)
    g, o = check_case(beg, csr, [dict(algo=SOPR, length=30, alpha=0.2, term=0.15, sampler=1)], 42, 6000, wpv=2,
                      launch=dict(variant=variant))
    assert 0 < o["totals"]["mix_steps"] < o["totals"]["steps"]


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_sopr_proposal_sampler(gpu, variant):
    beg, csr = G.power_law_directed(3000, 40000, seed=2, multi=False)
    g, o = check_case(beg, csr, [dict(algo=SOPR, length=30, alpha=0.2, term=0.15)], 42, 6000, wpv=2,
                      launch=dict(variant=variant))
    assert o["totals"]["mix_steps"] == 0 and o["totals"]["probes"] > 0


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_node2vec_both_variants(gpu, variant):
    beg, csr = G.rmat(12, 30000, seed=6, permute=True)
    check_case(beg, csr, [dict(algo=NODE2VEC, length=30, p=0.5, q=2.0)], 8, 4096, block_bytes=8192,
               launch=dict(variant=variant))


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_mixed_batch(gpu, variant):
    beg, csr = G.rmat(11, 20000, seed=9, undirected=False)
    types = [dict(algo=NODE2VEC, length=20, p=1.0, q=0.25), dict(algo=SOPR, length=20, alpha=0.2, term=0.2)]
    check_case(beg, csr, types, 11, 4096, wpv=2, block_bytes=8192, launch=dict(variant=variant))


@pytest.mark.parametrize("name", ["triangle", "star", "path", "bridge", "dag_dead_end"])
def test_tiny_graphs(gpu, name):
    beg, csr = getattr(G, name)()
    n = len(beg) - 1
    for variant in (0, 1, 2):
        L = dict(variant=variant)
        check_case(beg, csr, [dict(algo=NODE2VEC, length=12, p=0.5, q=2.0)], 1, n * 8, wpv=8, launch=L)
        check_case(beg, csr, [dict(algo=SOPR, length=12, alpha=0.3, term=0.1)], 2, n * 8, wpv=8, launch=L)
        check_case(beg, csr, [dict(algo=SOPR, length=12, alpha=0.3, term=0.1, sampler=1)], 3, n * 8, wpv=8,
                   launch=L)


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_out_of_core_window(gpu, variant):
    beg, csr = G.rmat(10, 8000, seed=4, permute=True)
    g, o = check_case(beg, csr, [dict(algo=NODE2VEC, length=20, p=0.5, q=2.0)], 5, 1024, block_bytes=4096,
                      nslots=3, launch=dict(variant=variant))
    assert g["runs"] > 1
    assert g["counters"]["h2d_bytes"] > 0


def test_sources_and_launch_shapes(gpu):
    beg, csr = G.random_undirected(500, 3000, seed=8)
    src = np.random.default_rng(0).integers(0, 500, 3000).astype(np.uint32)
    types = [dict(algo=NODE2VEC, length=15, p=2.0, q=0.5)]
    for launch in (dict(block=64, waves_per_cu=1), dict(block=1024, waves_per_cu=32),
                   dict(block=64, waves_per_cu=3, variant=1), dict(block=256, variant=1),
                   dict(block=64, waves_per_cu=1, variant=2), dict(block=1024, variant=2)):
        check_case(beg, csr, types, 3, 3000, sources=src, launch=launch)


def test_autotune_keeps_walks_bit_identical(gpu):
    # the tuner runs every candidate shape on a prefix of the walkers and restores their
    # state; the walks that follow with the tuned shape must still equal the oracle's
    beg, csr = G.rmat(12, 30000, seed=12, permute=True)
    types = [dict(algo=NODE2VEC, length=25, p=0.5, q=2.0)]
    n = 8192
    bl = Blocks(beg, csr, 1 << 30)
    with Context() as ctx:
        ctx.set_blocks(bl)
        ctx.load_all()
        ctx.set_walk(types, 21)
        ctx.init_walkers(n, 2, outputs=OUT_HASH | OUT_VISITS)
        best, ms = ctx.autotune(2048)
        print("autotune best", best, "ms", ms)
        assert best["block"] in (64, 128, 256, 512, 1024) and ms > 0
        assert ctx.counters()["steps"] == 0
        st = ctx.run(block=0)
        assert st["walkers_left"] == 0
        hashes, state, visits = ctx.hashes(), ctx.state(), ctx.visits()
    o = O.run(beg, csr, O.walk_cfg(21, *oracle_types(types)), n, wpv=2, visits=True)
    np.testing.assert_array_equal(state, o["state"])
    np.testing.assert_array_equal(hashes, o["hashes"])
    np.testing.assert_array_equal(visits, o["visits"])


@pytest.mark.parametrize("graph", ["permuted", "skewed"])
def test_interpolation_search_order(gpu, graph):
    # membership by interpolation: walks identical, probe counts equal the oracle's
    # interpolation order; Node2vec, both SOPR samplers, resident and out-of-core windows
    if graph == "permuted":
        beg, csr = G.rmat(12, 40000, seed=31, permute=True)
    else:
        beg, csr = G.rmat(12, 40000, seed=31, permute=False)
    for variant in (0, 2):
        L = dict(interp=True, variant=variant)
        check_case(beg, csr, [dict(algo=NODE2VEC, length=20, p=0.5, q=2.0)], 3, 4096, launch=L)
        check_case(beg, csr, [dict(algo=SOPR, length=20, alpha=0.2, term=0.1)], 4, 4096, launch=L)
        check_case(beg, csr, [dict(algo=SOPR, length=20, alpha=0.2, term=0.1, sampler=1)], 5, 4096, launch=L)
        check_case(beg, csr, [dict(algo=NODE2VEC, length=20, p=0.5, q=2.0)], 6, 2048, block_bytes=8192, nslots=3,
                   launch=L)


@pytest.mark.parametrize("graph,interp", [("permuted", True), ("skewed", True), ("skewed", False)])
def test_window_search_order(gpu, graph, interp):
    # membership by aligned 64-byte windows (bisected or interpolated guesses): walks
    # identical, probe counts equal the oracle's window order, which places the windows
    # by the edge array's alignment -- slots keep each block's offset modulo 64 bytes, in
    # one resident block and in out-of-core windows of 8 KB blocks (odd start edges)
    beg, csr = G.rmat(12, 40000, seed=31, permute=graph == "permuted")
    bl = Blocks(beg, csr, 8192)
    assert any(int(bl.ecut[b]) % 16 for b in range(bl.nblocks))
    for variant in (0, 2):
        L = dict(interp=interp, window=True, variant=variant)
        check_case(beg, csr, [dict(algo=NODE2VEC, length=20, p=0.5, q=2.0)], 3, 4096, launch=L)
        check_case(beg, csr, [dict(algo=SOPR, length=20, alpha=0.2, term=0.1)], 4, 4096, launch=L)
        check_case(beg, csr, [dict(algo=SOPR, length=20, alpha=0.2, term=0.1, sampler=1)], 5, 4096, launch=L)
        check_case(beg, csr, [dict(algo=NODE2VEC, length=20, p=0.5, q=2.0)], 6, 2048, block_bytes=8192, nslots=3,
                   launch=L)
        check_case(beg, csr, [dict(algo=SOPR, length=20, alpha=0.3, term=0.1)], 7, 2048, block_bytes=8192, nslots=3,
                   launch=L)


@pytest.mark.parametrize("interp", [False, True])
def test_window_trial_skips(gpu, interp):
    # the window lane kernel rejects a Node2vec trial without reading its candidate when
    # only x == prev could accept it and the step's earlier reads place prev's positions
    # elsewhere: walks and trial / probe counts equal the oracle's on directed graphs (prev
    # often not in N(cur)), multigraphs (prev repeated in N(cur)), tiny graphs, both
    # threshold orders and mixed walker types
    L = dict(variant=2, window=True, interp=interp)
    for multi in (False, True):
        beg, csr = G.power_law_directed(3000, 40000, seed=2, multi=multi)
        for p, q in ((0.5, 2.0), (2.0, 0.25), (0.25, 4.0)):
            check_case(beg, csr, [dict(algo=NODE2VEC, length=20, p=p, q=q)], 12, 6000, wpv=2, launch=L)
    for name in ("star", "bridge", "triangle"):
        beg, csr = getattr(G, name)()
        n = len(beg) - 1
        check_case(beg, csr, [dict(algo=NODE2VEC, length=12, p=0.5, q=2.0)], 13, n * 8, wpv=8, launch=L)
    beg, csr = G.rmat(11, 20000, seed=9, undirected=False)
    types = [dict(algo=NODE2VEC, length=20, p=0.5, q=2.0), dict(algo=SOPR, length=20, alpha=0.2, term=0.2)]
    check_case(beg, csr, types, 14, 4096, wpv=2, block_bytes=8192, launch=L)


@pytest.mark.parametrize("variant", [0, 2])
def test_block_tables_in_lds_and_global(gpu, variant):
    # up to 128 blocks the step kernels stage the table, its cut points and a vertex->block
    # lookup table in LDS; beyond that they read the table from global memory.  Degrees
    # stay below blocksize/4, so no single-vertex block is truncated (engine/cache.hpp:170-172)
    beg, csr = G.random_undirected(8000, 24000, seed=17)
    for block_bytes in (16384, 1024):
        bl = Blocks(beg, csr, block_bytes)
        print("block_bytes", block_bytes, "nblocks", bl.nblocks)
        assert int(np.diff(beg.astype(np.int64)).max()) <= block_bytes // 4
        assert (bl.nblocks <= 128) == (block_bytes == 16384)
        check_case(beg, csr, [dict(algo=NODE2VEC, length=25, p=0.5, q=2.0)], 19, 4096, block_bytes=block_bytes,
                   launch=dict(variant=variant, interp=True))
        check_case(beg, csr, [dict(algo=SOPR, length=25, alpha=0.3, term=0.1)], 20, 4096, block_bytes=block_bytes,
                   nslots=8, launch=dict(variant=variant))
