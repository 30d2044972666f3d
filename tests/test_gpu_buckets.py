"""The bucket matrix W[prev_blk, cur_blk] of unfinished walkers (gw_bucket_counts; the
reference counts it with atomicAdd(&block_offset[prev_blk * ncblock + cur_blk]) at the
exit of gpu_update, gpu/gpuwalk.hpp:177-184, and scans it in `prefix`, :321-338) against an
independent count: the walker state read back with gw_read_state (cur, prev, hop,
status), each unfinished walker put in the bucket of its (prev, cur) blocks by a binary
search of the block table's cut points in numpy.  Checked after initialisation, after
runs with a part of the blocks resident (the count left by the run's fused bucket pass,
GW_RUN_BUCKETS, and the separate pass of a later call), for tables staged in LDS and read
from global memory, Node2vec and mixed Node2vec + SOSR batches with termination, and
once the walk is over (all zero)."""
import numpy as np
import pytest

import graphs as G
from gowalker_amd.engine import NODE2VEC, OUT_HASH, SOPR, Blocks, Context

pytestmark = pytest.mark.gpu


def histogram(ctx, bl):
    s = ctx.state()
    live = s[:, 3] == 0
    cut = bl.vcut[:-1].astype(np.int64)
    cb = np.searchsorted(cut, s[live, 0].astype(np.int64), side="right") - 1
    pb = np.searchsorted(cut, s[live, 1].astype(np.int64), side="right") - 1
    H = np.zeros((bl.nblocks, bl.nblocks), dtype=np.uint64)
    np.add.at(H, (pb, cb), 1)
    return H, int(live.sum())


def check(ctx, bl, what):
    W = ctx.bucket_counts()
    H, live = histogram(ctx, bl)
    print(what, "walkers left", live, "buckets used", int((H > 0).sum()), flush=True)
    np.testing.assert_array_equal(W, H, err_msg=what)
    assert int(W.sum()) == live == ctx.walkers_left()
    return live


# (graph, block bytes, mixed batch): degrees stay below blocksize/4, so no single-vertex
# block is truncated (engine/cache.hpp:170-172); 1024-byte blocks give > 128 blocks (table
# read from global memory), the others <= 128 (table staged in LDS)
CASES = [("undirected", 16384, False), ("undirected", 1024, False), ("directed", 4096, True),
         ("undirected", 1024, True)]


@pytest.mark.parametrize("graph,block_bytes,mixed", CASES)
def test_bucket_matrix_matches_walker_histogram(gpu, graph, block_bytes, mixed):
    if graph == "undirected":
        beg, csr = G.random_undirected(8000, 24000, seed=5)
    else:
        beg, csr = G.rmat(13, 30000, seed=8, undirected=False)   # dead ends too
    n = len(beg) - 1
    assert int(np.diff(beg.astype(np.int64)).max()) <= block_bytes // 4
    bl = Blocks(beg, csr, block_bytes)
    nb = bl.nblocks
    assert nb >= 8 and (nb > 128) == (block_bytes == 1024)
    types = [dict(algo=NODE2VEC, length=25, p=0.5, q=2.0)]
    if mixed:
        types.append(dict(algo=SOPR, length=25, alpha=0.2, term=0.15))
    wpv = 2 if mixed else 1
    rng = np.random.default_rng(block_bytes + int(mixed))
    with Context() as ctx:
        ctx.set_blocks(bl)
        ctx.set_walk(types, 31)
        ctx.init_walkers(n * wpv, wpv, outputs=OUT_HASH)
        assert check(ctx, bl, "after init") == n * wpv
        k = max(2, nb // 3)
        ctx.reserve_slots(k)
        for it in range(6):
            # a different part of the table resident each time; the fused pass of the run
            # leaves W, then the separate pass of the next call counts it again
            resident = sorted(rng.choice(nb, k, replace=False).tolist())
            for s in range(k):
                ctx.evict(s)
            for s, b in enumerate(resident):
                ctx.upload(s, b)
            ctx.wait_uploads()
            st = ctx.run(buckets=(it % 2 == 0))
            print("run", it, "resident", resident, "walkers run", st["walkers_run"], flush=True)
            check(ctx, bl, "after run %d (%s pass)" % (it, "fused" if it % 2 == 0 else "separate"))
        ctx.reserve_slots(nb)
        for b in range(nb):
            ctx.upload(b, b)
        ctx.wait_uploads()
        assert ctx.run(buckets=True)["walkers_left"] == 0
        assert check(ctx, bl, "after the walk") == 0


@pytest.mark.parametrize("variant,window", [(0, False), (2, True)])
def test_short_worklist_takes_rounds_with_the_same_walks(gpu, variant, window):
    """A worklist shorter than a run's runnable walkers (gw_set_worklist_cap) makes gw_run
    take further filter + step rounds for the walkers it could not hold.  After every run,
    with a part of the blocks resident each time, the walker state, the run's step / walker
    counts and the bucket matrix equal those of the full-size list, and the final hashes
    too; the short lists took extra rounds and the full one none."""
    beg, csr = G.rmat(13, 30000, seed=8, undirected=False)
    n = len(beg) - 1
    bl = Blocks(beg, csr, 4096)
    nb = bl.nblocks
    types = [dict(algo=NODE2VEC, length=25, p=0.5, q=2.0), dict(algo=SOPR, length=25, alpha=0.2, term=0.15)]
    logs, finals = [], []
    for cap in (0, 64, 1000):
        rng = np.random.default_rng(77)
        log = []
        with Context() as ctx:
            ctx.set_worklist_cap(cap)
            ctx.set_blocks(bl)
            ctx.set_walk(types, 31)
            ctx.init_walkers(n * 2, 2, outputs=OUT_HASH)
            wcap, rounds0 = ctx.worklist_stats()
            assert wcap == (cap or n * 2)
            k = max(2, nb // 3)
            ctx.reserve_slots(k)
            for it in range(6):
                resident = sorted(rng.choice(nb, k, replace=False).tolist())
                for s in range(k):
                    ctx.evict(s)
                for s, b in enumerate(resident):
                    ctx.upload(s, b)
                ctx.wait_uploads()
                st = ctx.run(buckets=True, variant=variant, window=window)
                log.append((st["steps"], st["trials"], st["walkers_run"], st["walkers_left"]))
                check(ctx, bl, "cap %d run %d" % (cap, it))
                log.append(ctx.state().copy())
            ctx.reserve_slots(nb)
            for b in range(nb):
                ctx.upload(b, b)
            ctx.wait_uploads()
            assert ctx.run(buckets=True, variant=variant, window=window)["walkers_left"] == 0
            _, rounds = ctx.worklist_stats()
            print("cap", cap, "capacity", wcap, "extra rounds", rounds - rounds0, flush=True)
            assert (rounds - rounds0 > 0) == (cap != 0)
            finals.append(ctx.hashes())
        logs.append(log)
    for log in logs[1:]:
        assert len(log) == len(logs[0])
        for x, y in zip(logs[0], log):
            if isinstance(x, tuple):
                assert x == y
            else:
                np.testing.assert_array_equal(x, y)
    for h in finals[1:]:
        np.testing.assert_array_equal(finals[0], h)


@pytest.mark.parametrize("mixed", [False, True])
def test_walkers_whose_previous_block_was_evicted(gpu, mixed):
    """After an eviction, walkers whose current block is resident and previous one is not are
    run by the lane kernel (variant 2) until a step needs the previous vertex's list (a
    Node2vec class test inside the membership band, an SOSR proposal from it); such a step
    stops, unchanged, and the walker is not picked up again before that block is resident.
    Variant 0 reads both rows at pickup and runs only walkers with both blocks resident.  With
    the same resident sets in turn (half the blocks, then the other half twice, then all),
    the lane kernel steps more in the second run, neither kernel finds work in the third, the
    bucket matrix is the walker-state histogram after every run, and the finished walks,
    steps and trials are the oracle's for both kernels."""
    import _oracle as O

    beg, csr = G.rmat(13, 40000, seed=8, undirected=True, permute=True)
    n = len(beg) - 1
    bl = Blocks(beg, csr, 16384)
    nb = bl.nblocks
    assert nb >= 6
    t0 = dict(algo=NODE2VEC, length=20, p=0.5, q=2.0)
    if mixed:
        types = [t0, dict(algo=SOPR, length=20, alpha=0.2, term=0.15)]
        cfg = O.walk_cfg(5, O.type_cfg(O.N2V, 20, p=0.5, q=2.0), O.type_cfg(O.SOPR, 20, alpha=0.2, term=0.15))
        wpv, ntypes = 2, 2
    else:
        types = [t0]
        cfg = O.walk_cfg(5, O.type_cfg(O.N2V, 20, p=0.5, q=2.0))
        wpv, ntypes = 1, 1
    nw = n * wpv
    o = O.run(beg, csr, cfg, nw, wpv=wpv, ntypes=ntypes, state=False)
    half = nb // 2
    A, B = list(range(half)), list(range(half, nb))
    out = {}
    for variant, window in ((0, False), (2, True)):
        with Context() as ctx:
            ctx.set_blocks(bl)
            ctx.set_walk(types, 5)
            ctx.init_walkers(nw, wpv, outputs=OUT_HASH)
            ctx.reserve_slots(max(len(A), len(B)))
            runs = []
            for it, resident in enumerate((A, B, B)):
                if it < 2:
                    for s in range(max(len(A), len(B))):
                        ctx.evict(s)
                    for s, b in enumerate(resident):
                        ctx.upload(s, b)
                    ctx.wait_uploads()
                runs.append(ctx.run(buckets=True, variant=variant, window=window))
                check(ctx, bl, "variant %d run %d" % (variant, it))
            ctx.reserve_slots(nb)
            for b in range(nb):
                ctx.upload(b, b)
            ctx.wait_uploads()
            runs.append(ctx.run(buckets=True, variant=variant, window=window))
            assert runs[-1]["walkers_left"] == 0
            print("variant", variant, [(r["steps"], r["trials"], r["walkers_run"]) for r in runs], flush=True)
            out[variant] = (runs, ctx.hashes())
    for variant, (runs, h) in out.items():
        np.testing.assert_array_equal(h, o["hashes"], err_msg="variant %d" % variant)
        assert sum(r["steps"] for r in runs) == o["totals"]["steps"]
        assert sum(r["trials"] for r in runs) == o["totals"]["trials"]
        assert runs[2]["steps"] == 0 and runs[2]["walkers_run"] == 0   # nothing new is runnable
    # the second run: the lane kernel also steps walkers that came from the evicted half
    assert out[2][0][1]["steps"] > out[0][0][1]["steps"]
