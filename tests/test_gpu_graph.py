"""Runs replayed from captured HIP graphs (gw_run): a sequence of out-of-core iterations --
the resident blocks change between runs, every run filters the runnable walkers, steps them
and counts the bucket matrix of the walkers left -- leaves walker state, counters and bucket
matrix after every run identical to the same sequence launched kernel by kernel (the
default; GW_RUN_GRAPH=1 replays graphs); the graph is captured once and replayed for the runs whose launches
recur; the finished walks equal the oracle's."""
import numpy as np
import pytest

import _oracle as O
import graphs as G
from gowalker_amd.engine import NODE2VEC, OUT_HASH, OUT_VISITS, Blocks, Context

pytestmark = pytest.mark.gpu

TYPES = [dict(algo=NODE2VEC, length=30, p=0.5, q=2.0)]


def iterate(monkeypatch, graphs, beg, csr, block_bytes, launch):
    monkeypatch.setenv("GW_RUN_GRAPH", "1" if graphs else "0")
    n = len(beg) - 1
    bl = Blocks(beg, csr, block_bytes)
    nb = bl.nblocks
    rng = np.random.default_rng(3)
    snaps = []
    with Context() as ctx:
        ctx.set_blocks(bl)
        ctx.set_walk(TYPES, 9)
        ctx.init_walkers(n, 1, outputs=OUT_HASH | OUT_VISITS)
        ctx.reserve_slots(nb)
        for it in range(14):
            for s in range(nb):
                ctx.evict(s)
            for s, b in enumerate(sorted(rng.choice(nb, 4, replace=False))):
                ctx.upload(s, int(b))
            ctx.wait_uploads()
            st = ctx.run(buckets=True, **launch)
            assert st["kernel_ms"] > 0 and st["filter_ms"] > 0   # events, or device-clock stamps in graphs
            cnt = {k: st[k] for k in ("steps", "trials", "probes", "walkers_run", "walkers_ended", "walkers_left")}
            snaps.append((cnt, ctx.state().copy(), ctx.bucket_counts().copy()))
        ctx.load_all()   # every block resident: the rest of the walk in one run
        st = ctx.run(buckets=True, **launch)
        assert st["walkers_left"] == 0
        snaps.append(({"steps": st["steps"]}, ctx.state().copy(), ctx.hashes().copy()))
        stats = ctx.graph_stats()
    return snaps, stats, nb


@pytest.mark.parametrize("launch", [dict(variant=2, interp=True, window=True), dict(variant=0)],
                         ids=["lane-kernel", "per-step-kernel"])
def test_graph_replays_match_eager_runs(gpu, monkeypatch, launch):
    beg, csr = G.rmat(13, 60000, seed=21, permute=True)
    block_bytes = 16384
    eager, est, nb = iterate(monkeypatch, False, beg, csr, block_bytes, launch)
    graph, gst, _ = iterate(monkeypatch, True, beg, csr, block_bytes, launch)
    print("blocks", nb, "eager (replays, captures)", est, "graphs", gst, flush=True)
    assert est == (0, 0)
    assert gst[0] >= 10 and gst[1] <= 4   # one sequence for the partial runs, one for the last
    for i, (e, g) in enumerate(zip(eager, graph)):
        assert e[0] == g[0], (i, e[0], g[0])
        np.testing.assert_array_equal(e[1], g[1], err_msg="state after run %d" % i)
        np.testing.assert_array_equal(e[2], g[2], err_msg="buckets / hashes after run %d" % i)
    assert sum(s[0]["steps"] for s in graph[:-1]) > 0   # the partial runs did step walkers
    o = O.run(beg, csr, O.walk_cfg(9, O.type_cfg(O.N2V, 30, p=0.5, q=2.0)), len(beg) - 1)
    np.testing.assert_array_equal(graph[-1][2], o["hashes"])
