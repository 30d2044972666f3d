"""Packed uploads (gw_upload_block_packed / gw_stage_block_packed; the engine with a
gwh_packed graph): the GPU decodes each block's transfer encoding into its HBM slot, so
walks must equal the oracle's bit for bit -- with every block packed, with packed and raw
blocks mixed, with staged packed copies, for a truncated giant vertex, vertex-limited
blocks, directed multigraphs and mixed Node2vec + SOSR batches -- while the bytes that cross
PCIe drop below the blocks' size in the reference's format."""
import numpy as np
import pytest

import _oracle as O
import graphs as G
from gowalker_amd import engine as E
from gowalker_amd.engine import NODE2VEC, OUT_HASH, OUT_VISITS, SOPR, Blocks, Context, pack_block

pytestmark = pytest.mark.gpu

N2V = [dict(algo=NODE2VEC, length=20, p=0.5, q=2.0)]
MIXED = [dict(algo=NODE2VEC, length=20, p=1.0, q=0.25), dict(algo=SOPR, length=20, alpha=0.2, term=0.2)]


def oracle(beg, csr, types, seed, n, wpv):
    t0 = types[0]
    tc0 = O.type_cfg(O.N2V, t0["length"], p=t0["p"], q=t0["q"])
    if len(types) == 1:
        return O.run(beg, csr, O.walk_cfg(seed, tc0), n, wpv=wpv, visits=True)
    t1 = types[1]
    tc1 = O.type_cfg(O.SOPR, t1["length"], alpha=t1["alpha"], term=t1["term"])
    return O.run(beg, csr, O.walk_cfg(seed, tc0, tc1), n, wpv=wpv, ntypes=2, visits=True)


def effective_graph(beg, csr, blocks):
    """The graph the engine walks: a single-vertex block longer than blocksize/4 edges is
    cut to its first blocksize/4 targets (engine/cache.hpp:170-172)."""
    deg = np.diff(beg.astype(np.int64))
    for i in range(blocks.nblocks):
        d = blocks.desc[i]
        if d.nverts == 1:
            deg[d.start_vert] = d.nedges
    nb = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    out = np.concatenate([csr[int(beg[v]):int(beg[v]) + int(deg[v])] for v in range(len(deg))])
    return nb, out.astype(np.uint32)


def giant_graph():
    n = 3000
    src = np.concatenate([np.zeros(5000, np.int64), np.arange(1, n, dtype=np.int64)])
    dst = np.concatenate([np.arange(1, 5001, dtype=np.int64) % n, (np.arange(1, n) * 7) % n])
    return G.from_edges(n, src, dst, undirected=True, dedup=True, drop_self=True)


CASES = [
    ("rmat-permuted", lambda: G.rmat(13, 60000, seed=3, permute=True), 16384, N2V, 1),
    ("directed-mixed", lambda: G.rmat(12, 30000, seed=8, undirected=False), 8192, MIXED, 2),
    ("giant-vertex", giant_graph, 4096, N2V, 1),
]


@pytest.mark.parametrize("name,make,block_bytes,types,wpv", CASES)
@pytest.mark.parametrize("how", ["all", "mixed", "staged", "pieces"])
def test_packed_uploads_match_oracle(gpu, name, make, block_bytes, types, wpv, how):
    beg, csr = make()
    n = (len(beg) - 1) * wpv
    bl = Blocks(beg, csr, block_bytes)
    nb = bl.nblocks
    biggest = max(len(pack_block(*bl.arrays(b))) * 4 for b in range(nb))
    # "pieces": a staging buffer of the largest block's headers and list starts plus 1 KiB,
    # so every payload crosses in pieces of a few chunks
    head = max(16 * ((bl.desc[b].nedges + 255) // 256) + 4 * bl.desc[b].nverts for b in range(nb))
    with Context() as ctx:
        ctx.set_blocks(bl)
        ctx.set_walk(types, 13)
        ctx.init_walkers(n, wpv, outputs=OUT_HASH | OUT_VISITS)
        ctx.set_pack_staging(head + 1024 if how == "pieces" else biggest)
        ctx.reserve_slots(nb)
        sent = 0
        for b in range(nb):
            if bl.desc[b].nedges == 0 or (how == "mixed" and b % 2):
                ctx.upload(b, b)
                sent += 4 * bl.desc[b].nedges
            elif how == "staged":
                sent += ctx.stage_packed(b, b)
            else:
                sent += ctx.upload_packed(b, b)
        ctx.wait_uploads()
        if how == "staged":
            for b in range(nb):
                if bl.desc[b].nedges:
                    ctx.commit(b)
        st = ctx.run()
        assert st["walkers_left"] == 0
        h, v = ctx.hashes(), ctx.visits()
        cnt = ctx.counters()
    eb, ec = effective_graph(beg, csr, bl)
    o = oracle(eb, ec, types, 13, n, wpv)
    print(name, how, "blocks", nb, "steps", st["steps"], "sent", sent, "raw csr", 4 * len(csr),
          "h2d", cnt["h2d_bytes"], "raw", cnt["h2d_raw_bytes"], flush=True)
    np.testing.assert_array_equal(h, o["hashes"])
    np.testing.assert_array_equal(v, o["visits"])
    assert st["steps"] == o["totals"]["steps"]
    assert cnt["h2d_raw_bytes"] == sum(8 * (bl.desc[b].nverts + 1) + 4 * bl.desc[b].nedges for b in range(nb))
    assert cnt["h2d_bytes"] < cnt["h2d_raw_bytes"]


def test_engine_with_packed_graph(gpu):
    """The out-of-core engine with a packed graph: the same walks as without it and as the
    oracle, under a cap that forces evictions and staged copies, with fewer PCIe bytes."""
    beg, csr = G.rmat(14, 150000, seed=5, permute=True)
    n = len(beg) - 1
    b = E.PinnedArray(len(beg), np.uint64)
    c = E.PinnedArray(len(csr), np.uint32)
    b.a[:] = beg
    c.a[:] = csr
    block_bytes = 40000
    with Context() as ctx:
        bl = Blocks(beg, csr, block_bytes)
        ctx.set_blocks(bl)
        slot = ctx.slot_bytes()
    pk = E.PackedGraph(b.a, c.a, block_bytes)
    info = pk.info()
    print("packed", info, flush=True)
    assert info["blocks_packed"] > 0 and info["packed_bytes"] < info["raw_bytes"]
    mem = n * (5 * 4 + 8) + 8 * slot + 65536 + 4 * (info["raw_bytes"] // max(1, info["nblocks"]))
    out = {}
    for name, sched in (("plain", E.SCHED_PLAIN), ("staging", E.SCHED_STAGING)):
        for packed in (None, pk):
            rep, h, _ = E.run_out_of_core(b.a, c.a, N2V, 21, n, memory_bytes=mem, block_bytes=block_bytes,
                                          scheduler=sched, hashes=True, packed=packed)
            out[(name, packed is not None)] = (rep, h)
            print(name, "packed" if packed else "raw", {k: rep[k] for k in ("h2d_bytes", "h2d_raw_bytes", "uploads",
                                                                             "staged", "nslots", "nblocks")},
                  flush=True)
    o = oracle(beg, csr, N2V, 21, n, 1)
    for (name, packed), (rep, h) in out.items():
        np.testing.assert_array_equal(h, o["hashes"], err_msg="%s packed=%s" % (name, packed))
        assert rep["nslots"] < rep["nblocks"]
    # packed uploads send fewer bytes than the blocks' size in the reference's format; with
    # as many slots, the plain scheduler makes the same decisions, so the same blocks move
    rr, rp = out[("plain", False)][0], out[("plain", True)][0]
    assert rr["h2d_bytes"] == rr["h2d_raw_bytes"]
    assert rp["h2d_bytes"] < 0.9 * rp["h2d_raw_bytes"]
    if rp["nslots"] == rr["nslots"]:
        assert rp["h2d_raw_bytes"] == rr["h2d_bytes"]
    pk.close()
    b.close()
    c.close()
