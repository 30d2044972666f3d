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


def head_bytes(pk, ne):
    """bytes of a packed block before its arc payload: header, beg chunks and payload, arc
    chunk headers"""
    if ne == 0:
        return 0
    return 4 * (4 + 4 * int(pk[1]) + int(pk[2]) + 4 * ((ne + 255) // 256))


def unsorted_graph():
    """lists in random order: gaps wrap modulo 2^32, a chunk's gap sums pass 2^32"""
    rng = np.random.default_rng(4)
    n = 600
    deg = rng.integers(0, 90, n)
    beg = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    csr = rng.integers(0, n, int(beg[-1])).astype(np.uint32)
    return beg, csr


CASES = [
    ("rmat-permuted", lambda: G.rmat(13, 60000, seed=3, permute=True), 16384, N2V, 1),
    ("directed-mixed", lambda: G.rmat(12, 30000, seed=8, undirected=False), 8192, MIXED, 2),
    ("giant-vertex", giant_graph, 4096, N2V, 1),
]


@pytest.mark.parametrize("name,make,block_bytes,types,wpv", CASES)
@pytest.mark.parametrize("how", ["all", "mixed", "staged", "pieces", "lanes", "ring", "ring-pieces"])
def test_packed_uploads_match_oracle(gpu, name, make, block_bytes, types, wpv, how):
    beg, csr = make()
    n = (len(beg) - 1) * wpv
    bl = Blocks(beg, csr, block_bytes)
    nb = bl.nblocks
    biggest = max(len(pack_block(*bl.arrays(b))) * 4 for b in range(nb))
    # "pieces": a staging buffer of the largest block's headers and packed offsets plus the
    # largest possible chunk (1088 B), so every payload crosses in pieces of a few chunks
    head = max(head_bytes(pack_block(*bl.arrays(b)), bl.desc[b].nedges) for b in range(nb))
    with Context() as ctx:
        ctx.set_blocks(bl)
        ctx.set_walk(types, 13)
        ctx.init_walkers(n, wpv, outputs=OUT_HASH | OUT_VISITS)
        if how == "lanes":   # two copy lanes, uploads and staged copies alternating
            ctx.set_copy_lanes(2)
        if how == "ring":            # the staging ring with room for whole blocks
            ctx.set_pack_staging(8 * biggest + 65536)
        elif how == "ring-pieces":   # the staging ring, blocks in pieces
            ctx.set_pack_staging(max(8 * head, 16384))
        else:
            ctx.set_pack_staging(head + 1088 if how == "pieces" else (2 if how == "lanes" else 1) * biggest)
        ctx.reserve_slots(nb)
        sent = 0
        for b in range(nb):
            if bl.desc[b].nedges == 0 or (how == "mixed" and b % 2):
                ctx.upload(b, b)
                sent += 4 * bl.desc[b].nedges
            elif how == "staged" or (how in ("lanes", "ring", "ring-pieces") and b % 3 == 1):
                sent += ctx.stage_packed(b, b)
            else:
                sent += ctx.upload_packed(b, b)
        ctx.wait_uploads()
        if how in ("staged", "lanes", "ring", "ring-pieces"):
            for b in range(nb):
                if bl.desc[b].nedges and (how == "staged" or b % 3 == 1):
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


UPLOAD_MODES = {
    "in place+worker": {},                                        # the default
    "ring+worker": {"GW_PACK_INPLACE": "0"},
    "in place, caller's thread": {"GW_UPLOAD_THREAD": "0"},
    "ring, caller's thread": {"GW_UPLOAD_THREAD": "0", "GW_PACK_INPLACE": "0"},
    "copy-stream decode+worker": {"GW_PACK_INLINE_DECODE": "1"},
}


@pytest.mark.parametrize("mode", sorted(UPLOAD_MODES))
def test_engine_upload_modes(gpu, mode, monkeypatch):
    """Block uploads enqueued by the context's upload worker or by the caller's thread,
    decoded on the decode stream through the staging ring or on the copy stream: the
    engine's walks equal the oracle's under a cap with evictions, staged copies, slots
    rewritten while runs may still read them, and the zero-copy tail."""
    for k, v in UPLOAD_MODES[mode].items():
        monkeypatch.setenv(k, v)
    beg, csr = G.rmat(14, 150000, seed=12, permute=True)
    n = len(beg) - 1
    b = E.PinnedArray(len(beg), np.uint64)
    c = E.PinnedArray(len(csr), np.uint32)
    b.a[:] = beg
    c.a[:] = csr
    block_bytes = 40000
    with Context() as ctx:
        ctx.set_blocks(Blocks(beg, csr, block_bytes))
        slot = ctx.slot_bytes()
    pk = E.PackedGraph(b.a, c.a, block_bytes)
    mem = n * (4 * 4 + 8) + 8 * slot + 65536 + 4 * (pk.info()["raw_bytes"] // max(1, pk.info()["nblocks"]))
    o = oracle(beg, csr, N2V, 5, n, 1)
    for sched, zt in ((E.SCHED_LEARNED, 0), (E.SCHED_PLAIN, block_bytes // 3072)):
        rep, h, _ = E.run_out_of_core(b.a, c.a, N2V, 5, n, memory_bytes=mem, block_bytes=block_bytes,
                                      scheduler=sched, zero_threshold=zt, hashes=True, packed=pk)
        print(mode, sched, {k: rep[k] for k in ("uploads", "staged", "nslots", "nblocks", "zc_touches", "h2d_bytes",
                                                "inplace_blocks", "ring_blocks", "inline_blocks")}, flush=True)
        assert rep["nslots"] < rep["nblocks"] and rep["uploads"] + rep["staged"] > rep["nblocks"]
        np.testing.assert_array_equal(h, o["hashes"])
        # the mode's decode path: every block in place or on the copy stream; through the
        # ring, every block whose headers fit a head area (the rest on the copy stream)
        blocks = rep["uploads"] + rep["staged"]
        if "copy-stream" in mode:
            assert rep["inline_blocks"] == blocks, rep
        elif "ring" in mode:
            assert rep["inplace_blocks"] == 0 and rep["ring_blocks"] > blocks // 2, rep
        else:
            assert rep["inplace_blocks"] == blocks, rep
    pk.close()
    b.close()
    c.close()


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
    mem = n * (4 * 4 + 8) + 8 * slot + 65536 + 4 * (info["raw_bytes"] // max(1, info["nblocks"]))
    out = {}
    for name, sched in (("plain", E.SCHED_PLAIN), ("staging", E.SCHED_STAGING)):
        for packed in (None, pk):
            rep, h, _ = E.run_out_of_core(b.a, c.a, N2V, 21, n, memory_bytes=mem, block_bytes=block_bytes,
                                          scheduler=sched, hashes=True, packed=packed)
            out[(name, packed is not None)] = (rep, h)
    # the same packed walk on two copy lanes (half of the staging buffer each at this cap)
    rep, h, _ = E.run_out_of_core(b.a, c.a, N2V, 21, n, memory_bytes=mem, block_bytes=block_bytes,
                                  scheduler=E.SCHED_STAGING, hashes=True, packed=pk, flags=E.TWO_LANES)
    assert rep["copy_lanes"] == 2
    out[("staging two lanes", True)] = (rep, h)
    for (name, packed), (rep, _) in out.items():
        print(name, "packed" if packed else "raw", {k: rep[k] for k in ("h2d_bytes", "h2d_raw_bytes", "uploads", "staged",
                                                                         "nslots", "nblocks", "copy_lanes")},
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


SLOT_CASES = [
    ("rmat-permuted", lambda: G.rmat(14, 120000, seed=6, permute=True), 65536),
    ("unsorted", unsorted_graph, 8192),
    ("giant-vertex", giant_graph, 4096),
    ("power-law-multi", lambda: G.power_law_directed(3000, 40000, seed=2, multi=True), 8192),
]


@pytest.mark.parametrize("name,make,block_bytes", SLOT_CASES)
def test_packed_slots_bit_exact(gpu, name, make, block_bytes, monkeypatch):
    """Every block decoded into its slot, uploaded or staged, in place, through the staging
    ring in one piece or in pieces, or on the copy stream, holds exactly the block's beg
    (global offsets) and csr."""
    beg, csr = make()
    bl = Blocks(beg, csr, block_bytes)
    nb = bl.nblocks
    pks = [pack_block(*bl.arrays(b)) for b in range(nb)]
    biggest = max(4 * len(p) for p in pks)
    head = max(head_bytes(pks[b], bl.desc[b].nedges) for b in range(nb))
    # one lane decodes through the staging ring when a block's headers fit an eighth of the
    # buffer: with room for whole blocks, with pieces of a fraction of a block (the ring
    # wraps within blocks), and with a buffer that fits some blocks' headers and not others
    # (ring and copy-stream decodes alternate)
    heads = sorted(head_bytes(pks[b], bl.desc[b].nedges) for b in range(nb) if bl.desc[b].nedges)
    ring_roomy = 8 * biggest + 65536
    ring_tight = max(8 * head, 16384)
    ring_mixed = max(8 * heads[len(heads) // 2], 16384)
    checked = 0
    paths = [0, 0, 0]
    for lanes, staging, inplace in ((1, biggest, True), (1, head + 1088, True), (2, biggest, True),
                                    (2, head + 1088, True), (1, ring_roomy, False), (1, ring_tight, False),
                                    (1, ring_mixed, False), (1, ring_mixed, True)):
        monkeypatch.setenv("GW_PACK_INPLACE", "1" if inplace else "0")
        with Context() as ctx:
            ctx.set_blocks(bl)
            ctx.set_walk(N2V, 3)
            ctx.init_walkers(len(beg) - 1, 1, outputs=OUT_HASH)
            ctx.set_copy_lanes(lanes)
            ctx.set_pack_staging(lanes * staging)
            ctx.reserve_slots(nb)
            for b in range(nb):
                if bl.desc[b].nedges == 0:
                    continue
                if b % 2:
                    ctx.stage_packed(b, b)
                else:
                    ctx.upload_packed(b, b)
            ctx.wait_uploads()
            for b in range(nb):
                if bl.desc[b].nedges == 0:
                    continue
                eb, ec = bl.arrays(b)
                gb, gc = ctx.read_slot(b, b)
                np.testing.assert_array_equal(gb, eb.astype(np.uint64), err_msg="%s block %d beg" % (name, b))
                np.testing.assert_array_equal(gc, ec, err_msg="%s block %d csr" % (name, b))
                checked += 1
            cnt = ctx.counters()
            up = ctx.upload_stats()
            paths = [x + y for x, y in zip(paths, up)]
        print(name, "lanes", lanes, "staging", staging, "in place", inplace, "blocks", nb, "h2d", cnt["h2d_bytes"],
              "raw", cnt["h2d_raw_bytes"], "decoded (in place, ring, copy stream)", up, flush=True)
    assert checked >= 8
    assert all(p > 0 for p in paths), paths   # every decode path was taken


def test_malformed_packed_block_rejected(gpu):
    """A packed block whose headers disagree with the block (wrong magic, payload offsets,
    an Elias-Fano width or upper-bit count out of range, a first vertex past the block) is
    refused on the host before anything is copied; the context stays usable."""
    beg, csr = G.rmat(12, 30000, seed=9, permute=True)
    bl = Blocks(beg, csr, 1 << 20)
    blk = next(b for b in range(bl.nblocks) if bl.desc[b].nedges > 512)
    pk = pack_block(*bl.arrays(blk))
    nv = bl.desc[blk].nverts
    ah = 4 + 4 * int(pk[1]) + int(pk[2])
    bad = []
    x = pk.copy(); x[0] ^= 1; bad.append(("magic", x))
    x = pk.copy(); x[ah + 4] += 1; bad.append(("chunk offset", x))
    x = pk.copy(); x[ah + 1] = (x[ah + 1] & ~np.uint32(0xFF)) | np.uint32(40); bad.append(("L above 31", x))
    x = pk.copy(); x[ah + 1] = (x[ah + 1] & np.uint32(0xFF)) | np.uint32(600 << 8); bad.append(("nz", x))
    x = pk.copy(); x[ah + 3] = nv; bad.append(("first vertex", x))
    x = pk[:-1].copy(); bad.append(("truncated", x))
    with Context() as ctx:
        ctx.set_blocks(bl)
        ctx.set_walk(N2V, 3)
        ctx.init_walkers(len(beg) - 1, 1, outputs=OUT_HASH)
        ctx.set_pack_staging(4 * len(pk))
        ctx.reserve_slots(1)
        for what, x in bad:
            st = E.lib().gw_upload_block_packed(ctx.h, 0, blk, E._p(x), 4 * len(x))
            assert st != 0, what
        ctx.upload_packed(0, blk)
        ctx.wait_uploads()
        gb, gc = ctx.read_slot(0, blk)
        eb, ec = bl.arrays(blk)
        np.testing.assert_array_equal(gb, eb.astype(np.uint64))
        np.testing.assert_array_equal(gc, ec)


@pytest.mark.parametrize("lanes", [2, 1])
def test_lanes_same_slot_rewrites(gpu, lanes):
    """Blocks written into one slot back to back (uploads and staged copies, no wait between
    them) leave the slot holding the last one exactly: with two copy lanes, a copy into a slot
    whose last copy may still run stays on that copy's lane; with one lane and the staging
    ring, the decodes of one stream write the slot in order."""
    beg, csr = G.rmat(14, 120000, seed=11, permute=True)
    bl = Blocks(beg, csr, 65536)
    nb = bl.nblocks
    blocks = [b for b in range(nb) if bl.desc[b].nedges > 0]
    assert len(blocks) >= 6
    biggest = max(len(pack_block(*bl.arrays(b))) * 4 for b in blocks)
    with Context() as ctx:
        ctx.set_blocks(bl)
        ctx.set_walk(N2V, 3)
        ctx.init_walkers(len(beg) - 1, 1, outputs=OUT_HASH)
        ctx.set_copy_lanes(lanes)
        ctx.set_pack_staging(2 * biggest if lanes == 2 else 8 * biggest + 65536)
        ctx.reserve_slots(2)
        # slot 0: six blocks in a row; slot 1: one block, sent between them
        for i, b in enumerate(blocks[:6]):
            if i % 2:
                ctx.stage_packed(0, b)
            else:
                ctx.upload_packed(0, b)
            if i == 2:
                ctx.upload_packed(1, blocks[-1])
        ctx.wait_uploads()
        for slot, b in ((0, blocks[5]), (1, blocks[-1])):
            gb, gc = ctx.read_slot(slot, b)
            eb, ec = bl.arrays(b)
            np.testing.assert_array_equal(gb, eb.astype(np.uint64))
            np.testing.assert_array_equal(gc, ec)
