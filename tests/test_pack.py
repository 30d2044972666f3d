"""The transfer encoding of a block's adjacency (gw_pack_block, host code, no GPU) decoded
by a numpy restatement of its documented layout (include/gw.h): chunk headers {payload word
offset, bit width, carry, first vertex}, per-vertex list starts, bit-packed differences.
The decoded arcs must equal the block's csr exactly, for chunks that start inside lists,
empty lists, a truncated giant vertex, duplicate targets (zero differences, width 0),
unsorted lists (differences that wrap modulo 2^32, width 32) and vertex-limited blocks; the
packed size is what the layout predicts.  The GPU decoder is checked against the same
blocks through walks (tests/test_gpu_pack.py)."""
import numpy as np
import pytest

import graphs as G
from gowalker_amd.engine import Blocks, pack_block

CHUNK = 256


def decode(pk, beg, ne):
    """numpy restatement of the decoder: arcs from (headers, starts, payload)"""
    beg = beg.astype(np.int64)
    nv = len(beg) - 1
    se = int(beg[0])
    nch = (ne + CHUNK - 1) // CHUNK
    hdr = pk[:nch * 4].reshape(nch, 4).astype(np.int64) if nch else np.zeros((0, 4), np.int64)
    starts = pk[nch * 4:nch * 4 + nv]
    payload = pk[nch * 4 + nv:]
    out = np.zeros(ne, dtype=np.uint32)
    first = beg[:-1] - se
    nonempty = (beg[1:] - beg[:-1]) > 0
    start_of = {}
    for v in np.nonzero(nonempty)[0]:
        if first[v] < ne:
            start_of[int(first[v])] = int(v)
    for c in range(nch):
        word, w, carry, vfirst = (int(x) for x in hdr[c])
        a0, a1 = c * CHUNK, min(c * CHUNK + CHUNK, ne)
        bits = np.unpackbits(payload[word:word + 8 * w].view(np.uint8), bitorder="little") if w else None
        prev = carry
        for a in range(a0, a1):
            if a in start_of:
                x = int(starts[start_of[a]])
            else:
                g = 0
                if w:
                    j = a - a0
                    g = int(np.dot(bits[j * w:(j + 1) * w].astype(np.uint64), (1 << np.arange(w, dtype=np.uint64))))
                x = (prev + g) & 0xFFFFFFFF
            out[a] = x
            prev = x
        # the header's vertex holds the chunk's first arc
        assert first[vfirst] <= a0 < beg[vfirst + 1] - se
    return out, nch


def check_graph(beg, csr, block_bytes, nblocks_min=2):
    bl = Blocks(beg, csr, block_bytes)
    assert bl.nblocks >= nblocks_min
    total_raw = total_packed = 0
    for b in range(bl.nblocks):
        d = bl.desc[b]
        bb = np.ascontiguousarray(beg[d.start_vert:d.start_vert + d.nverts + 1])
        cc = np.ascontiguousarray(csr[d.start_edge:d.start_edge + d.nedges])
        pk = pack_block(bb, cc)
        if d.nedges == 0:
            assert len(pk) == 0
            continue
        got, nch = decode(pk, bb, d.nedges)
        np.testing.assert_array_equal(got, cc, err_msg="block %d" % b)
        widths = pk[:nch * 4].reshape(nch, 4)[:, 1].astype(np.int64)
        assert len(pk) == nch * 4 + d.nverts + 8 * int(widths.sum())
        total_raw += 4 * d.nedges
        total_packed += 4 * len(pk)
    return total_packed / total_raw


def test_pack_round_trip_random_graphs():
    # permuted RMAT: sorted lists, chunks across lists, many empty lists
    r = check_graph(*G.rmat(13, 60000, seed=3, permute=True), 16384)
    assert r < 0.85, r
    # directed power law with multi-edges (duplicate targets: zero differences)
    beg, csr = G.power_law_directed(3000, 40000, seed=2, multi=True)
    check_graph(beg, csr, 8192)


def test_pack_giant_vertex_and_vertex_limited_blocks():
    # vertex 0 holds 5000 arcs: a 4096-byte block is cut at 1024 of them (engine/cache.hpp:170-172);
    # low-degree tail: blocks limited by blocksize / 8 vertices
    n = 3000
    src = np.concatenate([np.zeros(5000, np.int64), np.arange(1, n, dtype=np.int64)])
    dst = np.concatenate([np.arange(1, 5001, dtype=np.int64) % n, (np.arange(1, n) * 7) % n])
    beg, csr = G.from_edges(n, src, dst, undirected=False, dedup=False, drop_self=True)
    bl = Blocks(beg, csr, 4096)
    # the split cuts before and after the giant vertex 0: an empty block 0, then vertex 0 alone
    assert bl.desc[0].nverts == 0 and bl.desc[1].nverts == 1 and bl.desc[1].nedges == 1024
    assert int(beg[1] - beg[0]) > 1024
    check_graph(beg, csr, 4096)


@pytest.mark.parametrize("kind", ["unsorted", "dense", "duplicates"])
def test_pack_extreme_widths(kind):
    rng = np.random.default_rng(5)
    n, deg = 40, 700
    beg = np.arange(0, (n + 1) * deg, deg, dtype=np.uint64)
    if kind == "unsorted":       # differences wrap modulo 2^32: width 32
        csr = rng.integers(0, 2**32, n * deg, dtype=np.uint64).astype(np.uint32)
    elif kind == "dense":        # consecutive ids: width 1
        csr = np.tile(np.arange(deg, dtype=np.uint32), n)
    else:                        # one repeated id per list: width 0
        csr = np.repeat(np.arange(n, dtype=np.uint32) * 3, deg)
    pk = pack_block(beg, csr)
    got, nch = decode(pk, beg, len(csr))
    np.testing.assert_array_equal(got, csr)
    widths = set(pk[:nch * 4].reshape(nch, 4)[:, 1].tolist())
    assert widths == {"unsorted": {32}, "dense": {1}, "duplicates": {0}}[kind]
