"""The transfer encoding of a block (gw_pack_block, host code, no GPU) decoded by a numpy
restatement of its documented layout (include/gw.h): a header, beg chunks of bit-packed
offset differences, arc chunks of Elias-Fano-coded gap sums.  The decoded beg and csr must
equal the block's exactly, for chunks that start inside lists, empty lists, a truncated
giant vertex, duplicate targets (zero gaps), unsorted lists (gaps that wrap modulo 2^32,
64-bit gap sums) and vertex-limited blocks; the packed size is what the layout predicts.
The GPU decoders are checked against the same blocks through walks and slot contents
(tests/test_gpu_pack.py)."""
import numpy as np
import pytest

import graphs as G
from gowalker_amd.engine import Blocks, pack_block

CHUNK = 256
MAGIC = 0x32505747


def bits_of(words):
    """little-endian bit array of u32 words"""
    return np.unpackbits(np.ascontiguousarray(words, dtype=np.uint32).view(np.uint8), bitorder="little")


def fields(words, n, w):
    """n w-bit little-endian fields packed from bit 0 of words"""
    if w == 0:
        return np.zeros(n, dtype=np.uint64)
    b = bits_of(words)[:n * w].reshape(n, w).astype(np.uint64)
    return (b << np.arange(w, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)


def decode(pk, nv, ne):
    """numpy restatement of the decoders: (beg, csr, per-chunk L) from a packed block"""
    assert pk[0] == MAGIC
    nvc = (nv + 1 + CHUNK - 1) // CHUNK
    assert pk[1] == nvc
    vpw = int(pk[2])
    vh = pk[4:4 + 4 * nvc].reshape(nvc, 4).astype(np.int64)
    vp = pk[4 + 4 * nvc:4 + 4 * nvc + vpw]
    beg = np.zeros(nv + 1, dtype=np.uint64)
    for k in range(nvc):
        base = np.uint64(vh[k, 0]) | (np.uint64(vh[k, 1]) << np.uint64(32))
        off = int(vh[k, 2])
        d = []
        for g in range(4):
            w = (int(vh[k, 3]) >> (8 * g)) & 0xFF
            d.append(fields(vp[off:off + 2 * w], 64, w))
            off += 2 * w
        d = np.concatenate(d)
        assert d[0] == 0
        cnt = min(CHUNK, nv + 1 - k * CHUNK)
        beg[k * CHUNK:k * CHUNK + cnt] = base + np.cumsum(d, dtype=np.uint64)[:cnt]
    nch = (ne + CHUNK - 1) // CHUNK
    ah_off = 4 + 4 * nvc + vpw
    ah = pk[ah_off:ah_off + 4 * nch].reshape(nch, 4).astype(np.int64)
    ap = pk[ah_off + 4 * nch:]
    se = int(beg[0])
    starts = set(int(x) - se for x, y in zip(beg[:-1], beg[1:]) if y > x)
    out = np.zeros(ne, dtype=np.uint32)
    Ls = []
    words = 0
    for c in range(nch):
        off, lz, carry, vfirst = (int(x) for x in ah[c])
        L, nz = lz & 0xFF, lz >> 8
        Ls.append(L)
        a0 = c * CHUNK
        n = min(CHUNK, ne - a0)
        assert off == words
        hw = (n + nz + 31) // 32
        words += 8 * L + hw
        low = fields(ap[off:off + 8 * L], n, L)
        ones = np.nonzero(bits_of(ap[off + 8 * L:off + 8 * L + hw]))[0]
        assert len(ones) == n
        S = ((ones - np.arange(n)).astype(np.uint64) << np.uint64(L)) | low
        assert int(S[-1]) >> L == nz
        prev_S, base = 0, carry          # a continued list: carry + S_j
        for j in range(n):
            if a0 + j in starts:
                prev_S, base = (int(S[j - 1]) if j else 0), 0
            out[a0 + j] = (base + int(S[j]) - prev_S) & 0xFFFFFFFF
        # the header's vertex holds the chunk's first arc
        assert int(beg[vfirst]) - se <= a0 < int(beg[vfirst + 1]) - se
    assert len(pk) == ah_off + 4 * nch + words
    return beg, out, np.array(Ls)


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
        gb, got, _ = decode(pk, d.nverts, d.nedges)
        np.testing.assert_array_equal(gb, bb.astype(np.uint64), err_msg="block %d beg" % b)
        np.testing.assert_array_equal(got, cc, err_msg="block %d" % b)
        total_raw += 8 * (d.nverts + 1) + 4 * d.nedges
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
def test_pack_extreme_gaps(kind):
    rng = np.random.default_rng(5)
    n, deg = 40, 700
    beg = np.arange(0, (n + 1) * deg, deg, dtype=np.uint64)
    if kind == "unsorted":       # gaps wrap modulo 2^32; chunk gap sums pass 2^32
        csr = rng.integers(0, 2**32, n * deg, dtype=np.uint64).astype(np.uint32)
    elif kind == "dense":        # consecutive ids: gap sums below the chunk's arc count
        csr = np.tile(np.arange(deg, dtype=np.uint32), n)
    else:                        # one repeated id per list: zero gaps
        csr = np.repeat(np.arange(n, dtype=np.uint32) * 3, deg)
    pk = pack_block(beg, csr)
    gb, got, Ls = decode(pk, n, len(csr))
    np.testing.assert_array_equal(gb, beg)
    np.testing.assert_array_equal(got, csr)
    if kind == "unsorted":
        assert Ls.min() >= 29 and Ls.max() <= 31
    else:
        assert Ls.max() == 0


def test_pack_size_follows_mean_gap():
    """An arc costs about log2(its chunk's mean gap) + 2 bits: lists of a fixed gap 2^k
    pack to L = k per chunk, with 8k words of lower bits and at most 24 of upper bits."""
    for k in (0, 3, 11, 20):
        nv, deg = 8, 512
        beg = np.arange(0, (nv + 1) * deg, deg, dtype=np.uint64)
        csr = np.tile((np.arange(deg, dtype=np.uint64) << np.uint64(k)) + np.uint64(1 << k), nv).astype(np.uint32)
        pk = pack_block(beg, csr)
        gb, got, Ls = decode(pk, nv, len(csr))
        np.testing.assert_array_equal(got, csr)
        # lists start at 2^k and step by 2^k: every chunk's gap sums are 256 x 2^k (two
        # lists per chunk at most, each restarting at 2^k)
        assert set(Ls.tolist()) <= {max(k - 1, 0), k}, (k, Ls)
        assert len(pk) * 32 <= len(csr) * (k + 2) + 64 * 48 + 32 * 80, (k, len(pk))
