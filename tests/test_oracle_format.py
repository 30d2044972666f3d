"""Known answers for split_blocks (preprocess/split.hpp:15-87) and the plain walker-count
scheduler (engine/schedule.hpp:70-138), computed by hand from the reference's rules,
for the oracle and for the product's implementations."""
import numpy as np
import pytest

import _oracle as O


SPLIT_CASES = [
    # (beg, block_size, vblocks, eblocks)
    # edge-limited: 4 vertices x 3 edges, max 8 edges -> cut at v-1 = 2 when beg[3]=9 > 8
    ([0, 3, 6, 9, 12], 32, [0, 2, 4], [0, 6, 12]),
    # vertex-limited: max 4 vertices -> cuts at 4 and 8, last cut = entries-1 = 10
    (list(range(11)), 32, [0, 4, 8, 10], [0, 4, 8, 10]),
    # giant vertex 0: duplicate cut at 0 (empty block), then the giant vertex alone
    ([0, 20, 21, 22], 32, [0, 0, 1, 3], [0, 0, 20, 22]),
    # leading zero-degree ids (min_vert > 0 padding of graph_converter.hpp:195-202); the
    # loop also visits the sentinel index entries-1 = 4, where 4 - 0 >= max_nverts cuts,
    # so the reference ends with an empty block
    ([0, 0, 0, 2, 4], 32, [0, 4, 4], [0, 4, 4]),
    # a single block
    ([0, 1, 3, 6], 1 << 20, [0, 3], [0, 6]),
]


@pytest.mark.parametrize("beg,bs,vb,eb", SPLIT_CASES)
def test_split_blocks_known_answers(beg, bs, vb, eb):
    v, e = O.split_blocks(np.array(beg, dtype=np.uint64), bs)
    assert list(v) == vb
    assert list(e) == eb


@pytest.mark.parametrize("beg,bs,vb,eb", SPLIT_CASES)
def test_product_split_and_truncation(beg, bs, vb, eb):
    from gowalker_amd.engine import Blocks
    b = Blocks(np.array(beg, dtype=np.uint64), np.zeros(beg[-1], dtype=np.uint32), bs)
    assert list(b.vcut) == vb and list(b.ecut) == eb
    for i in range(b.nblocks):
        d = b.desc[i]
        assert d.start_vert == vb[i] and d.nverts == vb[i + 1] - vb[i]
        ne = eb[i + 1] - eb[i]
        if d.nverts == 1 and ne > bs // 4:   # engine/cache.hpp:170-172
            ne = bs // 4
        assert d.nedges == ne


def test_giant_block_truncated():
    from gowalker_amd.engine import Blocks
    b = Blocks(np.array([0, 20, 21, 22], dtype=np.uint64), np.zeros(22, dtype=np.uint32), 32)
    assert [b.desc[i].nedges for i in range(3)] == [0, 8, 2]


def W(n, entries):
    m = np.zeros((n, n), dtype=np.uint64)
    for (a, b), v in entries.items():
        m[a, b] = v
    return m


def test_scheduler_not_full_picks_max_then_lowest_id():
    # score(c) = sum_{j in C} W[c,j] + W[j,c] + W[c,c]; c=1: 5+2+0 = 7, c=2: 1+0+6 = 7 -> tie -> 1
    w = W(3, {(1, 0): 5, (0, 1): 2, (2, 2): 6, (2, 0): 1})
    assert O.max_choose(w, [0], 2) == (1, [0, 1])


def test_scheduler_threshold():
    w = W(3, {(1, 0): 5, (0, 1): 2, (2, 2): 6, (2, 0): 1})
    assert O.max_choose(w, [0], 2, zero_threshold=7) == (0, [0])
    assert O.max_choose(w, [0], 2, zero_threshold=6) == (1, [0, 1])


def test_scheduler_full_replaces_weakest_partner():
    # C = [0, 1] full; c=2: pair sums s0 = 3+1 = 4, s1 = 10 -> score 14-4 = 10, replace slot 0
    w = W(4, {(2, 0): 3, (0, 2): 1, (2, 1): 10})
    assert O.max_choose(w, [0, 1], 2) == (1, [2, 1])
    # diagonal of the candidate does not count when full
    w2 = W(4, {(3, 3): 50, (2, 1): 1})
    assert O.max_choose(w2, [0, 1], 2) == (1, [2, 1])


def test_scheduler_nothing_to_do():
    assert O.max_choose(np.zeros((3, 3)), [0], 2) == (0, [0])
    assert O.max_choose(np.zeros((2, 2)), [0, 1], 2) == (0, [0, 1])


def test_scheduler_random_invariants():
    rng = np.random.default_rng(0)
    for _ in range(200):
        n = int(rng.integers(2, 9))
        ncb = int(rng.integers(1, n + 1))
        w = rng.integers(0, 5, (n, n)).astype(np.uint64) * (rng.random((n, n)) < 0.4)
        k = int(rng.integers(0, ncb + 1))
        b = list(rng.permutation(n)[:k])
        r, nb = O.max_choose(w, b, ncb)
        assert len(set(nb)) == len(nb) and len(nb) <= max(ncb, len(b))
        if r:
            assert set(b) != set(nb)
            if len(b) < ncb:
                assert nb[:-1] == b
            else:
                assert sum(x != y for x, y in zip(b, nb)) == 1
