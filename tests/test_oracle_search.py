"""The oracle's membership search orders: the bisection of the reference's CPU search
(engine/engine.hpp:224-243), the guarded interpolation order, and the two window orders
(16-entry reads aligned in the edge array, bisected or interpolated guesses).  All must
give the answer of a brute-force scan on every sorted list (with duplicates, skewed or
evenly spread ids, absent values below, inside and above the range, any alignment);
interpolation must stay within about twice the bisection's probes on any list and need
fewer on evenly spread ids, the case it is meant for; a bisected window stays within
bisection's worst case (each probe at least halves the range), and hand-worked cases pin
where the windows fall."""
import math

import numpy as np

import _oracle as O


def lists(rng):
    yield np.array([], dtype=np.uint32)
    yield np.array([5], dtype=np.uint32)
    yield np.array([3, 3, 3, 7, 7, 9], dtype=np.uint32)
    for n in (2, 15, 16, 17, 100, 1000, 5000):
        yield np.sort(rng.choice(1 << 20, n, replace=False)).astype(np.uint32)             # even spread
        yield np.sort((rng.pareto(0.7, n) * 50).astype(np.uint64) % (1 << 20)).astype(np.uint32)   # skewed, dups
        yield np.sort(rng.integers(0, 4 * n + 1, n)).astype(np.uint32)                     # dense, dups


def test_search_orders_agree_with_brute_force():
    rng = np.random.default_rng(3)
    end = 1 << 20
    for a in lists(rng):
        vals = set(a.tolist())
        qs = list(vals)[:64] + rng.integers(0, end, 64).tolist() + [0, end - 1]
        for x in qs:
            want = x in vals
            f0, p0 = O.member(a, x, 0, end)
            f1, p1 = O.member(a, x, 1, end)
            assert f0 == want and f1 == want, (len(a), x)
            n = len(a)
            assert p0 <= max(1, math.ceil(math.log2(n + 1)))
            assert p1 <= 2 * math.ceil(math.log2(n + 1)) + 2
            for head in (0, 5, 15, 1 << 33):
                f2, p2 = O.member(a, x, 2, end, head)
                f3, p3 = O.member(a, x, 3, end, head)
                assert f2 == want and f3 == want, (len(a), x, head)
                assert p2 <= max(1, math.ceil(math.log2(n + 1))), (len(a), x, head)
                assert p3 <= 2 * math.ceil(math.log2(n + 1)) + 2


def test_interpolation_needs_fewer_probes_on_even_ids():
    rng = np.random.default_rng(5)
    end = 1 << 24
    t0 = t1 = 0
    for _ in range(200):
        a = np.sort(rng.choice(end, 4096, replace=False)).astype(np.uint32)
        for x in rng.choice(a, 8):
            t0 += O.member(a, x, 0, end)[1]
            t1 += O.member(a, x, 1, end)[1]
    assert t1 < 0.6 * t0, (t0, t1)


def test_walks_do_not_depend_on_the_search_order():
    import graphs as G
    beg, csr = G.rmat(11, 15000, seed=4, permute=True)
    n = len(beg) - 1
    outs = []
    for search in (0, 1):
        cfg = O.walk_cfg(9, O.type_cfg(O.N2V, 15, p=0.5, q=2.0, search=search))
        outs.append(O.run(beg, csr, cfg, n, visits=True))
    np.testing.assert_array_equal(outs[0]["hashes"], outs[1]["hashes"])
    np.testing.assert_array_equal(outs[0]["visits"], outs[1]["visits"])
    t0, t1 = outs[0]["totals"], outs[1]["totals"]
    assert t0["steps"] == t1["steps"] and t0["trials"] == t1["trials"]
    assert t1["probes"] < t0["probes"]


def test_window_orders_known_answers():
    # a = 0, 10, ..., 390 starting at entry 5 of the edge array: the aligned 16-entry
    # pieces hold list indices [-5, 11), [11, 27) and [27, 43)
    a = np.arange(0, 400, 10, dtype=np.uint32)
    # bisected guesses: first guess 19 -> piece [11, 27)
    assert O.member(a, 250, 2, 400, 5) == (True, 1)
    assert O.member(a, 120, 2, 400, 5) == (True, 1)    # index 12: the piece is placed by head
    assert O.member(a, 120, 2, 400, 0) == (True, 2)    # head 0: pieces [0, 16), [16, 32): 160 > 120, then [0, 16)
    assert O.member(a, 255, 2, 400, 5) == (False, 1)   # between two entries read
    assert O.member(a, 5, 2, 400, 5) == (False, 2)     # below: [0, 11), guess 5 -> piece [-5, 11)
    assert O.member(a, 395, 2, 400, 5) == (False, 2)   # above: [27, 40), guess 33 -> piece [27, 43)
    # interpolated guesses: 120 * 40 / 400 = 12 -> piece [11, 27)
    assert O.member(a, 120, 3, 400, 5) == (True, 1)
    assert O.member(a, 400, 3, 400, 5) == (False, 0)   # not below id_end: no probe


def test_window_orders_need_fewer_probes():
    rng = np.random.default_rng(7)
    end = 1 << 24
    t = [0, 0, 0, 0]
    for _ in range(200):
        a = np.sort(rng.choice(end, 2048, replace=False)).astype(np.uint32)
        head = int(rng.integers(0, 1 << 30))
        for x in list(rng.choice(a, 4)) + list(rng.integers(0, end, 4)):
            for s in range(4):
                t[s] += O.member(a, x, s, end, head)[1]
    assert t[2] < 0.7 * t[0] and t[3] < 0.6 * t[1], t
