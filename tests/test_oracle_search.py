"""The oracle's two membership search orders: the bisection of the reference's CPU search
(engine/engine.hpp:224-243) and the guarded interpolation order.  Both must give the
answer of a brute-force scan on every sorted list (with duplicates, skewed or evenly
spread ids, absent values below, inside and above the range); interpolation must stay
within about twice the bisection's probes on any list and need fewer on evenly spread
ids, the case it is meant for."""
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
