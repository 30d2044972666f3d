"""Node2vec's membership band: the oracle searches prev's list only when the accept word
lies between the thresholds of the in-N(prev) and the other class.  Its walks must equal,
path for path, those of a plain-Python walker that classifies every candidate by a set
lookup and accepts iff the accept word is below its class threshold (the rule of
gpu/gpuwalk.hpp:118-150 without any shortcut), for every ordering of 1/p, 1 and 1/q."""
import numpy as np
import pytest

import _oracle as O
import graphs as G


def py_walks(beg, csr, tc, seed, n, length):
    key = [seed & 0xFFFFFFFF, seed >> 32]
    out = np.full((n, length + 1), 0xFFFFFFFF, dtype=np.uint32)
    for w in range(n):
        cur = prev = w
        out[w, 0] = w
        for hop in range(length):
            dv = int(beg[cur + 1] - beg[cur])
            if dv == 0:
                break
            nprev = set(csr[int(beg[prev]):int(beg[prev + 1])].tolist())
            k = 0
            while True:
                if k % 2 == 0:
                    d = O.philox([w, hop, k >> 1, 0], key)
                rp, ra = (int(d[0]), int(d[1])) if k % 2 == 0 else (int(d[2]), int(d[3]))
                x = int(csr[int(beg[cur]) + ((rp * dv) >> 32)])
                t = tc.t_p if x == prev else (tc.t_one if x in nprev else tc.t_q)
                if ra < t:
                    break
                k += 1
            prev, cur = cur, x
            out[w, hop + 1] = x
    return out


@pytest.mark.parametrize("p,q", [(0.5, 2.0), (2.0, 0.5), (1.0, 0.25), (1.0, 1.0), (0.25, 4.0), (4.0, 1.0),
                                 (0.5, 0.5), (3.0, 3.0)])
@pytest.mark.parametrize("search", [0, 1])
def test_band_keeps_walks(p, q, search):
    beg, csr = G.rmat(8, 1500, seed=3, permute=True)
    n = len(beg) - 1
    tc = O.type_cfg(O.N2V, 10, p=p, q=q, search=search)
    seed = 0x1234_5678_9ABC
    o = O.run(beg, csr, O.walk_cfg(seed, tc), n, paths=True)
    np.testing.assert_array_equal(o["paths"], py_walks(beg, csr, tc, seed, n, 10))


def test_band_saves_searches_where_the_classes_agree():
    """p = 0.5, q = 2: T(1) = 2^31 and T(1/q) = 2^30, so only accept words in
    [2^30, 2^31) need the search -- a third of the trials that pass the fast path.
    p = 1, q = 0.25: T(1) = T(1/p) = 2^30 = min T, so every such trial searches."""
    beg, csr = G.rmat(9, 4000, seed=5, permute=True)
    n = len(beg) - 1
    seed = 77
    tc = O.type_cfg(O.N2V, 12, p=0.5, q=2.0)
    assert (tc.t_one, tc.t_q, tc.t_min) == (1 << 31, 1 << 30, 1 << 30)
    tc2 = O.type_cfg(O.N2V, 12, p=1.0, q=0.25)
    assert (tc2.t_one, tc2.t_q, tc2.t_min) == (1 << 30, 1 << 32, 1 << 30)
    # count the searches each rule makes, from the draws of the oracle's own walks
    for cfg, frac in ((tc, 1 / 3), (tc2, 1.0)):
        o = O.run(beg, csr, O.walk_cfg(seed, cfg), n, paths=True)
        key = [seed & 0xFFFFFFFF, seed >> 32]
        searched = past_fast = 0
        for w in range(0, n, 7):
            path = o["paths"][w]
            for hop in range(12):
                if path[hop + 1] == 0xFFFFFFFF:
                    break
                cur = int(path[hop])
                prev = int(path[hop - 1]) if hop else cur
                dv = int(beg[cur + 1] - beg[cur])
                nprev = set(csr[int(beg[prev]):int(beg[prev + 1])].tolist())
                k = 0
                while True:
                    if k % 2 == 0:
                        d = O.philox([w, hop, k >> 1, 0], key)
                    rp, ra = (int(d[0]), int(d[1])) if k % 2 == 0 else (int(d[2]), int(d[3]))
                    x = int(csr[int(beg[cur]) + ((rp * dv) >> 32)])
                    if ra >= cfg.t_min and x != prev:
                        past_fast += 1
                        searched += min(cfg.t_one, cfg.t_q) <= ra < max(cfg.t_one, cfg.t_q)
                    t = cfg.t_p if x == prev else (cfg.t_one if x in nprev else cfg.t_q)
                    if ra < t:
                        break
                    k += 1
        assert past_fast > 500
        assert abs(searched / past_fast - frac) < 0.05
