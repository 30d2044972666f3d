"""Scheduler support in libgw.so through the C ABI: staged prefetch (stage / commit), the
bucket matrix and pending work left by a run, and the 12 per-block features the learned
scheduler reads -- each checked against the oracle or a numpy restatement."""
import numpy as np
import pytest

import _oracle as O
import graphs as G
from gowalker_amd.engine import NODE2VEC, OUT_HASH, OUT_VISITS, Blocks, Context

pytestmark = pytest.mark.gpu

TYPES = [dict(algo=NODE2VEC, length=20, p=0.5, q=2.0)]


def oracle(beg, csr, seed, n):
    t = O.type_cfg(O.N2V, 20, p=0.5, q=2.0)
    return O.run(beg, csr, O.walk_cfg(seed, t), n, visits=True)


def features_ref(W, resident, age, desc, ahead):
    nb = W.shape[0]
    W = W.astype(np.float64)
    tot = W.sum()
    f = np.zeros((nb, 12))
    for b in range(nb):
        row, col, diag = W[b].sum(), W[:, b].sum(), W[b, b]
        pair = [W[b, j] + W[j, b] for j in range(nb) if j != b and resident[j]]
        res = sum(pair)
        mn = min(pair) if pair else 0.0
        d = desc[b]
        f[b] = [np.log1p(row), np.log1p(col), np.log1p(diag), np.log1p(res), 1.0 if resident[b] else 0.0,
                np.log1p(mn), np.log1p(res + diag), age[b] / 64.0, np.log1p(d.nedges) / 24.0, ahead,
                np.log1p(d.nedges / d.nverts if d.nverts else 0.0) / 8.0, (row + col - diag) / tot if tot else 0.0]
    return f.astype(np.float32)


def test_stage_commit_features_pending(gpu):
    beg, csr = G.rmat(12, 40000, seed=21, permute=True)
    n = len(beg) - 1
    bl = Blocks(beg, csr, 16384)
    nb = bl.nblocks
    assert nb >= 6
    with Context() as ctx:
        ctx.set_blocks(bl)
        ctx.set_walk(TYPES, 5)
        ctx.init_walkers(n, 1, outputs=OUT_HASH | OUT_VISITS)
        ctx.reserve_slots(4)
        w0, h0 = ctx.pending_work()
        assert (w0, h0) == (n, n * 20)
        ctx.upload(0, 0)
        ctx.upload(1, 1)
        ctx.stage(2, 2)                       # copy in flight, not resident
        assert ctx.staged(2)[0] == 2
        st = ctx.run(buckets=True)
        assert st["walkers_run"] > 0
        # walkers left: state vs pending work
        s = ctx.state()
        live = s[:, 3] == 0
        assert ctx.pending_work() == (int(live.sum()), int((20 - s[live, 2].astype(np.int64)).sum()))
        # features from the bucket matrix the run left, blocks 0 and 1 resident (2 staged)
        W = ctx.bucket_counts()
        resident = np.zeros(nb, bool)
        resident[[0, 1]] = True
        age = np.arange(nb, dtype=np.float32)
        wl, hops = ctx.pending_work()
        f = ctx.block_features(age)
        ref = features_ref(W, resident, age, bl.desc, hops / (wl * 20.0))
        np.testing.assert_allclose(f, ref, rtol=2e-6, atol=1e-6)
        ctx.wait_uploads()
        assert ctx.staged(2) == (2, True)
        ctx.commit(2)
        assert ctx.staged(2)[0] == -1
        # a window of two slots (2 and 3) moves over the heaviest (prev, cur) pairs for a
        # while, then every block is loaded and the walk finishes
        for _ in range(20):
            W = ctx.bucket_counts()
            a, b = (int(x) for x in np.unravel_index(np.argmax(W), W.shape))
            ctx.upload(3, a)
            if b != a:
                ctx.upload(2, b)
            ctx.wait_uploads()
            assert ctx.run()["walkers_run"] > 0
        ctx.reserve_slots(nb)
        for blk in range(nb):
            ctx.upload(blk, blk)
        ctx.wait_uploads()
        assert ctx.run()["walkers_left"] == 0
        assert ctx.bucket_counts().sum() == 0
        hashes, visits = ctx.hashes(), ctx.visits()
    o = oracle(beg, csr, 5, n)
    np.testing.assert_array_equal(hashes, o["hashes"])
    np.testing.assert_array_equal(visits, o["visits"])


def test_block_scores_equal_features_then_forward(gpu):
    """gw_block_scores (features, standardization and the network's forward pass on the
    device, one host wait) returns the features of gw_block_features and, bit for bit, the
    scores gw_mlp_forward gives for the same features standardized on the host."""
    import ctypes as C

    from gowalker_amd._abi import check, lib

    beg, csr = G.rmat(12, 40000, seed=23, permute=True)
    n = len(beg) - 1
    bl = Blocks(beg, csr, 16384)
    nb = bl.nblocks
    rng = np.random.default_rng(2)
    with Context() as ctx:
        ctx.set_blocks(bl)
        ctx.set_walk(TYPES, 5)
        ctx.init_walkers(n, 1, outputs=OUT_HASH)
        ctx.reserve_slots(3)
        for s in range(3):
            ctx.upload(s, s)
        ctx.run(buckets=True)
        mlp = C.c_void_p()
        check(lib().gw_mlp_create(ctx.h, 12, 64, 64, 9, C.byref(mlp)), ctx.h)
        npar = C.c_uint32()
        check(lib().gw_mlp_num_params(mlp, C.byref(npar)))
        p = rng.normal(0, 0.3, npar.value).astype(np.float32)
        check(lib().gw_mlp_set_params(ctx.h, mlp, p.ctypes.data), ctx.h)
        age = rng.uniform(0, 64, nb).astype(np.float32)
        mu = rng.normal(2, 1, 12).astype(np.float32)
        sig = rng.uniform(0.5, 3, 12).astype(np.float32)
        feats = np.zeros((nb, 12), np.float32)
        scores = np.zeros(nb, np.float32)
        check(lib().gw_block_scores(ctx.h, mlp, age.ctypes.data, mu.ctypes.data, sig.ctypes.data, feats.ctypes.data,
                                    scores.ctypes.data, nb), ctx.h)
        f2 = ctx.block_features(age)
        x = ((f2 - mu) / sig).astype(np.float32)
        y = np.zeros(nb, np.float32)
        check(lib().gw_mlp_forward(ctx.h, mlp, x.ctypes.data, nb, y.ctypes.data), ctx.h)
        lib().gw_mlp_destroy(ctx.h, mlp)
    np.testing.assert_array_equal(feats, f2)
    np.testing.assert_array_equal(scores, y)
    assert np.std(scores) > 0
