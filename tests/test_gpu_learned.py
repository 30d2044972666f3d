"""What the trained network gains over the method without it: the out-of-core engine with
the learned scheduler against the reference's plain walker-count scheduler on the same
graph, walks and tight HBM cap.  Walks must be bit-identical (counter-based RNG); the
learned run must move clearly fewer PCIe bytes (block uploads + zero-copy reads) and must
actually have used the model."""
import numpy as np
import pytest

import graphs as G
from gowalker_amd import engine as E

pytestmark = pytest.mark.gpu


def pinned(beg, csr):
    b = E.PinnedArray(len(beg), np.uint64)
    c = E.PinnedArray(len(csr), np.uint32)
    b.a[:] = beg
    c.a[:] = csr
    return b, c


def test_learned_scheduler_moves_fewer_bytes(gpu):
    beg, csr = G.rmat(16, 900_000, seed=17, permute=True)
    n = len(beg) - 1
    block_bytes = (len(csr) * 4) // 40 + 64
    b, c = pinned(beg, csr)
    types = [dict(algo=E.NODE2VEC, length=40, p=0.5, q=2.0)]
    with E.Context() as ctx:
        bl = E.Blocks(beg, csr, block_bytes)
        ctx.set_blocks(bl)
        slot = ctx.slot_bytes()
        nb = bl.nblocks
    walker_bytes = n * (5 * 4 + 8)
    mem = walker_bytes + int(0.6 * nb) * slot + 4096
    out = {}
    for name, sched in (("plain", E.SCHED_PLAIN), ("learned", E.SCHED_LEARNED)):
        rep, h, _ = E.run_out_of_core(b.a, c.a, types, 23, n, memory_bytes=mem, block_bytes=block_bytes,
                                      scheduler=sched, hashes=True)
        out[name] = (rep, h)
        print(name, {k: rep[k] for k in ("steps", "h2d_bytes", "zc_touches", "uploads", "staged", "iterations",
                                         "learned_decisions", "tail_decisions", "fallbacks", "nslots", "nblocks")})
    (rp, hp), (rl, hl) = out["plain"], out["learned"]
    np.testing.assert_array_equal(hp, hl)
    assert rp["steps"] == rl["steps"]
    bytes_p = rp["h2d_bytes"] + 64 * rp["zc_touches"]
    bytes_l = rl["h2d_bytes"] + 64 * rl["zc_touches"]
    print("learned / plain PCIe bytes %.3f" % (bytes_l / bytes_p))
    assert rl["nslots"] < rl["nblocks"]
    assert rl["learned_decisions"] + rl["tail_decisions"] > 0
    assert bytes_l < 0.9 * bytes_p
