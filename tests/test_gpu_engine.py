"""The C++ out-of-core engine (libgwhost.so: scheduler, slot cache, H2D uploads, zero-copy
tail) over libgw.so: walks are identical to the oracle's whatever the cache budget and
scheduler; bytes moved follow the block sizes; failures come back as status codes."""
import os
import subprocess
import sys

import numpy as np
import pytest

import _oracle as O
import graphs as G
from gowalker_amd import engine as E
from gowalker_amd._abi import GwError

pytestmark = pytest.mark.gpu


def pinned_graph(beg, csr):
    b = E.PinnedArray(len(beg), np.uint64)
    c = E.PinnedArray(len(csr), np.uint32)
    b.a[:] = beg
    c.a[:] = csr
    return b, c


def oracle_hashes(beg, csr, types, seed, n, wpv):
    t = types[0]
    if t["algo"] == E.NODE2VEC:
        tc = O.type_cfg(O.N2V, t["length"], p=t["p"], q=t["q"], term=t.get("term", 0.0))
    else:
        tc = O.type_cfg(O.SOPR, t["length"], alpha=t["alpha"], term=t.get("term", 0.0))
    return O.run(beg, csr, O.walk_cfg(seed, tc), n, wpv=wpv, visits=True)


@pytest.mark.parametrize("sched", [E.SCHED_PLAIN, E.SCHED_LEARNED, E.SCHED_SOWALKER])
@pytest.mark.parametrize("memory_blocks", [3, 6, 100])
def test_out_of_core_matches_oracle(gpu, sched, memory_blocks):
    beg, csr = G.rmat(13, 60000, seed=3, permute=True)
    n = len(beg) - 1
    block_bytes = 16384
    b, c = pinned_graph(beg, csr)
    types = [dict(algo=E.NODE2VEC, length=20, p=0.5, q=2.0)]
    # budget = walker state (cur, prev, meta and a worklist entry: 4 u32, a u64 hash per walker,
    # a u32 visit count per vertex)
    # + `memory_blocks` slots
    with E.Context() as ctx:
        ctx.set_blocks(E.Blocks(beg, csr, block_bytes))
        slot = ctx.slot_bytes()
    walker_bytes = n * (4 * 4 + 8) + n * 4
    mem = walker_bytes + memory_blocks * slot + 4096
    rep, h, v = E.run_out_of_core(b.a, c.a, types, 9, n, memory_bytes=mem, block_bytes=block_bytes,
                                  scheduler=sched, hashes=True, visits=True)
    o = oracle_hashes(beg, csr, types, 9, n, 1)
    print(rep)
    np.testing.assert_array_equal(h, o["hashes"])
    np.testing.assert_array_equal(v, o["visits"])
    assert rep["steps"] == o["totals"]["steps"]
    assert rep["nslots"] == min(memory_blocks, rep["nblocks"])
    # every block is loaded at least once, by upload or by a prefetch that was committed
    assert rep["h2d_bytes"] > 0 and rep["uploads"] + rep["staged"] >= min(memory_blocks, rep["nblocks"])
    b.close()
    c.close()


def test_zero_copy_tail_and_sopr(gpu):
    beg, csr = G.power_law_directed(20000, 200000, seed=5, multi=False)
    n = len(beg) - 1
    b, c = pinned_graph(beg, csr)
    types = [dict(algo=E.SOPR, length=25, alpha=0.2, term=0.1)]
    # 64 KB blocks hold 16384 edges: no vertex is truncated (see test_giant_vertex_truncation);
    # a budget of 3 slots keeps blocks out, and a high zero_threshold ends the run by zero-copy
    with E.Context() as ctx:
        ctx.set_blocks(E.Blocks(beg, csr, 65536))
        slot = ctx.slot_bytes()
    mem = (n * 2) * (4 * 4 + 8) + n * 4 + 3 * slot + 4096
    rep, h, v = E.run_out_of_core(b.a, c.a, types, 4, n * 2, wpv=2, memory_bytes=mem, block_bytes=65536,
                                  zero_threshold=2000, hashes=True, visits=True)
    o = oracle_hashes(beg, csr, types, 4, n * 2, 2)
    np.testing.assert_array_equal(h, o["hashes"])
    np.testing.assert_array_equal(v, o["visits"])
    assert rep["zero_copy_runs"] >= 1
    b.close()
    c.close()


def test_tight_budget_over_block_sizes(gpu):
    """The engine sizes its slot count with gw_slot_bytes and gw_reserve_slots charges the
    same bytes per slot: with a budget of the walker state plus exactly three slots (64 B
    of slack), the walk must finish for every block size, whatever the largest block's
    edge count modulo 64 (the padding the slot's window alignment adds)."""
    beg, csr = G.rmat(12, 30000, seed=4, permute=True)
    n = len(beg) - 1
    b, c = pinned_graph(beg, csr)
    types = [dict(algo=E.NODE2VEC, length=10, p=0.5, q=2.0)]
    ref = None
    for block_bytes in range(16384, 16384 + 33 * 8, 8):
        with E.Context() as ctx:
            ctx.set_blocks(E.Blocks(beg, csr, block_bytes))
            slot = ctx.slot_bytes()
        mem = n * (4 * 4 + 8) + 3 * slot + 64
        rep, h, _ = E.run_out_of_core(b.a, c.a, types, 6, n, memory_bytes=mem, block_bytes=block_bytes,
                                      scheduler=E.SCHED_PLAIN, hashes=True)
        assert rep["nslots"] == 3, (block_bytes, rep["nslots"])
        if ref is None:
            ref = h
        np.testing.assert_array_equal(h, ref)
    b.close()
    c.close()


def test_budget_too_small_is_an_error(gpu):
    beg, csr = G.rmat(10, 5000, seed=1)
    b, c = pinned_graph(beg, csr)
    with pytest.raises(GwError) as e:
        E.run_out_of_core(b.a, c.a, [dict(algo=E.NODE2VEC, length=5, p=1.0, q=1.0)], 1, 1024,
                          memory_bytes=1024, block_bytes=4096)
    assert "budget" in str(e.value).lower()
    b.close()
    c.close()


def test_injected_upload_failure_returns_status(gpu):
    # the failure is injected from the host (GW_FAULT); the library returns a status and
    # the process carries on -- run in a child so the environment stays clean
    code = (
        "import sys; sys.path[:0]=[%r, %r];"
        "import numpy as np, graphs as G; from gowalker_amd import engine as E; from gowalker_amd._abi import GwError;"
        "beg,csr=G.rmat(10,5000,seed=1); b=E.PinnedArray(len(beg),np.uint64); c=E.PinnedArray(len(csr),np.uint32);"
        "b.a[:]=beg; c.a[:]=csr\n"
        "try:\n E.run_out_of_core(b.a,c.a,[dict(algo=0,length=5,p=1.0,q=1.0)],1,1024,block_bytes=4096)\n"
        "except GwError as e:\n print('STATUS', e.status, e); sys.exit(0)\n"
        "sys.exit(3)\n") % (os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, GW_FAULT="upload:3")
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "STATUS -7" in r.stdout


def test_apps_end_to_end(gpu, tmp_path):
    """gw_generate writes the on-disk format; Node2vec runs on it out of core and its path
    hashes equal the oracle's on the same files."""
    build = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gowalker_amd", "_build")
    base = str(tmp_path / "g")
    subprocess.run([os.path.join(build, "gw_generate"), base, "rmat", "scale", "16", "edges", "400000", "undirected",
                    "dedup", "permute", "seed", "3"], check=True, timeout=120)
    hf = str(tmp_path / "h.bin")
    r = subprocess.run([os.path.join(build, "Node2vec"), base + ".txt", "length", "15", "walkpersource", "2",
                        "memory", "1", "blocksize", "1", "p", "0.5", "q", "2", "gpu", "cpu", "GOWalker", "seed", "5",
                        "hashes", hf, "metrics.reporter.filename", str(tmp_path / "m.txt")],
                       capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    beg = np.fromfile(base + ".beg", dtype=np.uint64)
    csr = np.fromfile(base + ".csr", dtype=np.uint32)
    o = O.run(beg, csr, O.walk_cfg(5, O.type_cfg(O.N2V, 15, p=0.5, q=2.0)), (len(beg) - 1) * 2, wpv=2)
    np.testing.assert_array_equal(np.fromfile(hf, dtype=np.uint64), o["hashes"])
    m = open(str(tmp_path / "m.txt")).read()
    assert "0_Total_time=" in m and "walk_steps_per_s=" in m
    # a text edge list is converted on first use (preprocess/graph_converter.hpp:286-384):
    # each list written in reverse with a comment line and self-loops, it converts back to
    # the generator's .beg/.csr and the walk is unchanged
    deg = np.diff(beg.astype(np.int64))
    src = np.repeat(np.arange(len(deg), dtype=np.int64), deg)
    order = np.lexsort((-np.arange(len(csr)), src))
    lines = ["# text copy"] + [f"{a} {b}" for a, b in zip(src[order].tolist(), csr[order].tolist())]
    lines.insert(1, f"{int(src[0])} {int(src[0])}")
    (tmp_path / "t.txt").write_text("\n".join(lines) + "\n")
    ht = str(tmp_path / "ht.bin")
    r3 = subprocess.run([os.path.join(build, "Node2vec"), str(tmp_path / "t.txt"), "length", "15", "walkpersource", "2",
                         "memory", "1", "blocksize", "1", "p", "0.5", "q", "2", "gpu", "cpu", "GOWalker", "seed", "5",
                         "hashes", ht], capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    print(r3.stdout[-2000:], r3.stderr[-2000:])
    assert r3.returncode == 0
    tb = np.fromfile(str(tmp_path / "t.beg"), dtype=np.uint64)
    np.testing.assert_array_equal(np.fromfile(str(tmp_path / "t.csr"), dtype=np.uint32), csr)
    np.testing.assert_array_equal(tb, beg[:len(tb)])
    assert np.all(beg[len(tb):] == beg[-1])
    if len(tb) == len(beg):
        np.testing.assert_array_equal(np.fromfile(ht, dtype=np.uint64), o["hashes"])
    # the SOWalker scheduler (simulated annealing over resident sets) walks the same paths
    hs = str(tmp_path / "hs.bin")
    r4 = subprocess.run([os.path.join(build, "Node2vec"), base + ".txt", "length", "15", "walkpersource", "2",
                         "memory", "1", "blocksize", "1", "p", "0.5", "q", "2", "SOWalker", "seed", "5",
                         "hashes", hs], capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    print(r4.stdout[-1000:], r4.stderr[-1000:])
    assert r4.returncode == 0
    np.testing.assert_array_equal(np.fromfile(hs, dtype=np.uint64), o["hashes"])
    # packed block uploads and the staging scheduler walk the same paths; the metrics file
    # records the bytes that crossed PCIe below the blocks' size in the reference's format
    hp = str(tmp_path / "hp.bin")
    mp = str(tmp_path / "mp.txt")
    r5 = subprocess.run([os.path.join(build, "Node2vec"), base + ".txt", "length", "15", "walkpersource", "2",
                         "memory", "1", "blocksize", "1", "p", "0.5", "q", "2", "GOWalker", "sched", "staging", "pack",
                         "seed", "5", "hashes", hp, "metrics.reporter.filename", mp], capture_output=True, text=True,
                        timeout=300, cwd=str(tmp_path))
    print(r5.stdout[-1000:], r5.stderr[-1000:])
    assert r5.returncode == 0
    np.testing.assert_array_equal(np.fromfile(hp, dtype=np.uint64), o["hashes"])
    kv = dict(l.split("=", 1) for l in open(mp).read().splitlines() if "=" in l)

    def metric(key):   # keys read <run name>.<key>
        return float([v for k, v in kv.items() if k.endswith("." + key)][0])

    assert metric("pack_time") >= 0 and metric("h2d_bytes") < metric("h2d_raw_bytes"), kv
    assert "staging" in open(mp).read()
    # the reference requires one scheduler flag
    r2 = subprocess.run([os.path.join(build, "Node2vec"), base + ".txt", "length", "5"], capture_output=True,
                        text=True, timeout=60, cwd=str(tmp_path))
    assert "Please select one walker from GOWalker and SOWalker!" in r2.stdout


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


def test_giant_vertex_truncation(gpu):
    beg, csr = G.power_law_directed(20000, 200000, seed=5)
    n = len(beg) - 1
    bl = E.Blocks(beg, csr, 8192)
    giants = [i for i in range(bl.nblocks) if bl.desc[i].nverts == 1 and
              bl.ecut[i + 1] - bl.ecut[i] > 2048]
    assert giants, "the test graph must have a vertex longer than a block"
    eb, ec = effective_graph(beg, csr, bl)
    b, c = pinned_graph(beg, csr)
    types = [dict(algo=E.SOPR, length=25, alpha=0.2, term=0.1)]
    rep, h, v = E.run_out_of_core(b.a, c.a, types, 4, n * 2, wpv=2, memory_bytes=0, block_bytes=8192,
                                  zero_threshold=400, hashes=True, visits=True)
    o = oracle_hashes(eb, ec, types, 4, n * 2, 2)
    np.testing.assert_array_equal(h, o["hashes"])
    np.testing.assert_array_equal(v, o["visits"])
    b.close()
    c.close()
