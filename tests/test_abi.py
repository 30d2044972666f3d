"""libgw.so loads and exports every function include/gw.h declares (no GPU calls)."""
import os
import re
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header="gw.h", prefix="gw_"):
    src = open(os.path.join(REPO, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(%s[a-z0-9_]+)\s*\(" % prefix, src)))


def test_library_exports_every_declaration():
    from gowalker_amd import LIB_PATH, lib
    lib()
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB_PATH], text=True)
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if n not in exported]
    assert not missing, missing


def test_python_signatures_cover_header():
    from gowalker_amd._abi import SIGNATURES
    assert set(declared()) == set(SIGNATURES)


def test_host_library_exports_every_declaration():
    """libgwhost.so (include/gwh.h, the C++ engine behind one C call) exports every
    function its header declares, and the Python mirror names each of them."""
    from gowalker_amd._abi import HOST_LIB_PATH, HOST_SIGNATURES, host_lib
    host_lib()
    out = subprocess.check_output(["nm", "-D", "--defined-only", HOST_LIB_PATH], text=True)
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    names = declared("gwh.h", "gwh_")
    assert len(names) >= 5
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert set(names) == set(HOST_SIGNATURES)


def test_host_model_handle_without_gpu():
    """gwh_model holds the learned scheduler's state between walks; creating and querying
    one touches no GPU."""
    from gowalker_amd.engine import Model
    m = Model()
    info = m.info()
    assert info == dict(samples=0, walks=0, adam_steps=0, loss=0.0, has_params=0, scored=0, net_abs_err=0.0,
                        mean_abs_err=0.0)
    m.close()


def test_errors_are_returned_not_raised_by_library():
    import ctypes as C
    from gowalker_amd._abi import GW_OK, lib
    L = lib()
    # null arguments come back as GW_ERR_ARG with a message, no process exit
    assert L.gw_create(None, None) == -1
    assert b"null" in L.gw_last_error(None)
    assert L.gw_set_block_table(None, None, 0) == -1
    v = C.c_uint64(0)
    assert L.gw_split_blocks(None, 0, 0, None, None, 0, C.byref(v)) == -1
    assert GW_OK == 0


def test_split_blocks_product_matches_oracle():
    import numpy as np
    import _oracle as O
    import graphs as G
    from gowalker_amd.engine import Blocks
    for beg, csr, bs in [(*G.rmat(12, 30000, seed=1), 4096), (*G.random_undirected(3000, 9000), 1024),
                         (*G.star(40), 64)]:
        b = Blocks(beg, csr, bs)
        vb, eb = O.split_blocks(beg, bs)
        np.testing.assert_array_equal(b.vcut, vb)
        np.testing.assert_array_equal(b.ecut, eb)


def test_python_flag_constants_match_header():
    from gowalker_amd import _abi
    src = open(os.path.join(REPO, "include", "gw.h")).read()
    flags = {m.group(1): 1 << int(m.group(2)) for m in re.finditer(r"\b(GW_RUN_[A-Z]+)\s*=\s*1u\s*<<\s*(\d+)", src)}
    assert {"GW_RUN_BUCKETS", "GW_RUN_TRACE", "GW_RUN_INTERP", "GW_RUN_WINDOW"} <= set(flags)
    for name, value in flags.items():
        assert getattr(_abi, name) == value, name


def test_noalloc_walk_rejects_sizes_beyond_its_array():
    """gw_walk_resident_noalloc checks its sizes on the host before it touches a GPU: more
    walkers than its fixed device array holds come back as GW_ERR_ARG with the limit in
    the message; the header and the Python mirror state the same limit."""
    import ctypes as C
    from gowalker_amd._abi import GW_NOALLOC_MAX_WALKERS, GW_NOALLOC_SCRATCH_BYTES, lib
    from gowalker_amd.engine import walk_params
    hdr = open(os.path.join(REPO, "include", "gw.h")).read()
    assert "#define GW_NOALLOC_MAX_WALKERS (1u << 20)" in hdr
    assert "#define GW_NOALLOC_SCRATCH_BYTES (1024ull + 12ull * GW_NOALLOC_MAX_WALKERS)" in hdr
    assert GW_NOALLOC_SCRATCH_BYTES == 1024 + 12 * (1 << 20) and GW_NOALLOC_MAX_WALKERS == 1 << 20
    L = lib()
    wp = walk_params([dict(algo=0, length=10, p=1.0, q=1.0)])
    fake = C.c_void_p(0x1000)   # never dereferenced: the checks come first
    st = L.gw_walk_resident_noalloc(fake, fake, 1 << 22, 100, C.byref(wp), GW_NOALLOC_MAX_WALKERS + 1, 1, None,
                                    None, None, None)
    assert st == -1
    msg = L.gw_last_error(None).decode()
    assert str(GW_NOALLOC_MAX_WALKERS) in msg and "walkers" in msg, msg
    assert L.gw_walk_resident_noalloc(fake, fake, 1 << 22, 100, C.byref(wp), 0, 1, None, None, None, None) == -1
    assert L.gw_walk_resident_noalloc(fake, fake, 10, 100, C.byref(wp), 100, 1, None, None, None, None) == -1
    assert b"exceed" in L.gw_last_error(None)
