"""The product's plain scheduler (gowalker_amd/host/schedule.hpp, the reference's
GOwalker_scheduler_t::max_choose_blocks, engine/schedule.hpp:70-138) against the oracle's
orc_max_choose on random bucket matrices: both branches (free slots / full cache), the
zero-copy threshold on either side of the best score, ties (small value ranges), empty
rows, and counts past 2^32.  Decisions (accepted or not, and the resulting resident set in
slot order) must be identical."""
import os
import struct
import subprocess

import numpy as np

import _oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def cases(n, seed=11):
    rng = np.random.default_rng(seed)
    for i in range(n):
        nb = int(rng.integers(1, 40))
        nc = int(rng.integers(1, nb + 2))
        full = rng.random() < 0.5
        nbk = min(nb, nc) if full else int(rng.integers(0, min(nb, nc) + 1))
        buckets = rng.permutation(nb)[:nbk].astype(np.uint32)
        kind = i % 4
        if kind == 0:      # ties: a few distinct small values
            W = rng.integers(0, 3, (nb, nb))
        elif kind == 1:    # sparse
            W = rng.integers(0, 1000, (nb, nb)) * (rng.random((nb, nb)) < 0.1)
        elif kind == 2:    # wide range, some counts past 2^32
            W = rng.integers(0, 1 << 34, (nb, nb)) * (rng.random((nb, nb)) < 0.5)
        else:
            W = rng.integers(0, 100000, (nb, nb))
        W = W.astype(np.uint64)
        # a threshold near the best score some of the time, so both sides are exercised
        r, _ = O.max_choose(W, list(buckets), nc, 0)
        zt = 0
        if rng.random() < 0.4:
            scores = W.sum(axis=0) + W.sum(axis=1)
            zt = int(rng.integers(0, int(scores.max()) + 2)) if scores.max() > 0 else 0
        yield nb, nc, zt, buckets, W


def test_plain_scheduler_matches_oracle_on_random_matrices(tmp_path):
    exe = str(tmp_path / "sched_fuzz")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(REPO, "include"), "-I",
                           os.path.join(REPO, "gowalker_amd", "host"), "-o", exe,
                           os.path.join(REPO, "tests", "host", "sched_fuzz.cpp")])
    all_cases = list(cases(2000))
    blob = bytearray()
    for nb, nc, zt, buckets, W in all_cases:
        blob += struct.pack("<IIQI", nb, nc, zt, len(buckets)) + buckets.astype("<u4").tobytes()
        blob += W.astype("<u8").tobytes()
    inp, outp = tmp_path / "cases.bin", tmp_path / "out.txt"
    inp.write_bytes(bytes(blob))
    subprocess.check_call([exe, str(inp), str(outp)], timeout=60)
    lines = outp.read_text().splitlines()
    assert len(lines) == len(all_cases)
    accepted = {0: 0, 1: 0}
    branches = {"fill": 0, "full": 0}
    for (nb, nc, zt, buckets, W), line in zip(all_cases, lines):
        f = [int(x) for x in line.split()]
        r, out = O.max_choose(W, list(buckets), nc, zt)
        assert f[0] == r and f[2:] == [int(x) for x in out], (nb, nc, zt, list(buckets), line)
        accepted[r] += 1
        branches["full" if len(buckets) >= nc else "fill"] += 1
    # the generator reaches every outcome and both branches
    assert min(accepted.values()) > 100 and min(branches.values()) > 100, (accepted, branches)
