"""GPU graph generators: RMAT is bit-exact against a numpy restatement of its draws
(Philox per edge, quadrant thresholds, Feistel permutation, sort, dedup); Chung-Lu is
checked through its invariants (sorted lists, exact degree sequence, max degree)."""
import math

import numpy as np
import pytest

import np_ref as R
from gowalker_amd import engine as E

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scale,m,und,perm,dedup", [(10, 5000, True, False, True), (11, 9000, True, True, True),
                                                    (12, 20000, False, False, False), (9, 3000, False, True, True)])
def test_rmat_matches_numpy(gpu, scale, m, und, perm, dedup):
    with E.Context() as ctx:
        beg, csr = E.gen_rmat(ctx, scale, m, 0.57, 0.19, 0.19, seed=42, undirected=und, permute=perm, dedup=dedup)
    rb, rc = R.rmat_csr(scale, m, 0.57, 0.19, 0.19, 42, und, perm, dedup)
    np.testing.assert_array_equal(beg, rb)
    np.testing.assert_array_equal(csr, rc)


def test_rmat_multi_part_sort(gpu):
    # enough keys that the generator splits the sort into several source-range parts
    with E.Context() as ctx:
        beg, csr = E.gen_rmat(ctx, 22, 300_000_000, seed=7, undirected=True, permute=True, dedup=True)
    n = len(beg) - 1
    assert n == 1 << 22
    d = np.diff(beg.astype(np.int64))
    assert (d >= 0).all() and beg[-1] == len(csr)
    # sorted and duplicate-free within every list, no self loops, symmetric degree sum
    src = np.repeat(np.arange(n, dtype=np.int64), d)
    same = src[1:] == src[:-1]
    assert (csr[1:][same] > csr[:-1][same]).all()
    assert not (csr == src).any()
    assert len(csr) % 2 == 0


def test_chung_lu_invariants(gpu):
    n, m, gamma, dmax = 200_000, 3_000_000, 2.2, 40_000
    with E.Context() as ctx:
        beg, csr = E.gen_chung_lu(ctx, n, m, gamma, dmax, seed=42, dedup=False)
    d = np.diff(beg.astype(np.int64))
    assert beg[-1] == m == len(csr)
    assert abs(d[0] - dmax) <= 1 and d.argmax() == 0
    assert (np.diff(d) <= 1).all()        # degrees follow the decreasing weights (floor rounding)
    src = np.repeat(np.arange(n, dtype=np.int64), d)
    same = src[1:] == src[:-1]
    assert (csr[1:][same] >= csr[:-1][same]).all()
    assert csr.max() < n
    # targets follow the same weights as the out-degrees: the top 1% of ids receive the
    # share of arcs they send
    top = int(n * 0.01)
    indeg = np.bincount(csr, minlength=n)
    assert math.isclose(indeg[:top].sum() / m, d[:top].sum() / m, rel_tol=0.05)


def test_chung_lu_dedup(gpu):
    n, m = 200_000, 3_000_000
    with E.Context() as ctx:
        beg, csr = E.gen_chung_lu(ctx, n, m, 2.2, 40_000, seed=42, dedup=True)
    d = np.diff(beg.astype(np.int64))
    assert abs(int(beg[-1]) - m) < 0.03 * m
    src = np.repeat(np.arange(n, dtype=np.int64), d)
    same = src[1:] == src[:-1]
    assert (csr[1:][same] > csr[:-1][same]).all()     # strictly increasing: no duplicates
    assert not (csr.astype(np.int64) == src).any()    # no self-loops
