"""Small test graphs as sorted CSR (beg u64[n+1], csr u32) built with numpy."""
import numpy as np


def from_edges(n, src, dst, undirected=False, dedup=True, drop_self=True):
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    if undirected:
        src, dst = np.concatenate([src, dst]), np.concatenate([dst, src])
    if drop_self:
        keep = src != dst
        src, dst = src[keep], dst[keep]
    key = (src << 32) | dst
    key = np.unique(key) if dedup else np.sort(key)
    s = (key >> 32).astype(np.int64)
    d = (key & 0xFFFFFFFF).astype(np.uint32)
    beg = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(beg, s + 1, 1)
    beg = np.cumsum(beg).astype(np.uint64)
    return beg, d


def triangle():
    return from_edges(3, [0, 1, 2], [1, 2, 0], undirected=True)


def star(k=5):
    return from_edges(k + 1, [0] * k, list(range(1, k + 1)), undirected=True)


def path(n=6):
    return from_edges(n, list(range(n - 1)), list(range(1, n)), undirected=True)


def bridge(k=4):
    """Two k-cliques {0..k-1}, {k..2k-1} joined by the edge (k-1, k)."""
    s, d = [], []
    for base in (0, k):
        for i in range(k):
            for j in range(i + 1, k):
                s.append(base + i)
                d.append(base + j)
    s.append(k - 1)
    d.append(k)
    return from_edges(2 * k, s, d, undirected=True)


def dag_dead_end():
    """0->1, 0->2, 1->2, 1->3, 2->3 ; 3 has no out-edges."""
    return from_edges(4, [0, 0, 1, 1, 2], [1, 2, 2, 3, 3])


def random_undirected(n, m, seed=1):
    rng = np.random.default_rng(seed)
    return from_edges(n, rng.integers(0, n, m), rng.integers(0, n, m), undirected=True)


def rmat(scale, m, a=0.57, b=0.19, c=0.19, seed=1, undirected=True, permute=False):
    rng = np.random.default_rng(seed)
    n = 1 << scale
    src = np.zeros(m, dtype=np.int64)
    dst = np.zeros(m, dtype=np.int64)
    for _ in range(scale):
        r = rng.random(m)
        rb = (r >= a + b).astype(np.int64)
        cb = (((r >= a) & (r < a + b)) | (r >= a + b + c)).astype(np.int64)
        src = src * 2 + rb
        dst = dst * 2 + cb
    if permute:
        p = rng.permutation(n)
        src, dst = p[src], p[dst]
    return from_edges(n, src, dst, undirected=undirected)


def power_law_directed(n, m, seed=1, multi=True):
    """Directed graph with heavy-tailed in/out degrees; every vertex has out-degree >= 1."""
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, n + 1) ** 0.8
    w /= w.sum()
    src = np.concatenate([np.arange(n), rng.choice(n, m - n, p=w)])
    dst = rng.choice(n, m, p=w)
    return from_edges(n, src, dst, undirected=False, dedup=not multi, drop_self=True)
