"""Vectorised numpy references used by the tests: Philox-4x32-10 (checked against the
oracle's scalar version), the RMAT edge draw and Feistel permutation of the GPU
generator, and an fp32 MLP forward/backward with bf16-rounded weights."""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)


def philox(c0, c1, c2, c3, k0, k1):
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint32).copy() for x in (c0, c1, c2, c3))
    k0 = np.uint32(k0)
    k1 = np.uint32(k1)
    with np.errstate(over="ignore"):
        for r in range(10):
            if r:
                k0 = np.uint32(k0 + W0)
                k1 = np.uint32(k1 + W1)
            p0 = c0.astype(np.uint64) * M0
            p1 = c2.astype(np.uint64) * M1
            n0 = (p1 >> np.uint64(32)).astype(np.uint32) ^ c1 ^ k0
            n1 = (p1 & MASK32).astype(np.uint32)
            n2 = (p0 >> np.uint64(32)).astype(np.uint32) ^ c3 ^ k1
            n3 = (p0 & MASK32).astype(np.uint32)
            c0, c1, c2, c3 = n0, n1, n2, n3
    return c0, c1, c2, c3


def rmat_edges(scale, nedges, a, b, c, seed):
    th = lambda x: np.uint32(0xFFFFFFFF) if x >= 1.0 else np.uint32(int(np.floor(x * 4294967296.0)))
    ta, tab, tabc = th(a), th(a + b), th(a + b + c)
    e = np.arange(nedges, dtype=np.uint64)
    elo = (e & MASK32).astype(np.uint32)
    ehi = (e >> np.uint64(32)).astype(np.uint32)
    s = np.zeros(nedges, dtype=np.uint64)
    d = np.zeros(nedges, dtype=np.uint64)
    k0, k1 = seed & 0xFFFFFFFF, seed >> 32
    for l in range(0, scale, 4):
        rr = philox(elo, ehi, np.full(nedges, l >> 2, np.uint32), np.full(nedges, 7, np.uint32), k0, k1)
        for j in range(4):
            if l + j >= scale:
                break
            x = rr[j]
            row = (x >= tab).astype(np.uint64)
            col = (((x >= ta) & (x < tab)) | (x >= tabc)).astype(np.uint64)
            s = (s << np.uint64(1)) | row
            d = (d << np.uint64(1)) | col
    return s, d


def _fround(r, k, rnd):
    with np.errstate(over="ignore"):
        h = (r * np.uint32(0x9E3779B1) + np.uint32(k) + np.uint32(rnd) * np.uint32(0x85EBCA77)).astype(np.uint32)
        h ^= h >> np.uint32(15)
        h = (h * np.uint32(0x2C1B3C6D)).astype(np.uint32)
        h ^= h >> np.uint32(12)
        h = (h * np.uint32(0x297A2D39)).astype(np.uint32)
        h ^= h >> np.uint32(15)
    return h


def permute_ids(x, scale, seed):
    bits = scale + (scale & 1)
    h = bits // 2
    mask = np.uint32((1 << h) - 1)
    with np.errstate(over="ignore"):
        fk0 = np.uint32(((seed * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF) >> 32)
        fk1 = np.uint32(((seed * 0xC2B2AE3D27D4EB4F) & 0xFFFFFFFFFFFFFFFF) >> 32)
    x = np.asarray(x, dtype=np.uint32).copy()
    todo = np.ones(len(x), dtype=bool)
    while todo.any():
        y = x[todo]
        L, R = y >> np.uint32(h), y & mask
        for r in range(4):
            t = L ^ (_fround(R, fk1 if r & 1 else fk0, r) & mask)
            L, R = R, t
        y = (L << np.uint32(h)) | R
        x[todo] = y
        todo[todo] = (y >> np.uint32(scale)) != 0
    return x


def rmat_csr(scale, nedges, a, b, c, seed, undirected, permute, dedup):
    s, d = rmat_edges(scale, nedges, a, b, c, seed)
    s = s.astype(np.uint32)
    d = d.astype(np.uint32)
    if permute:
        s, d = permute_ids(s, scale, seed), permute_ids(d, scale, seed)
    keep = s != d
    s, d = s[keep].astype(np.uint64), d[keep].astype(np.uint64)
    if undirected:
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
    key = (s << np.uint64(scale)) | d
    key = np.unique(key) if dedup else np.sort(key)
    n = 1 << scale
    src = (key >> np.uint64(scale)).astype(np.int64)
    beg = np.zeros(n + 1, dtype=np.uint64)
    np.add.at(beg, src + 1, 1)
    return np.cumsum(beg).astype(np.uint64), (key & np.uint64(n - 1)).astype(np.uint32)


def bf16_round(x):
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)
    return (u.astype(np.uint32) << np.uint32(16)).view(np.float32)


def mlp_split(params, nin, h1, h2):
    o = 0
    W1 = params[o:o + h1 * nin].reshape(h1, nin); o += h1 * nin
    b1 = params[o:o + h1]; o += h1
    W2 = params[o:o + h2 * h1].reshape(h2, h1); o += h2 * h1
    b2 = params[o:o + h2]; o += h2
    w3 = params[o:o + h2]; o += h2
    b3 = params[o]
    return W1, b1, W2, b2, w3, b3


def mlp_forward(params, x, nin, h1, h2, dtype=np.float64):
    W1, b1, W2, b2, w3, b3 = (np.asarray(t, dtype=dtype) for t in mlp_split(params, nin, h1, h2))
    x = np.asarray(x, dtype=dtype)
    z1 = x @ W1.T + b1
    a1 = np.maximum(z1, 0)
    z2 = a1 @ W2.T + b2
    a2 = np.maximum(z2, 0)
    return a2 @ w3 + b3, (z1, a1, z2, a2)


def mlp_grad(params, x, t, nin, h1, h2):
    W1, b1, W2, b2, w3, b3 = (np.asarray(v, dtype=np.float64) for v in mlp_split(params, nin, h1, h2))
    y, (z1, a1, z2, a2) = mlp_forward(params, x, nin, h1, h2)
    n = len(t)
    dy = 2.0 * (y - t) / n
    dz2 = np.outer(dy, w3) * (z2 > 0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    g = np.concatenate([(dz1.T @ x).ravel(), dz1.sum(0), (dz2.T @ a1).ravel(), dz2.sum(0), dy @ a2, [dy.sum()]])
    return g, np.mean((y - t) ** 2)
