"""Pins the oracle's step logic against closed forms.

For a walker at (prev u, cur v) the normative samplers give an exact next-vertex law:
a position i of N(v) is proposed with probability c_i / 2^32, where
c_i = #{r in [0, 2^32) : (r * dv) >> 32 == i}, and accepted with probability T_i / 2^32.
So P(i) = c_i T_i / sum_j c_j T_j (rejection), or the two-component mixture law for the
SOPR mixture sampler.  These laws are computed here from the weight definitions of
gpu/gpuwalk.hpp:118-150 (node2vec: 1/p, 1, 1/q) and :241-266 (SOPR: (1-a)/dv +
a/du [x in N(u)]) by exact integer arithmetic, independently of the oracle's code,
and compared with the oracle's empirical two-step laws by chi-squared tests.  Each
test also checks that a planted mistake in the law (p and q swapped, alpha dropped,
the common-neighbour term dropped) is rejected by the same data.
"""
from fractions import Fraction

import numpy as np
import pytest
from scipy import stats

import _oracle as O
import graphs as G

TWO32 = 1 << 32


def ceil_div(a, b):
    return -((-a) // b)


def pos_counts(d):
    # c_i = #{r : i*2^32 <= r*d < (i+1)*2^32}
    return [ceil_div((i + 1) * TWO32, d) - ceil_div(i * TWO32, d) for i in range(d)]


def nbrs(beg, csr, v):
    return [int(x) for x in csr[int(beg[v]):int(beg[v + 1])]]


def n2v_thresholds(p, q):
    ip, iq = 1.0 / p, 1.0 / q
    m = max(ip, 1.0, iq)
    t = lambda w: TWO32 if w >= m else int(np.floor(w / m * 4294967296.0))
    return t(ip), t(1.0), t(iq)


def n2v_law(beg, csr, u, v, p, q):
    tp, t1, tq = n2v_thresholds(p, q)
    nu = set(nbrs(beg, csr, u))
    nv = nbrs(beg, csr, v)
    cs = pos_counts(len(nv))
    w = {}
    for i, x in enumerate(nv):
        t = tp if x == u else (t1 if x in nu else tq)
        w[x] = w.get(x, 0) + cs[i] * t
    tot = sum(w.values())
    return {x: Fraction(a, tot) for x, a in w.items()}


def sopr_law(beg, csr, u, v, alpha, use_mixture):
    A = int(round(alpha * 65536))
    B = 65536 - A
    nu_l = nbrs(beg, csr, u)
    nu = set(nu_l)
    nv = nbrs(beg, csr, v)
    du, dv = len(nu_l), len(nv)
    cs = pos_counts(dv)
    match = [x in nu and x != u for x in nv]
    mode = O.lib().orc_sopr_use_mixture(du, dv, A, B)
    assert mode == use_mixture, "test graph does not exercise the intended sampler"
    w = {}
    if not mode:
        num, den = B * du, B * du + A * dv
        tnon = ceil_div(num * TWO32, den)
        for i, x in enumerate(nv):
            w[x] = w.get(x, 0) + cs[i] * (TWO32 if match[i] else tnon)
        tot = sum(w.values())
        return {x: Fraction(a, tot) for x, a in w.items()}
    c = sum(match)
    pi = Fraction(ceil_div(B * du * TWO32, B * du + A * c), TWO32) if c else Fraction(1)
    kc = pos_counts(c) if c else []
    law = {}
    k = 0
    for i, x in enumerate(nv):
        pr = pi * Fraction(cs[i], TWO32)
        if match[i]:
            pr += (1 - pi) * Fraction(kc[k], TWO32)
            k += 1
        law[x] = law.get(x, 0) + pr
    return law


def sopr_proposal_law(beg, csr, u, v, alpha):
    """Exact law of the proposal sampler: an attempt takes position i of N(v) with
    probability T_B c_i(dv) / 2^64 (T_B = B 2^16), or proposes position p of N(u) with
    probability (2^32 - T_B) c_p(du) / 2^64 and keeps it iff its value is in N(v) and
    is not u."""
    A = int(round(alpha * 65536))
    B = 65536 - A
    tb = B << 16
    nu = nbrs(beg, csr, u)
    nv = nbrs(beg, csr, v)
    sv = set(nv)
    cv, cu = pos_counts(len(nv)), pos_counts(len(nu))
    w = {}
    for i, x in enumerate(nv):
        w[x] = w.get(x, 0) + tb * cv[i]
    for p, y in enumerate(nu):
        if y != u and y in sv:
            w[y] = w.get(y, 0) + (TWO32 - tb) * cu[p]
    tot = sum(w.values())
    return {x: Fraction(a, tot) for x, a in w.items()}


def joint_law(beg, csr, s, step_law):
    """Law of (x1, x2) for a walker starting at s (prev = cur = s)."""
    first = step_law(s, s)
    out = {}
    for x1, p1 in first.items():
        for x2, p2 in step_law(s, x1).items():
            out[(x1, x2)] = p1 * p2
    return out


def empirical_pairs(beg, csr, cfg, s, n, wid0=0):
    src = np.full(n, s, dtype=np.uint32)
    o = O.run(beg, csr, cfg, n, sources=src, paths=True, hashes=False, state=False)
    p = o["paths"][:, 1:3]
    keys, cnt = np.unique(p[:, 0].astype(np.uint64) << 32 | p[:, 1].astype(np.uint64), return_counts=True)
    return {(int(k >> 32), int(k & 0xFFFFFFFF)): int(c) for k, c in zip(keys, cnt)}


def chi2_pvalue(law, emp, n):
    keys = sorted(set(law) | set(emp))
    exp = np.array([float(law.get(k, 0)) * n for k in keys])
    obs = np.array([emp.get(k, 0) for k in keys], dtype=float)
    if ((exp == 0) & (obs > 0)).any():
        return 0.0
    keep = exp > 0
    exp, obs = exp[keep], obs[keep]
    order = np.argsort(exp)
    exp, obs = exp[order], obs[order]
    # pool the smallest bins until every expected count is >= 5
    e2, o2, ae, ao = [], [], 0.0, 0.0
    for e, o in zip(exp, obs):
        ae += e
        ao += o
        if ae >= 5:
            e2.append(ae)
            o2.append(ao)
            ae = ao = 0.0
    if ae > 0:
        e2[-1] += ae
        o2[-1] += ao
    e2, o2 = np.array(e2), np.array(o2)
    e2 *= o2.sum() / e2.sum()
    return stats.chisquare(o2, e2).pvalue


N = 200000


@pytest.mark.parametrize("graph,s,p,q", [
    ("bridge", 3, 0.5, 2.0),
    ("bridge", 3, 2.0, 0.5),
    ("star", 1, 0.5, 2.0),
    ("rand", 7, 0.25, 4.0),
])
def test_node2vec_two_step_law(graph, s, p, q):
    if graph == "rand":
        beg, csr = G.random_undirected(40, 160, seed=3)
    else:
        beg, csr = getattr(G, graph)()
    cfg = O.walk_cfg(123, O.type_cfg(O.N2V, 2, p=p, q=q))
    emp = empirical_pairs(beg, csr, cfg, s, N)
    law = joint_law(beg, csr, s, lambda u, v: n2v_law(beg, csr, u, v, p, q))
    assert abs(sum(law.values()) - 1) < Fraction(1, 10**12)
    assert chi2_pvalue(law, emp, N) > 1e-3
    # planted mistake: p and q swapped must be rejected by the same sample
    bad = joint_law(beg, csr, s, lambda u, v: n2v_law(beg, csr, u, v, q, p))
    assert chi2_pvalue(bad, emp, N) < 1e-6


def hub_graph(hub_deg=2000):
    """0 -> {1, 2, 3}; 1 (hub) -> {0, 2, 3, 4, ..}; every other vertex -> {1}."""
    n = hub_deg + 2
    src = [0, 0, 0] + [1] * (n - 2) + list(range(2, n))
    dst = [1, 2, 3] + [0] + list(range(2, n - 1)) + [1] * (n - 2)
    return G.from_edges(n, src, dst)


@pytest.mark.parametrize("graph", ["hub", "bridge"])
def test_sopr_proposal_two_step_law(graph):
    beg, csr = hub_graph(2000) if graph == "hub" else G.bridge(5)
    s = 0 if graph == "hub" else 4
    alpha = 0.2
    cfg = O.walk_cfg(78, O.type_cfg(O.SOPR, 2, alpha=alpha, sampler=0))
    emp = empirical_pairs(beg, csr, cfg, s, N)
    law = joint_law(beg, csr, s, lambda u, v: sopr_proposal_law(beg, csr, u, v, alpha))
    assert chi2_pvalue(law, emp, N) > 1e-3
    # on duplicate-free lists it is the reference's law (1-a)/dv + a/du [x in N(u), x != u]
    A = Fraction(int(round(alpha * 65536)), 65536)
    for u, v in [(s, s)] + [(s, x) for x in nbrs(beg, csr, s)][:5]:
        nu = set(nbrs(beg, csr, u))
        nv = nbrs(beg, csr, v)
        wt = {x: (1 - A) / len(nv) + (A / len(nu) if (x in nu and x != u) else 0) for x in nv}
        tot = sum(wt.values())
        pl = sopr_proposal_law(beg, csr, u, v, alpha)
        for x in nv:
            assert abs(float(pl[x] - wt[x] / tot)) < 1e-7
    bad = joint_law(beg, csr, s, lambda u, v: sopr_law(beg, csr, u, v, 0.0, False))
    assert chi2_pvalue(bad, emp, N) < 1e-6


@pytest.mark.parametrize("case", ["mixture", "rejection"])
def test_sopr_two_step_law(case):
    if case == "mixture":
        beg, csr = hub_graph(2000)
        s = 0
    else:
        beg, csr = G.bridge(5)
        s = 4
    alpha = 0.2
    cfg = O.walk_cfg(77, O.type_cfg(O.SOPR, 2, alpha=alpha, sampler=1))
    emp = empirical_pairs(beg, csr, cfg, s, N)
    mix_for = lambda u, v: bool(O.lib().orc_sopr_use_mixture(int(beg[u + 1] - beg[u]), int(beg[v + 1] - beg[v]),
                                                             int(round(alpha * 65536)), 65536 - int(round(alpha * 65536))))
    law = joint_law(beg, csr, s, lambda u, v: sopr_law(beg, csr, u, v, alpha, mix_for(u, v)))
    if case == "mixture":
        assert mix_for(0, 1)
    assert chi2_pvalue(law, emp, N) > 1e-3
    # planted mistakes: alpha ignored (first-order walk), or common neighbours ignored
    bad = joint_law(beg, csr, s, lambda u, v: sopr_law(beg, csr, u, v, 0.0, False))
    assert chi2_pvalue(bad, emp, N) < 1e-6


def test_sopr_law_normalisation_closed_form():
    # exact law equals (1-a)/dv + a/du [x in N(u), x != u], normalised (up to the
    # 2^-32 granularity of the integer thresholds)
    beg, csr = hub_graph(300)
    alpha = 0.2
    A = int(round(alpha * 65536))
    for u, v in [(0, 1), (1, 2), (2, 1)]:
        nu = set(nbrs(beg, csr, u))
        nv = nbrs(beg, csr, v)
        du, dv = len(nu), len(nv)
        mode = bool(O.lib().orc_sopr_use_mixture(du, dv, A, 65536 - A))
        law = sopr_law(beg, csr, u, v, alpha, mode)
        a = Fraction(A, 65536)
        w = {x: (1 - a) / dv + (a / du if (x in nu and x != u) else 0) for x in nv}
        tot = sum(w.values())
        for x in nv:
            assert abs(float(law[x] - w[x] / tot)) < 1e-7


def test_geometric_termination_law():
    beg, csr = G.path(4)   # undirected path: no dead ends
    tau, L = 0.25, 12
    cfg = O.walk_cfg(5, O.type_cfg(O.N2V, L, p=1.0, q=1.0, term=tau))
    n = 100000
    o = O.run(beg, csr, cfg, n, sources=np.zeros(n, dtype=np.uint32), hashes=False)
    hops = o["state"][:, 2]
    t = Fraction(int(np.floor(tau * 4294967296.0)), TWO32)
    law = {k: (1 - t) ** k * t for k in range(L)}
    law[L] = (1 - t) ** L
    emp = dict(zip(*np.unique(hops, return_counts=True)))
    emp = {int(k): int(v) for k, v in emp.items()}
    assert chi2_pvalue(law, emp, n) > 1e-3
    st = o["state"][:, 3]
    assert set(np.unique(st)) <= {1, 2}
    assert ((st == 1) == (hops == L)).all()


def test_dead_end_ends_without_moving():
    beg, csr = G.dag_dead_end()
    cfg = O.walk_cfg(1, O.type_cfg(O.N2V, 10, p=1.0, q=1.0))
    o = O.run(beg, csr, cfg, 400, sources=np.zeros(400, dtype=np.uint32), paths=True)
    st = o["state"]
    assert (st[:, 3] == 3).all() and (st[:, 0] == 3).all()
    # path 0 -> ... -> 3, then no more moves
    for row, s in zip(o["paths"], st):
        assert row[s[2]] == 3 and (row[s[2] + 1:] == 0xFFFFFFFF).all()


def test_counters_consistent_with_paths():
    beg, csr = G.random_undirected(300, 1500, seed=2)
    cfg = O.walk_cfg(9, O.type_cfg(O.N2V, 20, p=0.5, q=2.0))
    o = O.run(beg, csr, cfg, 300, paths=True, visits=True)
    moves = (o["paths"][:, 1:] != 0xFFFFFFFF).sum()
    assert o["totals"]["steps"] == moves == o["visits"].sum() == o["state"][:, 2].sum()
    assert o["totals"]["trials"] >= moves
    # path hash = FNV-1a over the vertex sequence
    h = np.uint64(1469598103934665603)
    P = np.uint64(1099511628211)
    row = o["paths"][0]
    with np.errstate(over="ignore"):
        for x in row[row != 0xFFFFFFFF]:
            h = (h ^ np.uint64(x)) * P
    assert h == o["hashes"][0]


def test_reference_transcription_fixed_mode_matches_build_distribution():
    """Per-vertex visit counts of the build's Philox walk vs the fixed-mode transcription of
    the reference CPU step: two-sample chi-squared at alpha = 0.01, L1 distance within 3x
    the L1 between two independent seeds of the transcription."""
    beg, csr = G.random_undirected(200, 800, seed=11)
    nv = len(beg) - 1
    wpv, L = 200, 20
    cfg = O.walk_cfg(42, O.type_cfg(O.N2V, L, p=2.0, q=0.5))
    build = O.run(beg, csr, cfg, nv * wpv, wpv=wpv, visits=True, hashes=False, state=False)["visits"]
    ref_a = O.ref_run(beg, csr, O.N2V, L, p=2.0, q=0.5, wpv=wpv, seed=1, visits=True)["visits"]
    ref_b = O.ref_run(beg, csr, O.N2V, L, p=2.0, q=0.5, wpv=wpv, seed=2, visits=True)["visits"]
    lit = O.ref_run(beg, csr, O.N2V, L, p=2.0, q=0.5, wpv=wpv, seed=1, literal=True, visits=True)["visits"]
    keep = (build + ref_a) > 10
    table = np.vstack([build[keep], ref_a[keep]])
    p = stats.chi2_contingency(table)[1]
    assert p > 0.01
    l1 = lambda a, b: np.abs(a / a.sum() - b / b.sum()).sum()
    assert l1(build, ref_a) <= 3 * l1(ref_a, ref_b)
    # the literal (type-punned) reference draws are nearly first-order; reported, not required
    print("literal-mode L1 vs build: %.4f (noise floor %.4f)" % (l1(build, lit), l1(ref_a, ref_b)))


@pytest.mark.parametrize("sampler", [0, 1])
def test_reference_transcription_sopr_matches_build_distribution(sampler):
    beg, csr = G.power_law_directed(150, 1500, seed=4, multi=False)
    nv = len(beg) - 1
    wpv, L = 100, 10
    cfg = O.walk_cfg(42, O.type_cfg(O.SOPR, L, alpha=0.2, sampler=sampler))
    o = O.run(beg, csr, cfg, nv * wpv, wpv=wpv, visits=True, hashes=False, state=False)
    build = o["visits"]
    if sampler == 0:
        assert o["totals"]["mix_steps"] == 0
    ref_a = O.ref_run(beg, csr, O.SOPR, L, alpha=0.2, wpv=wpv, seed=1, visits=True)["visits"]
    ref_b = O.ref_run(beg, csr, O.SOPR, L, alpha=0.2, wpv=wpv, seed=2, visits=True)["visits"]
    keep = (build + ref_a) > 10
    p = stats.chi2_contingency(np.vstack([build[keep], ref_a[keep]]))[1]
    assert p > 0.01
    l1 = lambda a, b: np.abs(a / a.sum() - b / b.sum()).sum()
    assert l1(build, ref_a) <= 3 * l1(ref_a, ref_b)
