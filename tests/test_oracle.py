"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against something that does not share its
code: the reference's own known answers on its shipped datasets, brute force
over k-subsets (C and pure Python, neither using the canonical DFS rule),
closed forms on complete graphs, stars, paths and cycles, and invariants.
"""
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from tests import graphs as G

GOLD = os.path.join(os.path.dirname(__file__), "golden")
KNOWN = json.load(open(os.path.join(GOLD, "known_answers.json")))


def csr(t, keep_loops=False):
    s, d, n = t
    return O.csr_from_edges(s, d, n, keep_loops)


def class_of(edges, k):
    """Class id of one small connected graph given as edges on 0..k-1."""
    s = np.array([e[0] for e in edges], np.int32)
    d = np.array([e[1] for e in edges], np.int32)
    g = O.csr_from_edges(s, d, k)
    tot, hist = O.brute_motif(g, k)
    assert tot == 1
    return int(np.nonzero(hist)[0][0])


# ---- graph input -----------------------------------------------------------
@pytest.mark.parametrize("name", ["test.edgelist", "citeseer.edgelist"])
def test_reader_invariants(name):
    g = O.load_edgelist(os.path.join(GOLD, "datasets", name))
    deg = g.degrees()
    assert deg.sum() == g.nnz and g.nnz % 2 == 0
    for v in range(g.n):
        nb = g.neighbours(v)
        assert np.all(np.diff(nb) > 0)
        assert not np.any(nb == v)
    # symmetry
    pairs = set()
    for v in range(g.n):
        for u in g.neighbours(v):
            pairs.add((v, int(u)))
    assert all((u, v) in pairs for (v, u) in pairs)
    if name == "citeseer.edgelist":
        kn = KNOWN[name]
        assert (g.n, g.nnz // 2, int(deg.max())) == (kn["n"], kn["m"], kn["maxdeg"])


def test_reader_rejects_malformed(tmp_path):
    p = tmp_path / "bad.edgelist"
    p.write_text("0 1\n1 2\n3\n")
    with pytest.raises(ValueError):
        O.load_edgelist(str(p))
    p.write_text("0 1\n1 x\n")
    with pytest.raises(ValueError):
        O.load_edgelist(str(p))
    p.write_text("0 1\n-1 2\n")
    with pytest.raises(ValueError):
        O.load_edgelist(str(p))


def test_reader_dedup_and_isolated(tmp_path):
    p = tmp_path / "g.edgelist"
    p.write_text("0 1\n1 0\n0 1\n4 2\n")
    g = O.load_edgelist(str(p))
    assert g.n == 5 and g.nnz == 4
    assert list(g.neighbours(3)) == []
    assert list(g.neighbours(0)) == [1] and list(g.neighbours(2)) == [4]


# ---- known answers on the reference datasets --------------------------------
@pytest.mark.parametrize("name", ["test.edgelist", "citeseer.edgelist"])
def test_known_answers(name):
    g = O.load_edgelist(os.path.join(GOLD, "datasets", name))
    kn = KNOWN[name]
    for k, want in kn["clique"].items():
        assert O.clique_count(g, int(k)) == want
        assert O.clique_count(g, int(k), faithful=True) == want
    for k, want in kn["motif"].items():
        for faithful in (False, True):
            tot, hist = O.motif_count(g, int(k), faithful=faithful)
            assert tot == want["total"]
            got = {str(i): int(x) for i, x in enumerate(hist) if x}
            assert got == want["hist"]
    if name == "test.edgelist":
        for k in (3, 4):
            assert O.brute_clique(g, k) == kn["clique"][str(k)]
            tot, hist = O.brute_motif(g, k)
            assert {str(i): int(x) for i, x in enumerate(hist) if x} == kn["motif"][str(k)]["hist"]


# ---- brute force ------------------------------------------------------------
@pytest.mark.parametrize("seed", range(6))
def test_c_brute_matches_python_brute(seed):
    t = G.gnp(9, 0.45, seed)
    g = csr(t)
    for k in (3, 4, 5):
        lut, _ = O.load_lut(k)
        pc, ph = G.py_brute(*t, k, lut)
        assert O.brute_clique(g, k) == pc
        tot, hist = O.brute_motif(g, k)
        assert {i: int(x) for i, x in enumerate(hist) if x} == ph
        assert tot == sum(ph.values())


@pytest.mark.parametrize("seed,n,p", [(s, n, p) for s in range(4) for (n, p) in
                                      ((14, 0.5), (20, 0.3), (24, 0.2), (16, 0.8))])
def test_dfs_matches_brute(seed, n, p):
    g = csr(G.gnp(n, p, 100 + seed))
    for k in (3, 4, 5, 6):
        assert O.clique_count(g, k) == O.brute_clique(g, k)
        assert O.clique_count(g, k, faithful=True) == O.brute_clique(g, k)
        bt, bh = O.brute_motif(g, k)
        for faithful in (False, True):
            t, h = O.motif_count(g, k, faithful=faithful)
            assert t == bt
            assert np.array_equal(h, bh)


# ---- closed forms ------------------------------------------------------------
@pytest.mark.parametrize("n", [5, 9, 13])
def test_complete_graph(n):
    g = csr(G.complete(n))
    for k in (2, 3, 4, 5, 6):
        assert O.clique_count(g, k) == math.comb(n, k)
    for k in (3, 4, 5, 6):
        tot, hist = O.motif_count(g, k)
        kk = class_of(list(itertools.combinations(range(k), 2)), k)
        assert tot == math.comb(n, k) and hist[kk] == tot


@pytest.mark.parametrize("m", [4, 7, 12])
def test_star_path_cycle(m):
    for k in (3, 4, 5, 6):
        star_cls = class_of([(0, i) for i in range(1, k)], k)
        path_cls = class_of([(i, i + 1) for i in range(k - 1)], k)
        tot, hist = O.motif_count(csr(G.star(m)), k)
        assert tot == math.comb(m, k - 1) and hist[star_cls] == tot
        n = m + k
        tot, hist = O.motif_count(csr(G.path(n)), k)
        assert tot == n - k + 1 and hist[path_cls] == tot
        tot, hist = O.motif_count(csr(G.cycle(n)), k)
        assert tot == n and hist[path_cls] == n
        assert O.clique_count(csr(G.star(m)), 3) == 0


# ---- invariants -------------------------------------------------------------
def test_relabel_invariance_and_clique_bucket():
    t = G.gnp(60, 0.15, 7)
    g = csr(t)
    gp = csr(G.permute(*t, seed=3))
    perm, gd = O.relabel(g, oriented=False)
    assert sorted(perm.tolist()) == list(range(g.n))
    for k in (3, 4, 5):
        a = O.motif_count(g, k)
        b = O.motif_count(gp, k)
        c = O.motif_count(gd, k)
        assert a[0] == b[0] == c[0]
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[1], c[1])
        kk = class_of(list(itertools.combinations(range(k), 2)), k)
        assert a[1][kk] == O.clique_count(g, k)
        assert O.clique_count(g, k) == O.clique_count(gp, k) == O.clique_count(gd, k)


def test_oriented_relabel_keeps_clique_counts():
    g = O.load_edgelist(os.path.join(GOLD, "datasets", "citeseer.edgelist"))
    perm, go = O.relabel(g, oriented=True)
    deg = g.degrees()
    # orientation follows (degree, id) rank
    for v in range(0, g.n, 97):
        for u in g.neighbours(v):
            a, b = (v, int(u)) if (deg[v], v) < (deg[u], u) else (int(u), v)
            assert perm[b] in go.neighbours(perm[a])
    for k in (3, 4, 5, 6):
        assert O.clique_count(go, k) == KNOWN["citeseer.edgelist"]["clique"][str(k)]


def test_self_loops_do_not_change_counts():
    s, d, n = G.gnp(30, 0.25, 11)
    loops = np.arange(0, n, 3, dtype=np.int32)
    s2, d2 = np.concatenate([s, loops]), np.concatenate([d, loops])
    a = O.csr_from_edges(s, d, n)
    b = O.csr_from_edges(s2, d2, n, keep_loops=True)
    assert b.nnz == a.nnz + len(loops)
    for k in (3, 4, 5):
        assert O.clique_count(a, k) == O.clique_count(b, k, faithful=True)
        ta, ha = O.motif_count(a, k)
        tb, hb = O.motif_count(b, k, faithful=True)
        assert ta == tb and np.array_equal(ha, hb)


def test_per_root_sums_and_root_lists():
    g = csr(G.gnp(80, 0.1, 5))
    for k in (3, 4):
        tot, pr = O.clique_count(g, k, per_root=True)
        assert int(pr.sum()) == tot
        roots = np.arange(g.n - 1, -1, -2, dtype=np.int32)
        t2, pr2 = O.clique_count_roots(g, k, roots)
        assert np.array_equal(pr2, pr[roots]) and t2 == int(pr[roots].sum())
        mt, mh, mpr = O.motif_count(g, k, per_root=True)
        assert int(mpr.sum()) == mt == int(mh.sum())
        t3, h3, pr3 = O.motif_count_roots(g, k, roots)
        assert np.array_equal(pr3, mpr[roots])


# ---- algorithmic bytes --------------------------------------------------------
@pytest.mark.parametrize("n", [6, 10])
def test_bytes_closed_forms(n):
    g = csr(G.complete(n))
    _, go = O.relabel(g, oriented=True)
    for k in (3, 4, 5):
        assert O.clique_bytes(go, k) == 4 * sum(math.comb(n, j) for j in range(2, k + 1))
        want = 4 * ((n - 1) * sum(j * math.comb(n, j) for j in range(1, k))
                    + (k - 1) * math.comb(n, k))
        assert O.motif_bytes(g, k) == want
    m = n
    gs = csr(G.star(m))
    for k in (3, 4, 5):
        inner = 2 * m + sum(math.comb(m, j - 1) * (m + j - 1) for j in range(2, k))
        assert O.motif_bytes(gs, k) == 4 * (inner + (k - 1) * math.comb(m, k - 1))


def _py_near_clique(src, dst, n, k):
    """Pure-Python count of the near-clique query by k-subsets."""
    adj = [set() for _ in range(n)]
    for a, b in zip(src.tolist(), dst.tolist()):
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    total = 0
    for S in itertools.combinations(range(n), k):
        missing = sum(1 for a, b in itertools.combinations(S, 2) if b not in adj[a])
        total += math.comb(k, 2) if missing == 0 else (1 if missing == 1 else 0)
    return total


@pytest.mark.parametrize("k", [4, 5, 6])
def test_near_clique_query_pinned(k):
    # q4_PA .. q13_PA: closed form on complete graphs, the C brute force and a
    # pure-Python brute force on random graphs, faithful = fast
    for n in (6, 9):
        g = O.csr_from_edges(*G.complete(n)[:2], n)
        assert O.near_clique_count(g, k) == math.comb(n, k) * math.comb(k, 2)
    for seed in range(4):
        s, d, n = G.gnp(13, 0.5, 40 + seed)
        g = O.csr_from_edges(s, d, n)
        want = _py_near_clique(s, d, n, k)
        assert O.brute_near_clique(g, k) == want
        assert O.near_clique_count(g, k) == want == O.near_clique_count(g, k, faithful=True)
    # one missing edge in K_k: exactly one occurrence
    s, d, n = G.complete(k)
    keep = ~((s == 0) & (d == 1))
    assert O.near_clique_count(O.csr_from_edges(s[keep], d[keep], n), k) == 1
