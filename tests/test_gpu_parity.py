"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle,
bit-exact, on the reference's datasets and seeded graphs, for every
virtual-wave width, with and without on-device work stealing."""
import os

import numpy as np
import pytest

import dumato_amd as dm
from oracle import pyoracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
WIDTHS = (8, 16, 32, 64)


def cfg(**kw):
    kw.setdefault("watchdog_s", 60.0)
    return dm.config(**kw)


@pytest.fixture(scope="module")
def citeseer():
    path = os.path.join(GOLD, "datasets", "citeseer.edgelist")
    return dm.Graph.from_edgelist(path), O.load_edgelist(path)


def test_tiny_test_edgelist():
    path = os.path.join(GOLD, "datasets", "test.edgelist")
    print("case test.edgelist", flush=True)
    g = dm.Graph.from_edgelist(path)
    ref = O.load_edgelist(path)
    for k in (2, 3, 4):
        assert g.clique_count(k, cfg()) == O.clique_count(ref, k)
    for k in (3, 4):
        tot, hist = g.motif_count(k, cfg())
        wt, wh = O.motif_count(ref, k)
        assert tot == wt and np.array_equal(hist, wh)


@pytest.mark.parametrize("w", WIDTHS)
def test_citeseer_cliques(citeseer, w):
    g, ref = citeseer
    for k in (3, 4, 5, 6):
        print("case citeseer clique k=%d w=%d" % (k, w), flush=True)
        want = O.clique_count(ref, k)
        assert g.clique_count(k, cfg(vwave=w)) == want
        assert g.clique_count(k, cfg(vwave=w, steal=0)) == want


@pytest.mark.parametrize("w", WIDTHS)
def test_citeseer_motifs(citeseer, w):
    g, ref = citeseer
    for k in (3, 4, 5):
        print("case citeseer motif k=%d w=%d" % (k, w), flush=True)
        wt, wh = O.motif_count(ref, k)
        for steal in (1, 0):
            tot, hist = g.motif_count(k, cfg(vwave=w, steal=steal))
            assert tot == wt and np.array_equal(hist, wh)


def test_citeseer_motif6(citeseer):
    g, ref = citeseer
    print("case citeseer motif k=6", flush=True)
    wt, wh = O.motif_count(ref, 6)
    tot, hist = g.motif_count(6, cfg(vwave=32))
    assert tot == wt and np.array_equal(hist, wh)


@pytest.mark.parametrize("seed", range(3))
def test_random_dense_graphs(seed):
    s, d, n = G.gnp(60, 0.3, 500 + seed)
    g = dm.Graph.from_edges(s, d, n)
    ref = O.csr_from_edges(s, d, n)
    for k in (3, 4, 5, 6, 7):
        print("case gnp seed=%d clique k=%d" % (seed, k), flush=True)
        assert g.clique_count(k, cfg(vwave=WIDTHS[seed % 4])) == O.clique_count(ref, k)
    for k in (3, 4, 5):
        print("case gnp seed=%d motif k=%d" % (seed, k), flush=True)
        wt, wh = O.motif_count(ref, k)
        tot, hist = g.motif_count(k, cfg(vwave=WIDTHS[(seed + 1) % 4]))
        assert tot == wt and np.array_equal(hist, wh)


def test_complete_graph_closed_form():
    import math
    s, d, n = G.complete(24)
    g = dm.Graph.from_edges(s, d, n)
    for k in (3, 4, 6, 8, 10, 12):
        print("case K24 clique k=%d" % k, flush=True)
        assert g.clique_count(k, cfg()) == math.comb(24, k)
    tot, hist = g.motif_count(5, cfg())
    assert tot == math.comb(24, 5) and hist[32] == tot


@pytest.mark.parametrize("bitmap", [1, 0])
def test_per_root_counts_and_bytes(citeseer, bitmap):
    g, ref = citeseer
    perm, ro = O.relabel(ref, oriented=True)
    for k in (3, 4, 5, 6):
        print("case per-root clique k=%d bitmap=%d" % (k, bitmap), flush=True)
        tot = g.clique_count(k, cfg(per_root=1, vwave=16, clique_bitmap=bitmap))
        pr = g.last_root_counts()
        st = g.last_stats()
        _, opr_rel = O.clique_count(ro, k, per_root=True)
        # roots are the vertices in degree order: map the oracle's oriented
        # per-root counts back to original ids
        assert np.array_equal(pr, opr_rel[perm]) and int(pr.sum()) == tot
        assert st["alg_bytes"] == O.clique_bytes(ro, k)
    for k in (3, 4):
        print("case per-root motif k=%d" % k, flush=True)
        tot, _ = g.motif_count(k, cfg(per_root=1, vwave=16))
        assert int(g.last_root_counts().sum()) == tot
        assert g.last_stats()["alg_bytes"] == O.motif_bytes(ref, k)


def test_root_range_subsets(citeseer):
    g, ref = citeseer
    print("case root range", flush=True)
    a = g.clique_count(4, cfg(root_begin=0, root_end=1500))
    b = g.clique_count(4, cfg(root_begin=1500, root_end=0))
    assert a + b == O.clique_count(ref, 4)
    ta, ha = g.motif_count(4, cfg(root_begin=0, root_end=1000))
    tb, hb = g.motif_count(4, cfg(root_begin=1000))
    wt, wh = O.motif_count(ref, 4)
    assert ta + tb == wt and np.array_equal(ha + hb, wh)


def test_errors_are_reported():
    g = dm.Graph.from_edges(*G.complete(6)[:2])
    with pytest.raises(dm.DMError):
        g.clique_count(1)
    with pytest.raises(dm.DMError):
        g.motif_count(8)
    with pytest.raises(dm.DMError):
        g.clique_count(4, cfg(vwave=12))
