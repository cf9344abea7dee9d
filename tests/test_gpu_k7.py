"""k = 7 motif histograms on the GPU against the CPU oracle and brute force.

The reference ships no dictionary for k = 7, so the class ids are this
library's own (its table ranks each isomorphism class by its largest quick
code; the CPU test test_pattern_table_k7_is_isomorphism_classes checks that the
classes are exactly the isomorphism classes).  The ids are unpinned against the
reference; the counts per class are pinned here: the oracle's DFS and the
brute force over 7-subsets classify with the same table, and closed forms fix
complete graphs, cycles and paths."""
from math import comb

import numpy as np
import pytest

import dumato_amd as dm
from oracle import pyoracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu

K = 7


@pytest.fixture(scope="module")
def table():
    q2c, _ = dm.pattern_table(K)
    return q2c, dm.num_classes(K)


def run(src, dst, n, vw):
    g = dm.Graph.from_edges(src, dst, n)
    return g.motif_count(K, dm.config(vwave=vw, watchdog_s=60.0))


@pytest.mark.parametrize("vw", [8, 16, 32, 64])
def test_k7_random_graphs_vs_oracle_and_brute(table, vw):
    for seed, (n, p) in enumerate([(22, 0.3), (26, 0.22), (40, 0.12)]):
        s, d, n = G.gnp(n, p, seed + 100 * vw)
        print("case k7 gnp n=%d p=%.2f vwave %d" % (n, p, vw), flush=True)
        tot, hist = run(s, d, n, vw)
        ref = O.csr_from_edges(s, d, n)
        wt, wh = O.motif_count(ref, K, table=table)
        assert tot == wt and np.array_equal(hist, wh)
        if n <= 26:
            bt, bh = O.brute_motif(ref, K, table=table)
            assert tot == bt and np.array_equal(hist, bh)


def test_k7_chung_lu_vs_oracle(table):
    g = dm.Graph.chung_lu(120, 330, 2.3, 15.0, 77)
    n, _, _ = g.info()
    ref = O.CSR(n, *g.csr())
    wt, wh = O.motif_count(ref, K, table=table)
    for vw in (16, 64):
        print("case k7 chung-lu n=120 vwave %d" % vw, flush=True)
        tot, hist = g.motif_count(K, dm.config(vwave=vw, watchdog_s=60.0))
        assert tot == wt and np.array_equal(hist, wh)


def test_k7_closed_forms(table):
    q2c, _ = table
    # K_9: every 7-subset is a K_7, all in one class
    tot, hist = run(*G.complete(9), 64)
    assert tot == comb(9, 7) and hist.max() == comb(9, 7)
    assert np.argmax(hist) == q2c[(1 << 20) - 1]
    # cycle C_10: the connected 7-subsets are its 10 paths; path P_12: 6 paths
    tot, hist = run(*G.cycle(10), 32)
    assert tot == 10 and np.count_nonzero(hist) == 1
    tp, hp = run(*G.path(12), 16)
    assert tp == 6 and np.array_equal(np.nonzero(hp)[0], np.nonzero(hist)[0])
