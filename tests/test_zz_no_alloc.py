"""dm_count_noalloc (counting whose device scratch is a static array) against
the allocating entry points dm_motif_count / dm_clique_count, through the C
ABI, on calls in a row with a smaller input second, and against the oracle.

Collected last: it shares the GPU's static array with nothing else in the
suite, and a failure here must not hide the tests before it."""
import numpy as np
import pytest

import dumato_amd as dm
from dumato_amd import workloads as W
from oracle import pyoracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu

CITESEER = "tests/golden/datasets/citeseer.edgelist"


def test_noalloc_equals_allocating_calls_larger_then_smaller():
    big = W.chunglu_18k()
    small = dm.Graph.from_edgelist(CITESEER)
    print("case noalloc chunglu_18k motif k=4", flush=True)
    wt, wh = big.motif_count(4)
    t, h = big.count_noalloc(True, 4)
    assert t == wt and np.array_equal(h, wh)
    print("case noalloc chunglu_18k clique k=4", flush=True)
    assert big.count_noalloc(False, 4)[0] == big.clique_count(4)
    # second call: a smaller input (the array keeps the first call's contents)
    ref = O.load_edgelist(CITESEER)
    for k in (3, 4, 5):
        print("case noalloc citeseer motif k=%d" % k, flush=True)
        wt, wh = small.motif_count(k)
        t, h = small.count_noalloc(True, k)
        assert t == wt and np.array_equal(h, wh)
        ot, oh = O.motif_count(ref, k)
        assert t == ot and np.array_equal(h, oh)
    for k in (3, 4, 5, 6):
        print("case noalloc citeseer clique k=%d" % k, flush=True)
        c = small.count_noalloc(False, k)[0]
        assert c == small.clique_count(k) == O.clique_count(ref, k)


def test_noalloc_widths_and_repeats():
    g = dm.Graph.from_edgelist(CITESEER)
    g.clique_count(3)  # uploads the graph
    ref = O.load_edgelist(CITESEER)
    ot, oh = O.motif_count(ref, 4)
    for vw in (8, 16, 32, 64):
        for rep in range(2):
            print("case noalloc citeseer motif k=4 vwave %d rep %d" % (vw, rep), flush=True)
            t, h = g.count_noalloc(True, 4, dm.config(vwave=vw))
            assert t == ot and np.array_equal(h, oh)


def test_noalloc_rejections():
    small = dm.Graph.from_edges(*G.complete(6)[:2])
    with pytest.raises(dm.DMError) as e:
        small.count_noalloc(True, 3)  # never uploaded
    assert "not on the GPU" in str(e.value)
    small.clique_count(3)
    with pytest.raises(dm.DMError):
        small.count_noalloc(True, 3, dm.config(sched=dm.DM_SCHED_LEARNED))
    with pytest.raises(dm.DMError):
        small.count_noalloc(False, 3, dm.config(per_root=1))
    # a hub of degree 2^20: the k=6 motif arena of one block exceeds the array
    s, d, n = G.star(1 << 20)
    hub = dm.Graph.from_edges(s, d, n)
    with pytest.raises(dm.DMError) as e:
        hub.count_noalloc(True, 6)
    assert "static scratch" in str(e.value)
