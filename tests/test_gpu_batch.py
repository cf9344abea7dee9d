"""Batched queries over many graphs in one hipGraph-replayed persistent
launch: every query's result must equal the single-graph path and the oracle,
on every replay, with eager work between replays."""
import numpy as np
import pytest

import dumato_amd as dm
from dumato_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

QUERIES_PER_GRAPH = [("clique", 3), ("clique", 4), ("clique", 5), ("clique", 6),
                     ("motif", 3), ("motif", 4), ("motif", 5)]


def small_specs(count):
    specs = [s for s in W.batch_specs(256) if s["n"] <= 1500 and s["m"] <= 12000]
    return specs[:count]


def test_batch_matches_single_and_oracle():
    specs = small_specs(10)
    assert len(specs) == 10
    graphs = [W.make_batch_graph(s) for s in specs]
    queries = [(gi, mode, k) for gi in range(len(graphs)) for (mode, k) in QUERIES_PER_GRAPH]
    b = dm.Batch(graphs)
    want = []
    for gi, mode, k in queries:
        g = graphs[gi]
        n, _, _ = g.info()
        ref = O.CSR(n, *g.csr())
        if mode == "clique":
            w = O.clique_count(ref, k)
            assert g.clique_count(k) == w
        else:
            w = O.motif_count(ref, k)[1]
            assert np.array_equal(g.motif_count(k)[1], w)
        want.append(w)
    for rep in range(3):
        print("case batch replay %d" % rep, flush=True)
        got, lat = b.run(queries)
        for (gi, mode, k), gv, wv in zip(queries, got, want):
            if mode == "clique":
                assert gv == wv, (gi, mode, k)
            else:
                assert np.array_equal(gv, wv), (gi, mode, k)
        assert np.all(lat >= 0) and lat.max() <= b.last_stats()["device_ms"] * 1.05 + 1.0
        # eager work between replays
        graphs[rep].clique_count(4)
        graphs[rep + 1].motif_count(3)


def test_batch_single_query_and_shapes():
    specs = small_specs(4)
    graphs = [W.make_batch_graph(s) for s in specs]
    b = dm.Batch(graphs)
    for w in (16, 64):
        got, _ = b.run([(2, "motif", 4)], dm.config(vwave=w))
        assert np.array_equal(got[0], graphs[2].motif_count(4)[1])
        got, _ = b.run([(1, "clique", 5), (3, "clique", 3)], dm.config(vwave=w, steal_min=8))
        assert got == [graphs[1].clique_count(5), graphs[3].clique_count(3)]


def test_batch_autotune_keeps_results():
    specs = small_specs(6)
    graphs = [W.make_batch_graph(s) for s in specs]
    queries = [(gi, mode, k) for gi in range(len(graphs)) for (mode, k) in QUERIES_PER_GRAPH]
    b = dm.Batch(graphs)
    base, _ = b.run(queries)
    best = b.autotune(queries, budget_ms=500.0)
    assert best.vwave in (16, 32, 64)
    got, _ = b.run(queries)
    for x, y in zip(base, got):
        assert np.array_equal(np.asarray(x), np.asarray(y))


def test_batch_rejects_bad_queries():
    b = dm.Batch([W.make_batch_graph(small_specs(1)[0])])
    with pytest.raises(dm.DMError):
        b.run([(1, "clique", 4)])
    with pytest.raises(dm.DMError):
        b.run([(0, "motif", 7)])
