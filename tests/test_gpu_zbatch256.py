"""The full mixed batch: 256 seeded graphs x {clique k=3..6, motif k=3..5}
(1,792 queries) in one hipGraph-captured launch, every result against the
oracle's values in tests/golden/batch256_counts.json (make_batch_golden.py).

Cliques and motifs k=3, 4 have full oracle values for every graph, and so
have k=5 motifs on the smaller graphs.  For the larger graphs' k=5 motifs the
oracle holds histograms of three root ranges each: the single-graph path is
checked on those ranges (cfg.root_begin / root_end), and the batch result must
equal the single-graph path's full histogram.

Collected after the single-graph workload tests: its graphs take a while to
build, and a failure here should not hide them."""
import hashlib
import json
import os

import numpy as np
import pytest

import dumato_amd as dm
from dumato_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "batch256_counts.json")))


def test_batch256_every_query_pinned():
    specs = W.batch_specs(256)
    assert len(GOLD["graphs"]) == len(specs) == 256
    assert [tuple(q) for q in GOLD["queries"]] == W.BATCH_QUERIES
    graphs = []
    for gi, spec in enumerate(specs):
        g = W.make_batch_graph(spec)
        rec = GOLD["graphs"][gi]
        n, m, md = g.info()
        off, adj = g.csr()
        assert (n, m, md) == (rec["n"], rec["m"], rec["maxdeg"]), gi
        assert hashlib.sha256(off.tobytes() + adj.tobytes()).hexdigest() == rec["csr_sha256"], gi
        graphs.append(g)
    print("case batch256: graphs built", flush=True)
    queries = [(gi, mode, k) for gi in range(len(graphs)) for (mode, k) in W.BATCH_QUERIES]
    assert len(queries) == 1792
    b = dm.Batch(graphs)
    print("case batch256: one launch of %d queries" % len(queries), flush=True)
    res, lat = b.run(queries)
    st = b.last_stats()
    assert np.all(lat > 0) and lat.max() <= st["device_ms"] * 1.05 + 1.0
    ranged = 0
    for (gi, mode, k), r in zip(queries, res):
        rec = GOLD["graphs"][gi]
        if mode == "clique":
            assert r == rec["clique"][k - 3], (gi, mode, k)
        elif k in (3, 4):
            assert np.array_equal(r, rec["motif%d" % k]), (gi, mode, k)
        elif rec["motif5"] is not None:
            assert np.array_equal(r, rec["motif5"]), (gi, mode, k)
        else:
            # the single-graph path: pinned on the oracle's root ranges, and
            # equal to the batch on the whole graph
            g = graphs[gi]
            for a, bb, want in rec["motif5_ranges"]:
                _, h = g.motif_count(5, dm.config(root_begin=a, root_end=bb))
                assert np.array_equal(h, want), (gi, a, bb)
            _, full = g.motif_count(5)
            assert np.array_equal(r, full), (gi, mode, k)
            ranged += 1
    print("case batch256: %d motif k=5 queries pinned through root ranges" % ranged, flush=True)
    assert ranged == sum(1 for rec in GOLD["graphs"] if rec["motif5"] is None)
