"""The DFS-wide step API (include/dumato_device.h) and the host-driven load
balancer (dm_dfs_run): kernels in the shape of the reference's clique_PA,
clique_DM_WCV8 and motifs_DM_WCLB against the oracle, at every worker width,
and with forced stop / snapshot / donate / relaunch cycles against the
full-size golden counts."""
import json
import os

import numpy as np
import pytest

import dumato_amd as dm
from dumato_amd import workloads as W
from oracle import pyoracle as O
from tests import graphs as G

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
WORK = json.load(open(os.path.join(GOLD, "workload_counts.json")))
CITESEER = os.path.join(GOLD, "datasets", "citeseer.edgelist")


@pytest.fixture(scope="module")
def citeseer():
    return dm.Graph.from_edgelist(CITESEER), O.load_edgelist(CITESEER)


@pytest.mark.parametrize("vw", [8, 16, 32, 64])
def test_clique_kernels_match_oracle(citeseer, vw):
    g, ref = citeseer
    for k in (3, 4, 5, 6):
        want = O.clique_count(ref, k)
        for kind in (dm.DM_DFS_CLIQUE_PA, dm.DM_DFS_CLIQUE_WC):
            print("case dfs clique k=%d kind %d vwave %d" % (k, kind, vw), flush=True)
            s = dm.Dfs(g, k, False, dm.dfs_config(vwave=vw, threshold=100))
            tot, out, st = s.run(dm.dfs_kernel(kind, vw))
            assert tot == want and st["launches"] == 1


@pytest.mark.parametrize("vw", [8, 16, 32, 64])
def test_motif_kernel_matches_oracle(citeseer, vw):
    g, ref = citeseer
    for k in (3, 4, 5, 6):
        if k == 6 and vw not in (16, 64):
            continue
        print("case dfs motif k=%d vwave %d" % (k, vw), flush=True)
        wt, wh = O.motif_count(ref, k)
        s = dm.Dfs(g, k, True, dm.dfs_config(vwave=vw, threshold=100))
        tot, hist, _ = s.run()
        assert tot == wt and np.array_equal(hist, wh)


def test_forced_preemption_motif_chunglu():
    # stops every 3 ms: the state goes to the host, the heaviest workers'
    # lowest-level tails move to idle workers, the kernel resumes
    g = W.chunglu_18k()
    gold = WORK["chunglu_18k"]["motif"]["4"]
    for vw in (32, 64):
        print("case dfs forced stops chunglu motif k=4 vwave %d" % vw, flush=True)
        s = dm.Dfs(g, 4, True, dm.dfs_config(vwave=vw, threshold=100, stop_every_us=3000, max_rebalances=12,
                                             poll_us=500))
        tot, hist, st = s.run()
        print("  rebalances %d donated %d launches %d total %.1f ms" %
              (st["rebalances"], st["donated_jobs"], st["launches"], st["total_ms"]), flush=True)
        assert tot == gold["total"]
        assert {str(i): int(x) for i, x in enumerate(hist) if x} == gold["hist"]
        assert st["rebalances"] >= 2 and st["launches"] == st["rebalances"] + 1


def test_idle_triggered_balancing_cliques_rmat():
    # the reference's mode: rebalance whenever more than threshold % idle
    g = W.rmat_s17()
    want = WORK["rmat_s17"]["clique"]["5"]
    for thr in (0, 30):
        print("case dfs rmat clique k=5 threshold %d" % thr, flush=True)
        s = dm.Dfs(g, 5, False, dm.dfs_config(vwave=32, threshold=thr, poll_us=500, max_rebalances=20))
        tot, _, st = s.run()
        print("  rebalances %d donated %d total %.1f ms" % (st["rebalances"], st["donated_jobs"], st["total_ms"]),
              flush=True)
        assert tot == want


def test_idle_balancing_small_graph_every_kind(citeseer):
    g, ref = citeseer
    for k in (4, 5):
        print("case dfs citeseer threshold 0 k=%d" % k, flush=True)
        cfg = dm.dfs_config(vwave=16, threshold=0, poll_us=100, jobs_per_worker=4)
        assert dm.Dfs(g, k, False, cfg).run()[0] == O.clique_count(ref, k)
        wt, wh = O.motif_count(ref, k)
        tot, hist, _ = dm.Dfs(g, k, True, cfg).run()
        assert tot == wt and np.array_equal(hist, wh)


@pytest.mark.parametrize("vw", [16, 64])
def test_near_clique_queries_match_oracle(citeseer, vw):
    # q4_PA .. q13_PA: the near-clique pattern of each size, without and with
    # host-driven balancing
    g, ref = citeseer
    for k in (4, 5, 6, 7):
        want = O.near_clique_count(ref, k)
        for thr in (100, 0):
            print("case dfs near-clique k=%d vwave %d threshold %d" % (k, vw, thr), flush=True)
            cfg = dm.dfs_config(vwave=vw, threshold=thr, poll_us=100)
            assert dm.Dfs(g, k, dm.DM_PATTERN_NEAR_CLIQUE, cfg).run()[0] == want
    # dense enough that the largest queries still have occurrences
    s, d, n = G.gnp(28, 0.85, 5)
    dense = dm.Graph.from_edges(s, d, n)
    dref = O.csr_from_edges(s, d, n)
    assert O.near_clique_count(dref, 13) > 0
    for k in range(4, 14):
        print("case dfs near-clique dense k=%d vwave %d" % (k, vw), flush=True)
        cfg = dm.dfs_config(vwave=vw, threshold=100)
        assert dm.Dfs(dense, k, dm.DM_PATTERN_NEAR_CLIQUE, cfg).run()[0] == O.near_clique_count(dref, k)


def test_near_clique_forced_preemption_rmat():
    g = W.rmat_s17()
    want = WORK["rmat_s17"]["near_clique"]["5"]
    print("case dfs near-clique rmat k=5 forced stops", flush=True)
    cfg = dm.dfs_config(vwave=32, threshold=100, stop_every_us=2000, max_rebalances=8, poll_us=500)
    tot, _, st = dm.Dfs(g, 5, dm.DM_PATTERN_NEAR_CLIQUE, cfg).run()
    print("  rebalances %d donated %d total %.1f ms" % (st["rebalances"], st["donated_jobs"], st["total_ms"]),
          flush=True)
    assert tot == want


def test_kernel_of_another_width_counts_nothing(citeseer):
    g, _ = citeseer
    s = dm.Dfs(g, 4, False, dm.dfs_config(vwave=32, threshold=100))
    print("case dfs width mismatch", flush=True)
    tot, _, _ = s.run(dm.dfs_kernel(dm.DM_DFS_CLIQUE_PA, 64))
    assert tot == 0
    assert s.run()[0] == O.clique_count(O.load_edgelist(CITESEER), 4)


def test_dfs_argument_errors(citeseer):
    g, _ = citeseer
    with pytest.raises(dm.DMError):
        dm.Dfs(g, 8, True)
    with pytest.raises(dm.DMError):
        dm.Dfs(g, 4, False, dm.dfs_config(vwave=12))
    with pytest.raises(dm.DMError):
        dm.dfs_kernel(7, 32)
