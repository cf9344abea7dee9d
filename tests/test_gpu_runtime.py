"""Runtime mechanics on the GPU: the device timeline in dm_stats, the
donation-check period, the worker stacks spilling to the global arena, the
static-partition cache, and heavy donation traffic.  Every case must keep
the counts exact (against the oracle or the committed full-size goldens)."""
import json
import os

import numpy as np
import pytest

import dumato_amd as dm
from dumato_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def cfg(**kw):
    kw.setdefault("watchdog_s", 60.0)
    return dm.config(**kw)


@pytest.fixture(scope="module")
def citeseer():
    path = os.path.join(GOLD, "datasets", "citeseer.edgelist")
    return dm.Graph.from_edgelist(path), O.load_edgelist(path)


@pytest.fixture(scope="module")
def rmat():
    return W.rmat_s17()


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLD, "workload_counts.json")) as f:
        return json.load(f)


def test_timeline_is_ordered(rmat):
    print("case timeline rmat k=4", flush=True)
    rmat.clique_count(4, cfg())
    st = rmat.last_stats()
    assert 0 < st["first_job_ms"] <= st["last_job_ms"] <= st["exit_ms"]
    assert st["static_done_ms"] <= st["exit_ms"]
    # the exit of the last block is inside the events bracketing the launch
    assert st["exit_ms"] <= st["device_ms"] * 1.05 + 0.05
    assert st["jobs"] >= 1 and st["busy_max_cycles"] > 0


def test_every_launched_worker_runs(rmat, golden):
    """A default launch holds only as many blocks as fit on the CUs together.
    gfx950 allocates a block's LDS in 1,280-byte units; a grid sized by the
    LDS request alone holds blocks that become resident only when others exit,
    and whose workers then find nothing left to run.  With more roots than
    workers, every worker of the launch runs jobs: cliques (3 blocks per CU),
    motifs k = 4 (4 per CU) and k >= 5 (the motif kernel at 6 waves)."""
    for k in (4, 5, 6):
        print("case all workers run: rmat clique k=%d" % k, flush=True)
        assert rmat.clique_count(k, cfg()) == golden["rmat_s17"]["clique"][str(k)]
        st = rmat.last_stats()
        assert st["active_workers"] == st["workers"], (k, st["active_workers"], st["workers"])
    g = dm.Graph.erdos_renyi(20000, 60000, 777)
    off, adj = g.csr()
    src = np.repeat(np.arange(len(off) - 1), np.diff(off)).astype(np.int32)
    keep = src < adj
    ref = O.csr_from_edges(src[keep], adj[keep], len(off) - 1)
    for k in (4, 5):
        print("case all workers run: ER 20000/60000 motif k=%d" % k, flush=True)
        tot, hist = g.motif_count(k, cfg())
        st = g.last_stats()
        wt, wh = O.motif_count(ref, k)
        assert tot == wt and np.array_equal(hist, wh)
        assert st["workers"] >= (6144 if k >= 5 else 4096)
        assert st["active_workers"] == st["workers"], (k, st["active_workers"], st["workers"])


@pytest.mark.parametrize("period", [0, 2, 50])
def test_donation_check_period_keeps_counts(citeseer, period):
    g, ref = citeseer
    print("case steal_period=%d" % period, flush=True)
    assert g.clique_count(4, cfg(steal_period=period, steal_min=2)) == O.clique_count(ref, 4)
    tot, hist = g.motif_count(4, cfg(steal_period=period, steal_min=2))
    wt, wh = O.motif_count(ref, 4)
    assert tot == wt and np.array_equal(hist, wh)


@pytest.mark.parametrize("lds", [256, 1024])
def test_small_lds_stacks_spill_to_the_arena(rmat, golden, lds):
    # tiny LDS slices put the level lists, leaf scratch and root bitmaps in
    # the global arena; counts must not change
    want = golden["rmat_s17"]["clique"]
    for k in (4, 5):
        print("case lds=%d clique k=%d" % (lds, k), flush=True)
        assert rmat.clique_count(k, cfg(lds_bytes=lds, vwave=64, steal_min=2)) == want[str(k)]


def hist_of(h):
    return {str(i): int(c) for i, c in enumerate(h) if c}


@pytest.mark.parametrize("lds", [256, 2048])
def test_small_lds_motif_leaf_paths(golden, lds):
    # little or no LDS for the motif leaf phase: its scratch and the level
    # lists go to the arena, the membership filter of U shrinks or is left
    # out (every lookup then goes through the search queue); histograms must
    # not change
    cl = W.chunglu_18k()
    want = golden["chunglu_18k"]["motif"]["4"]
    for w in (16, 64):
        print("case lds=%d motif k=4 chunglu vwave=%d" % (lds, w), flush=True)
        tot, h = cl.motif_count(4, cfg(lds_bytes=lds, vwave=w, steal_min=2))
        assert tot == want["total"] and hist_of(h) == want["hist"]
    ba = W.ba_100k_planted()
    for rng, want in golden["ba_100k_planted"]["motif5_root_ranges"].items():
        a, b = map(int, rng.split("-"))
        print("case lds=%d motif k=5 ba roots %s" % (lds, rng), flush=True)
        tot, h = ba.motif_count(5, cfg(lds_bytes=lds, vwave=64, root_begin=a, root_end=b))
        assert tot == want["total"] and hist_of(h) == want["hist"]


def test_heavy_donation_traffic_keeps_counts(rmat, golden):
    # donate as often and as finely as the runtime allows
    want = golden["rmat_s17"]["clique"]
    print("case heavy donation k=6", flush=True)
    assert rmat.clique_count(6, cfg(steal_min=2, steal_period=0, vwave=64)) == want["6"]
    assert rmat.last_stats()["donations"] > 0


def test_partition_cache_follows_root_range_and_model(citeseer):
    g, ref = citeseer
    n = g.info()[0]
    lo, hi = n // 4, n // 2
    print("case partition cache", flush=True)
    part = g.clique_count(4, cfg(root_begin=lo, root_end=hi))
    full = g.clique_count(4, cfg())
    assert full == O.clique_count(ref, 4)
    assert g.clique_count(4, cfg(root_begin=lo, root_end=hi)) == part
    # learned partitions are rebuilt after each training
    g.sched_train(False, 3, epochs=5, seed=1)
    a = g.clique_count(4, cfg(sched=dm.DM_SCHED_LEARNED))
    g.sched_train(False, 3, epochs=5, seed=2)
    b = g.clique_count(4, cfg(sched=dm.DM_SCHED_LEARNED))
    assert a == b == full


def test_cached_partitions_alternate_across_graphs(citeseer):
    """A repeated query reuses its partition and, when no other partition has
    replaced them since, the queues already on the device.  Alternating graphs,
    families, sizes, root ranges and schedulers on one device must upload
    whatever the next launch needs: every count stays exact."""
    g1, ref1 = citeseer
    g2 = dm.Graph.erdos_renyi(3000, 15000, 4242)
    off, adj = g2.csr()
    src = np.repeat(np.arange(len(off) - 1), np.diff(off)).astype(np.int32)
    keep = src < adj
    ref2 = O.csr_from_edges(src[keep], adj[keep], len(off) - 1)
    want = {}
    for name, ref in (("g1", ref1), ("g2", ref2)):
        for k in (3, 4):
            want[(name, "c", k)] = O.clique_count(ref, k)
            want[(name, "m", k)] = O.motif_count(ref, k)
    n1 = g1.info()[0]
    part_cfg = cfg(root_begin=n1 // 3, root_end=n1)
    part_want = g1.clique_count(4, part_cfg)
    g2.sched_train(False, 3, epochs=5, seed=3)
    for it in range(3):
        for name, g in (("g1", g1), ("g2", g2)):
            for k in (3, 4):
                print("case alternate it=%d %s k=%d" % (it, name, k), flush=True)
                assert g.clique_count(k, cfg()) == want[(name, "c", k)]
                tot, hist = g.motif_count(k, cfg())
                wt, wh = want[(name, "m", k)]
                assert tot == wt and np.array_equal(hist, wh)
        assert g1.clique_count(4, part_cfg) == part_want
        assert g2.clique_count(4, cfg(sched=dm.DM_SCHED_LEARNED)) == want[("g2", "c", 4)]
        # a recomputed partition of the same key replaces the one on the device
        assert g2.clique_count(4, cfg(repartition=1)) == want[("g2", "c", 4)]
