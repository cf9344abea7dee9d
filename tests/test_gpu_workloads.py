"""The benchmark workloads at full size against the oracle's counts
(tests/golden/workload_counts.json, written by make_workload_golden.py),
bit-exact.  Config 4 (about 1.2e12 subgraphs) is checked on root ranges."""
import hashlib
import json
import os

import numpy as np
import pytest

import dumato_amd as dm
from dumato_amd import workloads as W

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "workload_counts.json")))


def check_graph(g, want):
    n, m, md = g.info()
    off, adj = g.csr()
    assert (n, m, md) == (want["n"], want["m"], want["maxdeg"])
    assert hashlib.sha256(off.tobytes() + adj.tobytes()).hexdigest() == want["csr_sha256"]


def hist_of(h):
    return {str(i): int(x) for i, x in enumerate(h) if x}


def test_er_2000():
    g = W.er_2000()
    check_graph(g, GOLD["er_2000"])
    for k, want in GOLD["er_2000"]["clique"].items():
        print("case er_2000 clique k=%s" % k, flush=True)
        assert g.clique_count(int(k)) == want


def test_chunglu_18k_motif4():
    g = W.chunglu_18k()
    gold = GOLD["chunglu_18k"]
    check_graph(g, gold)
    for w in (16, 64):
        print("case chunglu motif k=4 w=%d" % w, flush=True)
        tot, h = g.motif_count(4, dm.config(vwave=w))
        assert tot == gold["motif"]["4"]["total"] and hist_of(h) == gold["motif"]["4"]["hist"]
    # the K4 class (cgLocal 9) of the histogram is the 4-clique count
    assert int(h[9]) == gold["clique"]["4"] == g.clique_count(4)
    for k in ("3", "5"):
        assert g.clique_count(int(k)) == gold["clique"][k]


def test_rmat_s17_cliques():
    g = W.rmat_s17()
    gold = GOLD["rmat_s17"]
    check_graph(g, gold)
    for k, want in gold["clique"].items():
        print("case rmat clique k=%s" % k, flush=True)
        assert g.clique_count(int(k)) == want
    for k in ("4", "5"):
        print("case rmat clique bytes k=%s list path" % k, flush=True)
        assert g.clique_count(int(k), dm.config(clique_bitmap=0)) == gold["clique"][k]
        assert g.last_stats()["alg_bytes"] == gold["clique_bytes"][k]
        assert g.clique_count(int(k)) == gold["clique"][k]
        assert g.last_stats()["alg_bytes"] == gold["clique_bytes"][k]


def test_ba_planted_motif5_root_ranges():
    g = W.ba_100k_planted()
    gold = GOLD["ba_100k_planted"]
    check_graph(g, gold)
    # the ranges cover the hubs (ids 0..10) and every member of one planted
    # clique; the members are pairwise adjacent
    off, adj = g.csr()
    members = gold["planted_members"]
    assert len(members) >= 8 and "0-11" in gold["motif5_root_ranges"]
    for i, a in enumerate(members):
        nb = set(adj[off[a]:off[a + 1]].tolist())
        assert all(b in nb for b in members[i + 1:])
        assert "%d-%d" % (a, a + 1) in gold["motif5_root_ranges"]
    for rng, want in gold["motif5_root_ranges"].items():
        a, b = map(int, rng.split("-"))
        print("case ba motif k=5 roots %s" % rng, flush=True)
        tot, h = g.motif_count(5, dm.config(root_begin=a, root_end=b))
        assert tot == want["total"] and hist_of(h) == want["hist"]
    for k, want in gold["clique"].items():
        print("case ba clique k=%s" % k, flush=True)
        assert g.clique_count(int(k)) == want
