"""Host-side parts of libdumato_amd.so that run without a GPU: the C ABI
exports, pattern-table regeneration, graph input / CSR, generators and error
behaviour.  Compared with the oracle and the reference's own dictionaries."""
import os
import re

import numpy as np
import pytest

import dumato_amd as dm
from oracle import pyoracle as O
from tests import graphs as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "dumato.h")).read()
    return sorted(set(re.findall(r"\b(dm_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    L = dm.lib()
    declared = header_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    bound = {s[0] for s in dm.SYMBOLS}
    assert set(declared) == bound


def test_config_defaults():
    c = dm.config()
    assert c.sched == dm.DM_SCHED_PLAIN and c.steal == 1 and c.steal_min == 0
    assert c.vwave == 0 and c.watchdog_s > 0
    assert c.steal_period == 2  # microseconds between donation checks


@pytest.mark.parametrize("bad", [dict(steal_min=1), dict(steal_min=-3), dict(steal_period=-1)])
def test_config_validation_before_any_device_work(bad):
    # rejected on the host, before the device is touched
    g = dm.Graph.erdos_renyi(50, 100, 1)
    with pytest.raises(dm.DMError) as e:
        g.clique_count(3, dm.config(**bad))
    assert "steal_min" in str(e.value)


@pytest.mark.parametrize("k", [3, 4, 5, 6])
def test_pattern_tables_match_reference_dictionaries(k):
    q2c, codes = dm.pattern_table(k)
    d = O.load_dict(k)
    assert q2c.tolist() == d["quick_to_cg_local"]
    assert codes.tolist() == d["cg_local_to_global"]
    assert dm.num_classes(k) == len(d["cg_local_to_global"])


def test_pattern_table_k7_is_isomorphism_classes():
    # no reference file for k=7: check class count (graphs on 7 vertices with
    # at least one edge) and that relabelled quick codes share a class
    q2c, codes = dm.pattern_table(7)
    assert dm.num_classes(7) == 1043 and len(q2c) == 1 << 20
    assert len(set(q2c.tolist())) == 1043
    rng = np.random.default_rng(0)

    def quick_of(adj, order):
        q = 0
        for s in range(2, 7):
            for t in range(s):
                if adj[order[s]][order[t]]:
                    q |= 1 << ((s - 2) * (s + 1) // 2 + t)
        return q

    for _ in range(200):
        A = np.triu(rng.random((7, 7)) < 0.45, 1)
        A = A | A.T
        A[0, 1] = A[1, 0] = True
        base = quick_of(A, list(range(7)))
        for _ in range(5):
            perm = list(rng.permutation(7))
            if not A[perm[0], perm[1]]:
                continue
            assert q2c[quick_of(A, perm)] == q2c[base]


def test_unsupported_pattern_sizes():
    with pytest.raises(dm.DMError) as e:
        dm.num_classes(8)
    assert e.value.code == -4
    with pytest.raises(dm.DMError):
        dm.num_classes(2)


@pytest.mark.parametrize("name", ["test.edgelist", "citeseer.edgelist"])
def test_reader_matches_oracle(name):
    path = os.path.join(GOLD, "datasets", name)
    g = dm.Graph.from_edgelist(path)
    ref = O.load_edgelist(path)
    off, adj = g.csr()
    assert g.n == ref.n and np.array_equal(off, ref.off) and np.array_equal(adj, ref.adj)
    assert g.info()[2] == int(ref.degrees().max())


def test_reader_errors(tmp_path):
    with pytest.raises(dm.DMError) as e:
        dm.Graph.from_edgelist(str(tmp_path / "missing.edgelist"))
    assert e.value.code == -2
    for bad in ("0 1\n2\n", "0 1\n1 x\n", "0 -1\n"):
        p = tmp_path / "bad.edgelist"
        p.write_text(bad)
        with pytest.raises(dm.DMError):
            dm.Graph.from_edgelist(str(p))


def test_from_edges_dedup_loops_isolated():
    g = dm.Graph.from_edges([0, 1, 0, 3, 2], [1, 0, 1, 3, 4], n=7)
    off, adj = g.csr()
    assert g.n == 7 and g.m == 2
    assert adj.tolist() == [1, 0, 4, 2]
    assert np.diff(off).tolist() == [1, 1, 1, 0, 1, 0, 0]


def test_generators_deterministic_and_sized():
    a = dm.Graph.erdos_renyi(2000, 20000, 7)
    b = dm.Graph.erdos_renyi(2000, 20000, 7)
    c = dm.Graph.erdos_renyi(2000, 20000, 8)
    assert a.m == 20000 and a.n == 2000
    assert np.array_equal(a.csr()[1], b.csr()[1]) and not np.array_equal(a.csr()[1], c.csr()[1])
    cl = dm.Graph.chung_lu(3000, 20000, 2.3, 150.0, 3)
    n, m, md = cl.info()
    assert n == 3000 and m == 20000 and 60 <= md <= 400
    r = dm.Graph.rmat(12, 8, 0.57, 0.19, 0.19, 5)
    n, m, md = r.info()
    assert n == 4096 and 20000 < m <= 8 * 4096
    ba = dm.Graph.barabasi_albert(3000, 5, 10, 8, 12, 9)
    n, m, md = ba.info()
    assert n == 3000 and m >= 5 * (3000 - 6)
    # planted cliques are present: at least 10 * C(8, 4) 4-cliques beyond BA's own
    ref = O.CSR(n, *ba.csr())
    assert O.clique_count(ref, 8) >= 10


def test_noalloc_limits_checked_on_the_host():
    # rejected before any device work: sizes beyond the static array and
    # options that would need allocation
    s, d, n = G.star(1 << 20)
    hub = dm.Graph.from_edges(s, d, n)
    with pytest.raises(dm.DMError) as e:
        hub.count_noalloc(True, 6)
    assert e.value.code == -1 and "static scratch" in str(e.value)
    g = dm.Graph.from_edges(*G.complete(5)[:2])
    for bad in (dict(sched=dm.DM_SCHED_LEARNED), dict(record_costs=1), dict(root_end=3)):
        with pytest.raises(dm.DMError) as e:
            g.count_noalloc(False, 3, dm.config(**bad))
        assert e.value.code == -1


def test_counting_without_gpu_fails_loudly():
    import ctypes
    L = dm.lib()
    n = ctypes.c_int()
    # only meaningful where no GPU is visible
    if os.environ.get("HIP_VISIBLE_DEVICES", None) == "" or not os.path.exists("/dev/kfd"):
        g = dm.Graph.from_edges(*G.complete(5)[:2])
        with pytest.raises(dm.DMError) as e:
            g.clique_count(3)
        assert e.value.code == -3
    del n, L
