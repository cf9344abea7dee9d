"""Drop-in command-line tools on the GPU: the stdout lines of clique_PA
(src/main/clique_PA.cu:34-130) and motifs (src/main/motifs_DM_WCLB.cu:38-140),
and their DM_VALIDATE hook against Peregrine-format files written from the
oracle's counts."""
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "..", "dumato_amd", "bin")
GOLD = os.path.join(HERE, "golden", "datasets")

pytestmark = pytest.mark.gpu


def _run(tool, gp, k, validate, host_lb=False):
    env = dict(os.environ, DM_VALIDATE="1" if validate else "0")
    if host_lb:
        env.update(DM_LB="host", DM_LB_POLL_US="200")
    print("case cli %s %s k=%d validate=%d host_lb=%d" % (tool, os.path.basename(gp), k, validate, host_lb),
          flush=True)
    r = subprocess.run([os.path.join(BIN, tool), gp, str(k), "102400", "256", "16", "30"],
                       capture_output=True, text=True, env=env, timeout=100)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("k", [4, 5, 6])
def test_q_pa_tools(citeseer, k):
    # q4_PA .. q13_PA drop-ins: the near-clique query of size k
    want = O.near_clique_count(O.load_edgelist(citeseer), k)
    for tool in ("q%d_PA" % k, "q13_PA"):  # the size comes from argv, as in the reference
        out = _run(tool, citeseer, k, False)
        assert re.search(r"^Result: %d\.$" % want, out, re.M), out
        assert re.search(r"^Time \(LB\): [0-9.]+ secs,[0-9.]+%$", out, re.M), out


@pytest.mark.parametrize("k", [4, 5])
def test_cli_host_driven_balancing(citeseer, k):
    # DM_LB=host: the reference's model (step-API kernel, argv
    # donationsPerWarp / threshold, host stop/snapshot/donate loop)
    ref = O.load_edgelist(citeseer)
    out = _run("clique_PA", citeseer, k, False, host_lb=True)
    assert re.search(r"^Result: %d\.$" % O.clique_count(ref, k), out, re.M), out
    assert re.search(r"^Time \(LB\): [0-9.]+ secs,[0-9.]+%$", out, re.M), out
    tot, hist = O.motif_count(ref, k)
    out = _run("motifs", citeseer, k, False, host_lb=True)
    lines = dict((int(a), int(b)) for a, b in re.findall(r"^(\d+), (\d+)$", out, re.M))
    assert lines == {i: int(v) for i, v in enumerate(hist) if v}, out
    assert re.search(r"^Aggregated: %d, %d patterns\.$" % (tot, np.count_nonzero(hist)), out, re.M), out


@pytest.fixture
def citeseer(tmp_path):
    dst = tmp_path / "citeseer.edgelist"
    dst.write_bytes(open(os.path.join(GOLD, "citeseer.edgelist"), "rb").read())
    return str(dst)


@pytest.mark.parametrize("k", [3, 4])
def test_clique_cli_result_and_validation(citeseer, k):
    want = O.clique_count(O.load_edgelist(citeseer), k)
    with open(citeseer + ".%d.aggregateCounter" % k, "w") as f:
        f.write("[1-2][1-3][2-3]: %d\n" % want)
    out = _run("clique_PA", citeseer, k, True)
    assert re.search(r"^Result: %d\.$" % want, out, re.M), out
    assert "validateAggregateCounter: MATCH" in out
    assert re.search(r"^Time \(GPU\): [0-9.]+ secs\.$", out, re.M)
    # validation stays off unless asked for, as in the reference
    assert "validate" not in _run("clique_PA", citeseer, k, False)


@pytest.mark.parametrize("k", [3, 4])
def test_motif_cli_histogram_and_validation(citeseer, k):
    tot, hist = O.motif_count(O.load_edgelist(citeseer), k)
    with open(citeseer + ".%d.aggregatePattern" % k, "w") as f:
        for i, v in enumerate(reversed([int(x) for x in hist if x > 0])):
            f.write("[%d-%d]: %d\n" % (i, i + 1, v))
    out = _run("motifs", citeseer, k, True)
    lines = dict((int(a), int(b)) for a, b in re.findall(r"^(\d+), (\d+)$", out, re.M))
    got = np.zeros(len(hist), np.uint64)
    for i, v in lines.items():
        got[i] = v
    assert np.array_equal(got, hist), out
    npat = int(np.count_nonzero(hist))
    assert re.search(r"^Aggregated: %d, %d patterns\.$" % (tot, npat), out, re.M), out
    assert "validateAggregatePattern (counting): MATCH" in out
    assert "validateAggregatePattern (per pattern): MATCH" in out
