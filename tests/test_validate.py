"""Validation against Peregrine output files (DuMatoCPU::validateAggregateCounter
/ validateAggregatePattern, src/main/DuMatoCPU.cu:564-661).  Host-only: the
expected counts come from the oracle, the files are written here in Peregrine's
"<pattern>: <count>" line format."""
import os
import random

import numpy as np
import pytest

import dumato_amd as dm
from oracle import pyoracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "datasets")


def _graph_copy(tmp_path, name="citeseer.edgelist"):
    dst = tmp_path / name
    dst.write_bytes(open(os.path.join(GOLD, name), "rb").read())
    return str(dst)


def _write_pattern_file(path, counts, shuffle_seed=0, extra_lines=()):
    """One '[...]: count' line per value, in shuffled order (Peregrine numbers
    its patterns its own way), plus lines the reader must skip."""
    vals = list(counts)
    random.Random(shuffle_seed).shuffle(vals)
    with open(path, "w") as f:
        f.write("Finished reading datagraph\n")
        for i, v in enumerate(vals):
            f.write("[%d-%d](1~2)[1-3]: %d\n" % (i, i + 1, v))
        for line in extra_lines:
            f.write(line + "\n")
        f.write("took 0.1s\n")


def test_counter_match_and_mismatch(tmp_path):
    gp = _graph_copy(tmp_path)
    want = O.clique_count(O.load_edgelist(gp), 4)
    # Peregrine prints one line per pattern; several lines are summed.
    with open(gp + ".4.aggregateCounter", "w") as f:
        f.write("All patterns finished after 0.2s\n[1-2][1-3][1-4][2-3][2-4][3-4]: %d\n" % (want - 5))
        f.write("[1-2][1-3][1-4][2-3][2-4][3-4]: 5\n")
    assert dm.validate_counter(gp, 4, want)
    assert not dm.validate_counter(gp, 4, want + 1)
    assert not dm.validate_counter(gp, 4, want - 1)


def test_counter_values_above_int32(tmp_path):
    gp = str(tmp_path / "g.edgelist")
    big = 8365866732  # k=6 clique count of R-MAT scale 17, > 2^31
    with open(gp + ".6.aggregateCounter", "w") as f:
        f.write("[1-2]: %d\n" % big)
    assert dm.validate_counter(gp, 6, big)
    assert not dm.validate_counter(gp, 6, big & 0xFFFFFFFF)


def test_missing_file_is_an_io_error(tmp_path):
    gp = str(tmp_path / "absent.edgelist")
    with pytest.raises(dm.DMError) as e:
        dm.validate_counter(gp, 4, 0)
    assert e.value.code == -2
    with pytest.raises(dm.DMError):
        dm.validate_pattern(gp, 4, np.zeros(10, np.uint64))


@pytest.mark.parametrize("k", [3, 4])
def test_pattern_multiset_ignores_class_ids(tmp_path, k):
    gp = _graph_copy(tmp_path)
    _, hist = O.motif_count(O.load_edgelist(gp), k)
    pos = [int(v) for v in hist if v > 0]
    # zero-count patterns in the file are dropped, as are lines without '['
    _write_pattern_file(gp + ".%d.aggregatePattern" % k, pos + [0, 0], shuffle_seed=k,
                        extra_lines=("pattern count: 99",))
    assert dm.validate_pattern(gp, k, hist) == (True, True)


def test_pattern_detects_moved_and_missing_counts(tmp_path):
    gp = _graph_copy(tmp_path)
    _, hist = O.motif_count(O.load_edgelist(gp), 4)
    pos = [int(v) for v in hist if v > 0]
    assert len(pos) >= 2
    # one subgraph moved between classes: totals agree, per-pattern does not
    moved = list(pos)
    moved[0] += 1
    moved[1] -= 1
    _write_pattern_file(gp + ".4.aggregatePattern", moved)
    assert dm.validate_pattern(gp, 4, hist) == (True, False)
    # one class missing from the file: both differ
    _write_pattern_file(gp + ".4.aggregatePattern", pos[1:])
    assert dm.validate_pattern(gp, 4, hist) == (False, False)
    # our zero slots do not count as patterns
    _write_pattern_file(gp + ".4.aggregatePattern", pos)
    padded = np.concatenate([hist, np.zeros(3, np.uint64)])
    assert dm.validate_pattern(gp, 4, padded) == (True, True)
