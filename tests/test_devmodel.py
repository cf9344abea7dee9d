"""The restated device algorithm (tests/devmodel.py) against the oracle.

Checks, on the CPU, that union-with-membership-masks leaf counting, range
jobs and donation of the lowest level's tail keep counts bit-exact under any
donation schedule and any vertex labelling.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import graphs as G
from tests.devmodel import CliqueModel, MotifModel


@pytest.mark.parametrize("seed", range(3))
def test_model_matches_oracle(seed):
    g = O.csr_from_edges(*G.gnp(36, 0.22, 40 + seed))
    _, gd = O.relabel(g, oriented=False)
    _, go = O.relabel(g, oriented=True)
    for K in (3, 4, 5, 6):
        lut, ncls = O.load_lut(K)
        ot, oh = O.motif_count(g, K)
        for gg in (g, gd):
            for dp in (0.0, 0.35):
                mt, mh = MotifModel(gg, K, lut, ncls).run(dp, seed)
                assert mt == ot and np.array_equal(mh, oh.astype(np.int64))
        oc = O.clique_count(g, K)
        for dp in (0.0, 0.35):
            assert CliqueModel(go, K).run(dp, seed) == oc


def test_model_citeseer_k4_with_donation():
    c = O.load_edgelist("tests/golden/datasets/citeseer.edgelist")
    _, cd = O.relabel(c, False)
    lut, ncls = O.load_lut(4)
    mt, mh = MotifModel(cd, 4, lut, ncls).run(0.05, 1)
    ot, oh = O.motif_count(c, 4)
    assert mt == ot and np.array_equal(mh, oh.astype(np.int64))
