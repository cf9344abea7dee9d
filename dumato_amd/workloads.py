"""The five benchmark workloads (seeded synthetic graphs) and their queries.

Seeds are fixed here: BASE_SEED + workload number.
  1  Erdos-Renyi G(n=2000, m=20000)                      clique k=4 (and k=3)
  2  Chung-Lu n=18000, 200000 edges, gamma 2.3, wmax 500  motif k=4
  3  R-MAT scale 17, edge factor 8, (.57,.19,.19,.05)     clique k=5, k=6
  4  Barabasi-Albert n=100000, m=10, + 200 planted cliques of size U[8,12]
                                                          motif k=5
  5  256 graphs, n log-uniform [500, 50000], average degree U[4, 40], half
     Erdos-Renyi, half Chung-Lu (gamma U[2.1, 2.8], wmax = min(n/8, 4*avg
     degree)); per graph clique k=3..6 and motif k=3..5
"""
import math

import numpy as np

from . import Graph

BASE_SEED = 20261202


def er_2000():
    return Graph.erdos_renyi(2000, 20000, BASE_SEED + 1)


def chunglu_18k():
    return Graph.chung_lu(18000, 200000, 2.3, 500.0, BASE_SEED + 2)


def rmat_s17():
    return Graph.rmat(17, 8, 0.57, 0.19, 0.19, BASE_SEED + 3)


def ba_100k_planted():
    return Graph.barabasi_albert(100000, 10, 200, 8, 12, BASE_SEED + 4)


def batch_specs(count=256, seed=BASE_SEED + 5):
    """Parameters of the mixed batch: list of dicts (kind, n, m, gamma, wmax, seed)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        n = int(round(math.exp(rng.uniform(math.log(500), math.log(50000)))))
        avg = float(rng.uniform(4, 40))
        m = int(round(n * avg / 2))
        if i % 2 == 0:
            specs.append(dict(kind="er", n=n, m=m, seed=int(seed * 1000 + i)))
        else:
            gamma = float(rng.uniform(2.1, 2.8))
            wmax = float(min(n / 8.0, 4.0 * avg))
            specs.append(dict(kind="chunglu", n=n, m=m, gamma=gamma, wmax=wmax, seed=int(seed * 1000 + i)))
    return specs


def make_batch_graph(spec):
    if spec["kind"] == "er":
        return Graph.erdos_renyi(spec["n"], spec["m"], spec["seed"])
    return Graph.chung_lu(spec["n"], spec["m"], spec["gamma"], spec["wmax"], spec["seed"])


BATCH_QUERIES = [("clique", 3), ("clique", 4), ("clique", 5), ("clique", 6),
                 ("motif", 3), ("motif", 4), ("motif", 5)]
