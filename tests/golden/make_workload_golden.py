"""Oracle counts for the seeded benchmark workloads -> workload_counts.json.

Run in the build container (takes a few minutes on 8 cores).  The GPU tests
compare the HIP kernels with these values at full size.  The graphs come from
the library's seeded generators; the SHA-256 of each CSR is stored so a
generator change is caught.  Config 4 (about 3.5e12 subgraphs) is covered by
root ranges: the oracle enumerates only the subgraphs whose smallest vertex in
degree-rank order has an original id in the range, exactly what the library
does with cfg.root_begin / root_end.  The ranges are three fixed spans, the
hubs (BA ids 0..10, the initial clique) and each member of one planted clique
(found from the graph: see planted_member_set).
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from dumato_amd import workloads as W  # noqa: E402
from oracle import pyoracle as O  # noqa: E402


def ref_of(g):
    n, m, md = g.info()
    off, adj = g.csr()
    h = hashlib.sha256(off.tobytes() + adj.tobytes()).hexdigest()
    return O.CSR(n, off, adj), {"n": n, "m": m, "maxdeg": md, "csr_sha256": h}


def hist_dict(h):
    return {str(i): int(x) for i, x in enumerate(h) if x}


def planted_member_set(ref, late=1000):
    """Original ids of one planted clique of the BA graph: the largest clique
    through the lowest id >= `late` that roots an 8-clique.  Growth attaches
    each vertex by m = 10 edges to earlier, mostly early, vertices: only the
    first vertices (the initial K_11 and those that joined it) form 8-cliques
    by growth, so an 8-clique through a late vertex is a planted one."""
    perm, ro = O.relabel(ref, oriented=True)
    _, per_root = O.clique_count(ro, 8, per_root=True)
    inv = np.argsort(perm)
    roots = sorted(int(inv[r]) for r in np.nonzero(per_root)[0] if inv[r] >= late)
    v = roots[0]
    nb = set(ref.neighbours(v).tolist())
    best = []

    def expand(clique, cand):
        nonlocal best
        if len(clique) > len(best):
            best = list(clique)
        for u in sorted(cand):
            if len(clique) + len(cand) <= len(best):
                return
            expand(clique + [u], cand & set(ref.neighbours(u).tolist()))
            cand = cand - {u}

    expand([v], nb)
    return sorted(best)


def main():
    out = {}
    t0 = time.time()
    r1, info = ref_of(W.er_2000())
    out["er_2000"] = dict(info, clique={str(k): O.clique_count(r1, k) for k in (3, 4)})
    print("er_2000", out["er_2000"]["clique"], flush=True)

    r2, info = ref_of(W.chunglu_18k())
    tot, h = O.motif_count(r2, 4)
    out["chunglu_18k"] = dict(info, motif={"4": {"total": tot, "hist": hist_dict(h)}},
                              clique={str(k): O.clique_count(r2, k) for k in (3, 4, 5)})
    print("chunglu_18k", tot, "%.0fs" % (time.time() - t0), flush=True)

    r3, info = ref_of(W.rmat_s17())
    _, ro = O.relabel(r3, oriented=True)
    out["rmat_s17"] = dict(info, clique={str(k): O.clique_count(ro, k) for k in (3, 4, 5, 6)},
                           clique_bytes={str(k): O.clique_bytes(ro, k) for k in (4, 5)},
                           near_clique={"5": O.near_clique_count(r3, 5)})
    print("rmat_s17", out["rmat_s17"]["clique"], "%.0fs" % (time.time() - t0), flush=True)

    r4, info = ref_of(W.ba_100k_planted())
    perm, rd = O.relabel(r4, oriented=False)
    members = planted_member_set(r4)
    ranges = {}
    spans = [(1000, 1005), (50000, 50020), (99990, 100000), (0, 11)] + [(v, v + 1) for v in members]
    for a, b in spans:
        tot, h, _ = O.motif_count_roots(rd, 5, perm[a:b].astype(np.int32))
        ranges["%d-%d" % (a, b)] = {"total": tot, "hist": hist_dict(h)}
    out["ba_100k_planted"] = dict(info, motif5_root_ranges=ranges, planted_members=members,
                                  clique={str(k): O.clique_count(r4, k) for k in (5, 8)})
    print("ba_100k_planted", "%.0fs" % (time.time() - t0), flush=True)
    with open(os.path.join(HERE, "workload_counts.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
