#!/usr/bin/env python3
"""Benchmark: subgraphs enumerated per second on one MI355X.

Headline workload (value): k=4 motif histogram on the seeded Chung-Lu graph
(n=18,000, 200,000 edges, gamma 2.3, wmax 500).  Also reported in the same
JSON line: k=5 clique count on the seeded R-MAT scale-17 graph, and the
learned-scheduler speedup where a model is trained.

A step is one full counting call (schedule + kernel + result copy) over the
whole graph, which is resident on the GPU before timing starts.  value =
subgraphs counted per second of wall time over the timed steps.  The roofline
uses the enumeration kernel's own duration (HIP events on the library's
stream) and its algorithmic bytes (counted on the device with the formula of
the measurement plan; checked against the CPU oracle in tests).

The CPU baseline is the oracle's port of the reference's DFS logic
(clique_DM_DFS / motifs_DM_DFS: thread per root, linear-scan adjacency),
OpenMP over all host cores, timed on a seeded random sample of roots for a
bounded time.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")


def cpu_baseline(graph, mode, k, budget_s, seed=7):
    """Oracle port timed on random roots until budget_s is used."""
    import numpy as np

    from oracle import pyoracle as O
    n, m, _ = graph.info()
    off, adj = graph.csr()
    ref = O.CSR(n, off, adj)
    order = np.random.default_rng(seed).permutation(n).astype(np.int32)
    threads = O.num_threads()
    done = 0
    count = 0
    chunk = 64
    t0 = time.perf_counter()
    while done < n and time.perf_counter() - t0 < budget_s:
        roots = order[done:done + chunk]
        if mode == "motif":
            c, _, _ = O.motif_count_roots(ref, k, roots, nthreads=threads, faithful=True)
        else:
            c, _ = O.clique_count_roots(ref, k, roots, nthreads=threads, faithful=True)
        count += c
        done += len(roots)
        el = time.perf_counter() - t0
        per_root = el / done
        left = budget_s - el
        chunk = int(max(16, min(4096, 0.9 * left / per_root if per_root > 0 else 4096)))
    dt = time.perf_counter() - t0
    return {"value": count / dt if dt > 0 else 0.0, "unit": "subgraphs/s", "cores": threads,
            "kind": "port",
            "sample": "%d of %d roots (seeded random order), %d subgraphs in %.1f s, "
                      "faithful linear-scan DFS of the reference (%s k=%d)" % (done, n, count, dt, mode, k)}


def run_steps(fn, warmup, steps):
    for _ in range(warmup):
        fn()
    dev, walls, out = [], [], None
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        out = fn()
        walls.append(time.perf_counter() - s0)
        dev.append(out[1])
    t = time.perf_counter() - t0
    return out, t, dev, walls


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--vwave", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(root subtrees would shard across GPUs with one final all-reduce of the counts)",
              file=sys.stderr)
        sys.exit(2)

    import dumato_amd as dm
    from dumato_amd import workloads as W

    g = W.chunglu_18k()
    K = 4
    cfg = dm.config(vwave=args.vwave)

    def step():
        tot, hist = g.motif_count(K, cfg)
        st = g.last_stats()
        return tot, st["device_ms"], st

    (total, _, st), elapsed, dev, walls = run_steps(step, args.warmup, args.steps)
    value = args.steps * total / elapsed
    kern_s = statistics.mean(dev) / 1e3
    achieved = st["alg_bytes"] / kern_s / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        tr = json.load(open(PROFILE_TRAFFIC))
        traffic = tr.get("motif4_chunglu18k")
    line = {
        "metric": "subgraphs enumerated/sec (k=4 motif on Chung-Lu 18K)",
        "value": value,
        "unit": "subgraphs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (seeded Chung-Lu n=18000 m=200000 gamma=2.3 wmax=500)",
        "config": {"workload": "motif k=4, Chung-Lu n=18000 m=200000 gamma=2.3 wmax=500, seed %d" % (W.BASE_SEED + 2),
                   "subgraphs": total, "vwave": st["vwave"], "workers": st["workers"],
                   "kernel_ms": statistics.mean(dev), "kernel_ms_median": statistics.median(dev),
                   "subgraphs_per_kernel_s": total / kern_s, "scheduler": "plain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes": st["alg_bytes"]},
    }
    if not args.no_extra:
        gr = W.rmat_s17()
        ccfg = dm.config(vwave=args.vwave)

        def cstep():
            c = gr.clique_count(5, ccfg)
            s = gr.last_stats()
            return c, s["device_ms"], s

        (c5, _, cst), cel, cdev, _ = run_steps(cstep, 1, max(2, args.steps // 2))
        line["clique5_rmat17"] = {"subgraphs": c5, "value": len(cdev) * c5 / cel, "unit": "subgraphs/s",
                                  "kernel_ms": statistics.mean(cdev),
                                  "subgraphs_per_kernel_s": c5 / (statistics.mean(cdev) / 1e3),
                                  "alg_GBs": cst["alg_bytes"] / (statistics.mean(cdev) / 1e3) / 1e9,
                                  "vwave": cst["vwave"]}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(g, "motif", K, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
