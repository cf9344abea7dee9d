#!/usr/bin/env python3
"""Benchmark: subgraphs enumerated per second on one MI355X.

Headline (value): k=4 motif histogram of the seeded Chung-Lu graph
(n=18,000, 200,000 edges, gamma 2.3, wmax 500), plain scheduler.

Same JSON line, extra fields:
  clique_rmat17   k=5 and k=6 clique counts of the seeded R-MAT scale-17
                  graph, plain vs learned scheduler (the cost model is trained
                  on per-root cycle counts of k=4 runs of the same graph; its
                  Spearman/MAPE against measured per-root cycles at k=4..6 and
                  the makespan reduction are reported);
  learned_speedup plain / learned time per query, where time = enumeration
                  kernel + host scheduling (inference, ordering, binning);
  motif5_ba       (--full) k=5 motif histogram of the Barabasi-Albert graph
                  with 200 planted cliques, plain vs learned.

A step is one full counting call over the whole graph (schedule, kernel,
result copy); the graph is resident on the GPU before timing starts and the
launch parameters are autotuned once before timing.  value = subgraphs per
second of wall time over the timed steps.  The roofline uses the enumeration
kernel's own duration (HIP events on the library's stream, median over the
steps) and its algorithmic bytes: 4 bytes per adjacency entry the DFS reads
(each member list of every node, plus k-1 adjacency tests per motif leaf;
for cliques the out-lists of the candidates of nodes with 1..k-2 vertices),
summed on the device and checked against the CPU oracle in tests.  The
traffic field is the measured HBM bytes per launch of the same kernel
(rocprofv3 FETCH_SIZE and WRITE_SIZE, calibrated on a copy kernel), and
traffic_GBs that traffic over the same kernel time.  The algorithmic fraction
can exceed 1: the graph and the DFS lists are served from LDS and L2/MALL, and
most leaf-level neighbour tests are answered by an LDS membership filter, so
the kernel moves far fewer HBM bytes than the algorithmic figure counts.

The CPU baseline is the oracle's port of the reference's DFS logic
(clique_DM_DFS / motifs_DM_DFS: thread per root, linear-scan adjacency), with
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


def progress(msg):
    """One flushed stderr line per finished stage (the JSON line stays the only stdout)."""
    print("[bench %.1fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


T0 = time.perf_counter()


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
        progress("cpu baseline: %d roots, %.1f s" % (done, el))
    dt = time.perf_counter() - t0
    return {"value": count / dt if dt > 0 else 0.0, "unit": "subgraphs/s", "cores": threads,
            "kind": "port",
            "sample": "%d of %d roots (seeded random order), %d subgraphs in %.1f s, "
                      "faithful linear-scan DFS of the reference (%s k=%d)" % (done, n, count, dt, mode, k)}


def timed(fn, graph, warmup, steps):
    """Run fn warmup+steps times; returns (last result, wall s, kernel ms list, sched ms list)."""
    for _ in range(warmup):
        fn()
    kms, sms = [], []
    out = None
    t0 = time.perf_counter()
    for _ in range(steps):
        out = fn()
        st = graph.last_stats()
        kms.append(st["device_ms"])
        sms.append(st["sched_ms"])
    return out, time.perf_counter() - t0, kms, sms


def query_block(graph, mode, k, sched, warmup, steps):
    import dumato_amd as dm
    cfg = dm.config(sched=sched)
    fn = (lambda: graph.motif_count(k, cfg)[0]) if mode == "motif" else (lambda: graph.clique_count(k, cfg))
    count, wall, kms, sms = timed(fn, graph, warmup, steps)
    st = graph.last_stats()
    km, sm = statistics.median(kms), statistics.median(sms)
    return {"subgraphs": count, "kernel_ms": km, "sched_ms": sm, "query_ms": km + sm,
            "subgraphs_per_s": count / ((km + sm) / 1e3), "wall_ms_per_step": 1e3 * wall / steps,
            "vwave": st["vwave"], "workers": st["workers"], "donations": st["donations"],
            "balance": st["busy_sum_cycles"] / max(1, st["workers"] * st["busy_max_cycles"]),
            "busy_max_mcycles": st["busy_max_cycles"] / 1e6,
            "alg_GBs": st["alg_bytes"] / (km / 1e3) / 1e9}


def rank_corr(a, b):
    """Spearman rank correlation (no tie correction)."""
    import numpy as np
    ra = np.argsort(np.argsort(a)).astype(np.float64)
    rb = np.argsort(np.argsort(b)).astype(np.float64)
    return float(np.corrcoef(ra, rb)[0, 1])


def model_quality(graph, mode, k):
    """Predicted (cost model) vs measured per-root cycles of a plain run
    without stealing at size k: Spearman of the ranks and MAPE of the cost
    (the model predicts log cycles of its training size, so MAPE is only
    meaningful at that size)."""
    import numpy as np

    import dumato_amd as dm
    pred = np.asarray(graph.sched_predict(), np.float64)
    cfg = dm.config(steal=0, record_costs=1)
    if mode == "motif":
        graph.motif_count(k, cfg)
    else:
        graph.clique_count(k, cfg)
    cost = graph.last_root_costs().astype(np.float64)
    has = cost > 0
    est = np.exp(pred[has])
    heavy = cost >= np.quantile(cost[has], 0.9)
    return {"k": k, "roots": int(has.sum()), "spearman": rank_corr(pred[has], cost[has]),
            "spearman_top10pct": rank_corr(pred[heavy], cost[heavy]),
            "mape": float(np.mean(np.abs(est - cost[has]) / cost[has]))}


def plain_vs_learned(graph, mode, k_list, k_train, warmup, steps, budget_ms=300.0):
    import dumato_amd as dm
    out = {}
    t = time.perf_counter()
    loss = graph.sched_train(mode == "motif", k_train, epochs=30, seed=1)
    out["train"] = {"k_train": k_train, "loss": loss, "seconds": time.perf_counter() - t}
    progress("%s: cost model trained on k=%d (loss %.3f)" % (mode, k_train, loss))
    out["model_vs_measured"] = [model_quality(graph, mode, k) for k in (k_train,) + tuple(k_list)]
    progress("%s: model quality measured" % mode)
    for k in k_list:
        graph.autotune(mode == "motif", k, budget_ms)
        progress("%s k=%d: autotuned" % (mode, k))
        p = query_block(graph, mode, k, dm.DM_SCHED_PLAIN, warmup, steps)
        progress("%s k=%d plain: kernel %.2f ms" % (mode, k, p["kernel_ms"]))
        lq = query_block(graph, mode, k, dm.DM_SCHED_LEARNED, warmup, steps)
        progress("%s k=%d learned: kernel %.2f ms" % (mode, k, lq["kernel_ms"]))
        assert p["subgraphs"] == lq["subgraphs"], "learned and plain counts differ"
        out["k%d" % k] = {"plain": p, "learned": lq,
                          "learned_speedup": p["query_ms"] / lq["query_ms"],
                          "learned_kernel_speedup": p["kernel_ms"] / lq["kernel_ms"],
                          "makespan_reduction": 1.0 - lq["busy_max_mcycles"] / p["busy_max_mcycles"]}
    return out


def batch_block(warmup, steps, check_graphs=8):
    """256 seeded graphs x {clique k=3..6, motif k=3..5} in one hipGraph-replayed launch."""
    import numpy as np

    import dumato_amd as dm
    from dumato_amd import workloads as W
    t = time.perf_counter()
    specs = W.batch_specs(256)
    graphs = [W.make_batch_graph(s) for s in specs]
    queries = [(gi, mode, k) for gi in range(len(graphs)) for (mode, k) in W.BATCH_QUERIES]
    b = dm.Batch(graphs)
    gen_s = time.perf_counter() - t
    progress("batch: %d graphs built in %.1f s" % (len(graphs), gen_s))
    best = b.autotune(queries, budget_ms=1500.0)
    progress("batch: autotuned (vwave %d)" % best.vwave)
    for _ in range(warmup):
        b.run(queries)
    kms, lats = [], []
    res = None
    t0 = time.perf_counter()
    for _ in range(steps):
        res, lat = b.run(queries)
        kms.append(b.last_stats()["device_ms"])
        lats.append(lat)
    wall = time.perf_counter() - t0
    progress("batch: %d replays timed" % steps)
    total = sum(int(np.asarray(r).sum()) for r in res)
    # spot-check: the first graphs' results against the single-graph path
    ok = True
    for gi in range(check_graphs):
        for (mode, k), r in zip(W.BATCH_QUERIES, res[gi * len(W.BATCH_QUERIES):(gi + 1) * len(W.BATCH_QUERIES)]):
            one = graphs[gi].motif_count(k)[1] if mode == "motif" else graphs[gi].clique_count(k)
            ok = ok and np.array_equal(np.asarray(one), np.asarray(r))
    km = statistics.median(kms)
    lat = np.median(np.stack(lats), axis=0)
    st = b.last_stats()
    return {"graphs": len(graphs), "queries": len(queries), "subgraphs": total,
            "kernel_ms": km, "subgraphs_per_s": total / (km / 1e3), "wall_ms_per_step": 1e3 * wall / steps,
            "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
            "vwave": st["vwave"], "workers": st["workers"], "steal_min": best.steal_min,
            "donations": st["donations"], "graph_build_s": gen_s,
            "spot_check_vs_single_graph_path": ok, "hipgraph": True}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=7)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--full", action="store_true", help="also the k=5 motif run on the BA graph (~1 min)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 256-graph batched launch")
    ap.add_argument("--no-autotune", action="store_true",
                    help="headline with the default launch shape (no tuning dispatches; for profiling)")
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
    if not args.no_autotune:
        g.autotune(True, K, 300.0)
    progress("headline graph built%s" % ("" if args.no_autotune else " and autotuned"))
    cfg = dm.config()

    def step():
        return g.motif_count(K, cfg)[0]

    total, elapsed, kms, _ = timed(step, g, args.warmup, args.steps)
    progress("headline: %d steps timed, kernel %.2f ms" % (args.steps, statistics.median(kms)))
    st = g.last_stats()
    value = args.steps * total / elapsed
    kern_ms = statistics.median(kms)
    achieved = st["alg_bytes"] / (kern_ms / 1e3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        rec = json.load(open(PROFILE_TRAFFIC)).get("motif4_chunglu18k")
        traffic = rec["bytes"] if rec else None
    line = {
        "metric": "subgraphs enumerated/sec (k=4 motif on Chung-Lu 18K; k=5 clique and learned speedup below)",
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
        "config": {"workload": "motif k=4, Chung-Lu n=18000 m=200000 gamma=2.3 wmax=500, seed %d"
                               % (W.BASE_SEED + 2),
                   "subgraphs": total, "scheduler": "plain", "vwave": st["vwave"],
                   "workers": st["workers"], "lds_bytes": st["lds_bytes"],
                   "kernel_ms_median": kern_ms, "kernel_ms_min": min(kms),
                   "subgraphs_per_kernel_s": total / (kern_ms / 1e3),
                   "balance": st["busy_sum_cycles"] / max(1, st["workers"] * st["busy_max_cycles"]),
                   "donations": st["donations"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "alg_bytes": st["alg_bytes"],
                     "bytes_per_subgraph": st["alg_bytes"] / max(1, total),
                     "traffic_GBs": (traffic / (kern_ms * 1e6)) if traffic else None},
    }
    if not args.no_extra:
        gr = W.rmat_s17()
        line["clique_rmat17"] = plain_vs_learned(gr, "clique", (5, 6), 4, args.warmup, args.steps)
        line["learned_speedup"] = {"clique_k5_rmat17": line["clique_rmat17"]["k5"]["learned_speedup"],
                                   "clique_k6_rmat17": line["clique_rmat17"]["k6"]["learned_speedup"]}
    if not args.no_batch:
        line["batch256"] = batch_block(min(args.warmup, 1), max(1, args.steps // 3))
    if args.full:
        gb = W.ba_100k_planted()
        line["motif5_ba"] = plain_vs_learned(gb, "motif", (5,), 4, 0, 1, budget_ms=2000.0)
        line.setdefault("learned_speedup", {})["motif_k5_ba"] = line["motif5_ba"]["k5"]["learned_speedup"]
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(g, "motif", K, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
