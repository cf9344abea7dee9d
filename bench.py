#!/usr/bin/env python3
"""Benchmark: subgraphs enumerated per second on one MI355X.

Headline (value): k=4 motif histogram of the seeded Chung-Lu graph
(n=18,000, 200,000 edges, gamma 2.3, wmax 500), plain scheduler, autotuned
launch shape.  A step is one full counting call over the whole graph
(schedule, kernel, result copy); the graph is resident on the GPU before
timing starts.  value = subgraphs per second of wall time over the timed steps.

Same JSON line, extra blocks (--no-extra skips the first three):
  er2000          k=3/k=4 cliques of the Erdos-Renyi n=2000 graph;
  clique_rmat17   k=5 and k=6 cliques of the seeded R-MAT scale-17 graph;
  motif5_ba       k=5 motifs of the Barabasi-Albert n=100000 graph with 200
                  planted cliques.
The last two compare the plain runtime at its default launch shape (full
64-lane waves, default LDS stack, default steal granularity: 2 for motifs,
32 for cliques), the plain runtime autotuned, and the learned scheduler
autotuned and at the default shape.  The
cost model is trained on per-root cycle counts of k=4 runs of the same graph;
for k=6 the learned query is also run with the model retrained on k=5 costs
(reported beside it, *_train_k_minus_1).  Query time = kernel + host scheduling, and every
timed query computes its partition afresh (learned: feature standardisation,
one MLP forward launch, splitting, binning), so inference and binning are
inside it.  learned_speedup = plain (default shape) query time / learned
(autotuned) query time; learned_kernel_speedup is the same ratio of kernel
times (a cached partition's gain on repeated queries).
  batch256        (--no-batch skips it) 256 seeded graphs x {clique k=3..6,
                  motif k=3..5} in one hipGraph-replayed launch.

roofline: the resource that bounds the headline kernel.  Its shader counters
(profiles/counters.json, from the rocprofv3 runs summarised in
profiles/enum_kernel_motif4_chunglu18k.md) show VALU issue as the most used
unit (latency-bound otherwise: waves wait about two thirds of their cycles),
so achieved = VALU wave-instructions per launch (profile) / kernel time (HIP
events on the library's stream, this run; frac_enumeration leaves out the idle
workers' steal loop, 9% of the count), against the MI355X issue peak of
256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction.  Beside it:
traffic = HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated
on a copy of known size) and the algorithmic bytes of the enumeration
(4 bytes per adjacency entry the DFS reads, plus k-1 adjacency tests per motif
leaf; summed on the device and checked against the CPU oracle in tests), each
as a rate and as a fraction of the 8 TB/s HBM peak.  The algorithmic bytes
are mostly served from L2/MALL and LDS, and the leaf phase never performs the
per-leaf adjacency tests the formula counts, so their rate is not an HBM
figure.

cpu_baseline: the oracle's port of the reference's DFS logic (clique_DM_DFS /
motifs_DM_DFS: thread per root, linear-scan adjacency tests) with OpenMP
(one root per claim, OMP_PROC_BIND=close) over the host cores, heaviest roots
first, timed on the whole headline graph.  cpu_baselines lists every
workload's CPU run (a time budget turns a run into a sample of roots,
labelled so).
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads stay close to their cores
os.environ.setdefault("OMP_PROC_BIND", "close")

HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles, 2.4 GHz
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4
PROFILE_COUNTERS = os.path.join(ROOT, "profiles", "counters.json")


def progress(msg):
    """One flushed stderr line per finished stage (the JSON line stays the only stdout)."""
    print("[bench %.1fs] %s" % (time.perf_counter() - T0, msg), file=sys.stderr, flush=True)


T0 = time.perf_counter()


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_run(graph, mode, k, budget_s, faithful=True, oriented=False, random_order=False, label="",
            chunk_cap=8192, chunk_s=8.0):
    """Oracle port over the graph's roots, heaviest (highest degree) first, in
    chunks, until all are done or budget_s is used.  Heaviest first, each
    root handed to the next free thread, keeps the hubs from meeting in one
    thread's share at the end of a chunk.  Chunks at most double and stay
    below chunk_s seconds of estimated work, so a chunk cannot run long
    unreported.  random_order: a seeded random order instead, for a sampled
    run whose heaviest roots alone would outlast the budget (a sample that
    hits few hubs).  oriented: the FAST variant's input, the degree-oriented
    relabelled CSR (cliques)."""
    import numpy as np

    from oracle import pyoracle as O
    n, m, _ = graph.info()
    off, adj = graph.csr()
    ref = O.CSR(n, off, adj)
    if oriented:
        _, ref = O.relabel(ref, oriented=True)
    if random_order:
        order = np.random.default_rng(7).permutation(n).astype(np.int32)
    else:
        # a root's work grows with its degree (its out-degree in the oriented CSR)
        deg = np.diff(np.asarray(ref.off, np.int64))
        order = np.argsort(-deg, kind="stable").astype(np.int32)
    total_roots = len(order)
    threads = O.num_threads()
    done = count = 0
    chunk = 16
    t0 = time.perf_counter()
    while done < total_roots and time.perf_counter() - t0 < budget_s:
        roots = order[done:done + chunk]
        if mode == "motif":
            c, _, _ = O.motif_count_roots(ref, k, roots, nthreads=threads, faithful=faithful)
        else:
            c, _ = O.clique_count_roots(ref, k, roots, nthreads=threads, faithful=faithful)
        count += c
        done += len(roots)
        el = time.perf_counter() - t0
        per_root = el / done
        est = chunk_s / per_root if per_root > 0 else chunk_cap
        chunk = int(max(16, min(chunk_cap, 2 * chunk, est)))
        progress("cpu %s %s k=%d: %d/%d roots, %.1f s" % (label, mode, k, done, total_roots, el))
    dt = time.perf_counter() - t0
    full = done >= n
    variant = "faithful linear-scan adjacency" if faithful else "binary-search adjacency (FAST)"
    if oriented:
        variant += " on the degree-oriented relabelled CSR"
    if full:
        sample = "all %d roots" % n
    elif random_order:
        sample = "%d of %d roots (seeded random order; time budget %.0f s)" % (done, n, budget_s)
    else:
        sample = "the %d highest-degree roots of %d (time budget %.0f s)" % (done, n, budget_s)
    return {"workload": label, "query": "%s k=%d" % (mode, k), "value": count / dt if dt > 0 else 0.0,
            "unit": "subgraphs/s", "cores": threads, "kind": "port", "variant": variant,
            "subgraphs": count, "seconds": dt, "roots": done, "of_roots": n, "full": full, "sample": sample}


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


def query_block(graph, mode, k, cfg, warmup, steps):
    fn = (lambda: graph.motif_count(k, cfg)[0]) if mode == "motif" else (lambda: graph.clique_count(k, cfg))
    count, wall, kms, sms = timed(fn, graph, warmup, steps)
    st = graph.last_stats()
    km, sm = statistics.median(kms), statistics.median(sms)
    return {"subgraphs": count, "kernel_ms": km, "sched_ms": sm, "query_ms": km + sm,
            "subgraphs_per_s": count / ((km + sm) / 1e3), "wall_ms_per_step": 1e3 * wall / steps,
            "vwave": st["vwave"], "workers": st["workers"], "lds_bytes": st["lds_bytes"],
            "donations": st["donations"],
            "balance": st["busy_sum_cycles"] / max(1, st["workers"] * st["busy_max_cycles"]),
            "busy_max_mcycles": st["busy_max_cycles"] / 1e6, "job_max_mcycles": st["job_max_cycles"] / 1e6,
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


def plain_vs_learned(graph, mode, k_list, warmup, steps, k_train=4, budget_ms=300.0, quality=True):
    """Plain at the default shape, plain autotuned, learned autotuned and
    learned at the default shape, with the cost model trained on per-root
    cycles of size k_train runs (4: the workload's definition).  Where
    k - 1 > k_train the learned query is also run with the model retrained
    on size k-1 (the ranking of a smaller size drifts further: R-MAT's
    densest high-degree roots are heavy at k=6 but rank low at k=4)."""
    import dumato_amd as dm
    out = {}
    motif = mode == "motif"
    for k in k_list:
        # the default shape, fixed before any tuning is cached in the graph
        out["k%d" % k] = {"plain_default": query_block(graph, mode, k, dm.config(vwave=64, repartition=1),
                                                       warmup, steps)}
        progress("%s k=%d plain default shape: kernel %.2f ms" % (mode, k, out["k%d" % k]["plain_default"]["kernel_ms"]))

    def train(kt):
        t = time.perf_counter()
        loss = graph.sched_train(motif, kt, epochs=30, seed=1)
        progress("%s: cost model trained on k=%d (loss %.3f)" % (mode, kt, loss))
        return {"k_train": kt, "loss": loss, "seconds": time.perf_counter() - t}

    out["train"] = [train(k_train)]
    if quality:
        out["model_vs_measured"] = [model_quality(graph, mode, kk) for kk in (k_train,) + tuple(k_list)]
        progress("%s: model quality measured" % mode)
    learned = dm.config(sched=dm.DM_SCHED_LEARNED, repartition=1)
    for k in k_list:
        rec = out["k%d" % k]
        graph.autotune(motif, k, budget_ms)
        progress("%s k=%d: autotuned" % (mode, k))
        rec["plain_tuned"] = query_block(graph, mode, k, dm.config(repartition=1), warmup, steps)
        progress("%s k=%d plain tuned: kernel %.2f ms" % (mode, k, rec["plain_tuned"]["kernel_ms"]))
        rec["learned_tuned"] = query_block(graph, mode, k, learned, warmup, steps)
        progress("%s k=%d learned tuned: kernel %.2f ms, sched %.2f ms" %
                 (mode, k, rec["learned_tuned"]["kernel_ms"], rec["learned_tuned"]["sched_ms"]))
        # the learned partition at the plain default's own shape
        rec["learned_default"] = query_block(graph, mode, k, dm.config(vwave=64, repartition=1,
                                                                       sched=dm.DM_SCHED_LEARNED), warmup, steps)
        progress("%s k=%d learned default shape: kernel %.2f ms" % (mode, k, rec["learned_default"]["kernel_ms"]))
        pd, pt, lt, ld = rec["plain_default"], rec["plain_tuned"], rec["learned_tuned"], rec["learned_default"]
        assert len({x["subgraphs"] for x in (pd, pt, lt, ld)}) == 1, "learned and plain counts differ"
        # query time = kernel + scheduling (the partition is rebuilt on every call)
        rec["learned_speedup"] = pd["query_ms"] / lt["query_ms"]
        rec["learned_vs_plain_tuned"] = pt["query_ms"] / lt["query_ms"]
        # kernel time only (what a cached partition gives on repeated queries)
        rec["learned_kernel_speedup"] = pd["kernel_ms"] / lt["kernel_ms"]
        rec["learned_kernel_speedup_same_shape"] = pd["kernel_ms"] / ld["kernel_ms"]
        rec["makespan_reduction"] = 1.0 - lt["busy_max_mcycles"] / pd["busy_max_mcycles"]
        # what scheduling could gain at most: the plain runtime's idle share
        rec["plain_balance_ceiling"] = 1.0 / max(pd["balance"], 1e-9)
    for k in k_list:
        if k - 1 <= k_train:
            continue
        rec = out["k%d" % k]
        tr = train(k - 1)
        out["train"].append(tr)
        if quality:
            out["model_vs_measured"].append(model_quality(graph, mode, k))
        lp = query_block(graph, mode, k, learned, warmup, steps)
        assert lp["subgraphs"] == rec["plain_default"]["subgraphs"], "learned and plain counts differ"
        rec["learned_tuned_train_k_minus_1"] = lp
        rec["learned_speedup_train_k_minus_1"] = rec["plain_default"]["query_ms"] / lp["query_ms"]
        rec["learned_kernel_speedup_train_k_minus_1"] = rec["plain_default"]["kernel_ms"] / lp["kernel_ms"]
        progress("%s k=%d learned tuned, model trained on k=%d: kernel %.2f ms" % (mode, k, k - 1, lp["kernel_ms"]))
    return out


def batch_block(warmup, steps):
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
    km = statistics.median(kms)
    lat = np.median(np.stack(lats), axis=0)
    st = b.last_stats()
    return {"graphs": len(graphs), "queries": len(queries), "subgraphs": total,
            "kernel_ms": km, "subgraphs_per_s": total / (km / 1e3), "wall_ms_per_step": 1e3 * wall / steps,
            "latency_ms_p50": float(np.percentile(lat, 50)), "latency_ms_p99": float(np.percentile(lat, 99)),
            "vwave": st["vwave"], "workers": st["workers"], "steal_min": best.steal_min,
            "donations": st["donations"], "graph_build_s": gen_s, "hipgraph": True,
            "parity": "every query pinned against the oracle by tests/test_gpu_zbatch256.py"}


def roofline(st, kern_ms, total):
    rec = {}
    if os.path.exists(PROFILE_COUNTERS):
        rec = json.load(open(PROFILE_COUNTERS)).get("motif4_chunglu18k", {})
    valu = rec.get("valu_insts")
    steal = rec.get("valu_insts_steal_loop", 0.0)
    traffic = rec.get("hbm_bytes")
    t = kern_ms / 1e3
    achieved = valu / t / 1e9 if valu else None
    alg = st["alg_bytes"]
    # frac counts every VALU instruction issued, the idle workers' steal
    # loop included; frac_enumeration leaves that loop's share out
    return {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_GINST, "unit": "Ginst/s",
            "frac": achieved / VALU_PEAK_GINST if achieved else None, "traffic": traffic,
            "frac_enumeration": (valu - steal) / t / 1e9 / VALU_PEAK_GINST if valu else None,
            "valu_insts_per_launch": valu, "valu_insts_steal_loop": steal,
            "counters": "profiles/enum_kernel_motif4_chunglu18k.md",
            "hbm_GBs": traffic / t / 1e9 if traffic else None,
            "hbm_frac": traffic / t / 1e9 / HBM_PEAK_GBS if traffic else None,
            "alg_bytes": alg, "alg_GBs": alg / t / 1e9, "alg_frac_of_hbm_peak": alg / t / 1e9 / HBM_PEAK_GBS,
            "bytes_per_subgraph": alg / max(1, total)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=7)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=150.0,
                    help="seconds per CPU baseline run before it becomes a sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
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
    kern_ms = statistics.median(kms)
    progress("headline: %d steps timed, kernel %.2f ms" % (args.steps, kern_ms))
    st = g.last_stats()
    value = args.steps * total / elapsed
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
        "roofline": roofline(st, kern_ms, total),
    }
    if not args.no_extra:
        ge = W.er_2000()
        line["er2000"] = {"k%d" % k: query_block(ge, "clique", k, dm.config(), 1, 3) for k in (3, 4)}
        progress("er2000: cliques k=3/4")
        gr = W.rmat_s17()
        line["clique_rmat17"] = plain_vs_learned(gr, "clique", (5, 6), args.warmup, args.steps)
        gb = W.ba_100k_planted()
        line["motif5_ba"] = plain_vs_learned(gb, "motif", (5,), 0, 1, budget_ms=2000.0, quality=False)
        recs = {"clique_k5_rmat17": line["clique_rmat17"]["k5"], "clique_k6_rmat17": line["clique_rmat17"]["k6"],
                "motif_k5_ba": line["motif5_ba"]["k5"]}
        line["learned_speedup"] = {n: r["learned_speedup"] for n, r in recs.items()}
        line["learned_kernel_speedup"] = {n: r["learned_kernel_speedup"] for n, r in recs.items()}
        line["learned_kernel_speedup_train_k_minus_1"] = {
            n: r["learned_kernel_speedup_train_k_minus_1"] for n, r in recs.items()
            if "learned_kernel_speedup_train_k_minus_1" in r}
    if not args.no_batch:
        line["batch256"] = batch_block(min(args.warmup, 1), max(1, args.steps // 3))
    if not args.no_cpu:
        runs = [cpu_run(g, "motif", K, args.cpu_budget, label="chunglu_18k")]
        line["cpu_baseline"] = dict(runs[0])
        if not args.no_extra:
            runs.append(cpu_run(ge, "clique", 4, args.cpu_budget, label="er_2000"))
            # the faithful port needs hours here (hub neighbourhoods scanned
            # linearly): a seeded sample, beside the FAST variant on all roots
            runs.append(cpu_run(gr, "clique", 5, 20.0, random_order=True, label="rmat_s17", chunk_cap=64, chunk_s=1.0))
            runs.append(cpu_run(gr, "clique", 5, args.cpu_budget, faithful=False, oriented=True, label="rmat_s17"))
        host = {"cpu_model": host_cpu(), "nproc": os.cpu_count(), "omp_threads": runs[0]["cores"],
                "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND")}
        line["cpu_baseline"].update(host)
        line["cpu_baselines"] = runs
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
