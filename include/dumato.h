/*
 * dumato_amd — MI355X (gfx950) graph pattern mining runtime, C ABI.
 *
 * The reference (samuelbferraz/DuMato) has no library boundary: callers use
 * its executables (clique_PA, motifs), its host class DuMatoCPU and its
 * device API DuMatoGPU.  Each entry point below replaces one of those; the
 * reference interface it stands in for is cited beside it.
 *
 * Conventions: every function returns DM_OK (0) or a negative DM_ERR_* code
 * and never throws; dm_last_error() gives the message of the last failure on
 * the calling thread.  Output buffers are owned by the caller.  A dm_graph
 * handle is not thread-safe: use it from one host thread at a time.  One GPU.
 */
#ifndef DUMATO_AMD_H
#define DUMATO_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DM_OK 0
#define DM_ERR_ARG (-1)         /* bad argument (k out of range, short buffer, NULL) */
#define DM_ERR_IO (-2)          /* edgelist cannot be opened or parsed */
#define DM_ERR_HIP (-3)         /* HIP runtime error (no GPU, launch failure) */
#define DM_ERR_UNSUPPORTED (-4) /* pattern size without a class table, etc. */
#define DM_ERR_NOMEM (-5)       /* host or device allocation failed */
#define DM_ERR_DEVICE (-6)      /* kernel reported overflow or watchdog expiry */

#define DM_SCHED_PLAIN 0   /* degree-sorted static partition + work stealing */
#define DM_SCHED_LEARNED 1 /* MLP cost-model ordering/binning + work stealing */

typedef struct dm_graph dm_graph;

/* Launch and scheduling knobs.  The reference's equivalents are the argv of
 * clique_PA / motifs (threads, blockSize, donationsPerWarp, threshold;
 * src/main/clique_PA.cu:35-60) and the hard-coded virtual-warp width of
 * DM_WCV (src/DM_WCV/motifs_DM_WCV8.cu:65). */
typedef struct dm_config {
    int32_t sched;           /* DM_SCHED_PLAIN or DM_SCHED_LEARNED */
    int32_t vwave;           /* lanes per DFS worker: 8, 16, 32, 64; 0 = autotuned */
    int32_t steal;           /* 1: idle workers receive donated subtrees */
    int32_t steal_min;       /* donate only levels with >= steal_min children left
                                (the last level: >= 4 * steal_min); 0 = autotuned
                                (untuned: 2 for motifs, 32 for cliques) */
    int32_t steal_period;    /* microseconds between a worker's checks for idle
                                workers (0: every DFS step) */
    int32_t lds_bytes;       /* LDS stack bytes per worker; 0 = autotuned */
    int32_t waves_per_block; /* wave64s per block: 1, 2 or 4 (0 = 4) */
    int32_t blocks;          /* 0 = as many as stay resident on the GPU */
    int32_t record_costs;    /* 1: per-root cycle counts (dm_last_root_costs) */
    int32_t per_root;        /* 1: per-root subgraph counts (dm_last_root_counts) */
    int32_t root_begin;      /* enumerate roots [root_begin, root_end) only... */
    int32_t root_end;        /* ...(original vertex ids); root_end <= 0: all */
    int32_t clique_bitmap;   /* 1: k=4..6 cliques of roots with out-degree <= 256 count
                                on a local adjacency bitmap; 0: list intersections */
    int32_t repartition;     /* 1: compute the static partition afresh (features,
                                inference, binning) instead of reusing the cached one */
    double watchdog_s;       /* device-side time limit per launch (seconds) */
} dm_config;

typedef struct dm_stats {
    double device_ms;         /* enumeration kernel time (HIP events) */
    double wall_ms;           /* host time of the whole call */
    double sched_ms;          /* learned scheduling (features + inference + binning) */
    uint64_t subgraphs;       /* count returned by the call */
    uint64_t alg_bytes;       /* algorithmic bytes of the enumeration */
    uint64_t donations;       /* subtrees donated to idle workers */
    uint64_t jobs;            /* jobs executed (static + donated) */
    uint64_t busy_sum_cycles; /* sum over workers of cycles spent inside jobs */
    uint64_t busy_max_cycles; /* largest per-worker busy cycles (makespan proxy) */
    int32_t vwave, workers, blocks, lds_bytes;
    /* timeline from the first block's start (ms): the latest worker to start
     * its first job, to find the static queues empty, to end its last job,
     * and to exit */
    double first_job_ms, static_done_ms, last_job_ms, exit_ms;
    uint64_t job_max_cycles;  /* cycles of the longest single job (a lower bound
                                 on the makespan that stealing cannot cut) */
    double infer_ms;          /* learned: feature scaling and the MLP forward launch */
    double bin_ms;            /* learned: splitting, ordering and binning the roots */
    uint64_t active_workers;  /* workers that ran at least one job (a batch:
                                 the larger of its two launches) */
} dm_stats;

/* Fills *cfg with the defaults (plain scheduler, autotuned width/LDS). */
int dm_config_init(dm_config *cfg);

/* ---- graphs ---- */
/* Reads "src dst" lines (reference: Graph::Graph, src/main/Graph.cpp:17-54):
 * symmetric, de-duplicated, ids 0..maxId, isolated ids kept; self-loops are
 * dropped (they change no clique or motif count). */
int dm_graph_load_edgelist(const char *path, dm_graph **out);
/* Same from arrays; n = max(n, max id + 1). */
int dm_graph_from_edges(const int32_t *src, const int32_t *dst, size_t m,
                        int32_t n, dm_graph **out);
/* Vertex count, undirected edge count, maximum degree
 * (reference: Graph::getNumberOfVertices/getNumberOfEdges/getMaxDegree). */
int dm_graph_info(const dm_graph *g, int32_t *n, int64_t *m, int32_t *maxdeg);
/* Copies the symmetric CSR in original ids: off[n+1], adj[2m]. */
int dm_graph_csr(const dm_graph *g, int64_t *off, size_t off_len,
                 int32_t *adj, size_t adj_len);
void dm_graph_free(dm_graph *g);

/* Seeded synthetic generators (no reference counterpart). */
int dm_gen_erdos_renyi(int32_t n, int64_t m, uint64_t seed, dm_graph **out);
int dm_gen_chung_lu(int32_t n, int64_t m_target, double gamma, double wmax,
                    uint64_t seed, dm_graph **out);
int dm_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                double c, uint64_t seed, dm_graph **out);
int dm_gen_barabasi_albert(int32_t n, int32_t m_per_vertex,
                           int32_t planted, int32_t kmin, int32_t kmax,
                           uint64_t seed, dm_graph **out);

/* ---- counting (the hot path) ---- */
/* Number of k-cliques, 2 <= k <= 12.  Replaces clique_PA's kernel + host
 * loop (src/main/clique_PA.cu:9-32, DuMatoCPU::runKernel/copyResult/
 * outputAggregateCounter, src/main/DuMatoCPU.cu:316-362,419-423). */
int dm_clique_count(dm_graph *g, int k, const dm_config *cfg, uint64_t *out);

/* Histogram of connected induced k-vertex subgraphs, 3 <= k <= 7, slot i =
 * the reference's class id cgLocal i (dictionaries/k.csv); hist_len must be
 * dm_motif_num_classes(k).  k = 7 has no reference dictionary: its 1043 slots
 * are this library's own class ids (isomorphism classes ranked by their
 * largest quick code), so only its per-class counts, not its ids, match the
 * reference's semantics.  Replaces motifs' kernel + host loop
 * (src/main/motifs_DM_WCLB.cu:9-36, DuMatoCPU::outputAggregatePattern,
 * src/main/DuMatoCPU.cu:364-384).  *total (nullable) = sum of hist. */
int dm_motif_count(dm_graph *g, int k, const dm_config *cfg, uint64_t *hist,
                   size_t hist_len, uint64_t *total);

/* Counting that allocates no device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * results as dm_clique_count (motif = 0: out[0]) and dm_motif_count
 * (motif = 1: out[0 .. dm_motif_num_classes(k)), *total = their sum).
 * Every device buffer beyond the graph and the result comes from one static
 * device array of DM_NOALLOC_SCRATCH_BYTES per GPU, which keeps the last
 * call's contents; calls on one GPU are serialised by the library, and other
 * work on that array while a call runs would race with it.  Limits: the
 * graph must already be on the GPU (its first counting or tuning call
 * uploads it); plain scheduler only, no per-root records or root ranges;
 * the worker count is capped so that the stack arenas fit, and a graph whose
 * arena for one block exceeds the array fails with DM_ERR_ARG. */
#define DM_NOALLOC_SCRATCH_BYTES (64u << 20)
int dm_count_noalloc(dm_graph *g, int motif, int k, const dm_config *cfg, uint64_t *out,
                     size_t out_len, uint64_t *total);

/* Statistics and optional per-root records of the last counting call on g.
 * Root arrays are indexed by original vertex id (n entries). */
int dm_last_stats(const dm_graph *g, dm_stats *st);
int dm_last_root_costs(const dm_graph *g, uint64_t *cycles, size_t n);
int dm_last_root_counts(const dm_graph *g, uint64_t *counts, size_t n);

/* ---- pattern tables (reference: QuickMapping, src/main/QuickMapping.cpp:5-43) ---- */
/* Number of classes for k (3, 10, 33, 155, 1043 for k = 3..7). */
int dm_motif_num_classes(int k, size_t *ncls);
/* Regenerated tables: quick code -> class id (2^((k+1)(k-2)/2) entries) and
 * class id -> canonical lower-triangle code (the reference's cg_global). */
int dm_pattern_table(int k, uint16_t *quick_to_cls, size_t nquick,
                     uint64_t *cls_to_code, size_t ncls);

/* ---- learned scheduler (no reference counterpart; the reference analogue is
 * the static partition of DuMatoCPU, src/main/DuMatoCPU.cu:121-135) ---- */
/* Records per-root cycle costs of a plain k_train run on the GPU and trains the
 * 8-64-64-1 MLP cost model on them with the library's HIP kernels.  The model
 * is kept in g and used by later calls with sched = DM_SCHED_LEARNED: their
 * partition gives the predicted-heaviest roots many jobs (shared-cursor jobs
 * over a prebuilt bitmap on the clique bitmap path), orders all jobs heaviest
 * first and sets per-job donation thresholds.  A clique query larger than
 * k_train carries the prediction up by each root's clique growth (out-degree
 * and oriented triangle density). */
int dm_sched_train(dm_graph *g, int mode_motif, int k_train, int epochs,
                   uint64_t seed, double *final_loss);
/* Autotuner (no reference counterpart; the reference fixes the virtual-warp
 * width per binary and takes donationsPerWarp/threshold from argv): times
 * candidate widths (8/16/32/64), LDS stack sizes and steal granularities on
 * a degree-stratified sample of the roots within about budget_ms of device
 * time, and keeps the fastest for (mode, k) in g.  Later calls with vwave,
 * lds_bytes or steal_min = 0 use it.  *best (nullable) receives it. */
int dm_autotune(dm_graph *g, int mode_motif, int k, double budget_ms,
                dm_config *best);
/* Predicted cost (model output, log-cycles) per original vertex id. */
int dm_sched_predict(dm_graph *g, float *pred, size_t n);
/* The MLP kernels on their own, for checking against a CPU reference.
 * params: 4801 floats, W1[8][64], b1[64], W2[64][64], b2[64], W3[64], b3.
 * dm_mlp_grad: loss = mean_b (f(X_b) - Y_b)^2, its gradient, f(X) (nullable).
 * dm_mlp_adam: `steps` full-batch Adam steps (beta 0.9/0.999, eps 1e-8) in
 * place on params; losses[steps] (nullable) before each step. */
int dm_mlp_grad(const float *params, const float *X, const float *Y, int n,
                float *out, float *grad, float *loss);
int dm_mlp_adam(float *params, const float *X, const float *Y, int n, int steps,
                float lr, float *losses);

/* ---- batches (no reference counterpart; the reference runs one graph and
 * one k per executable, src/main/clique_PA.cu:34-130) ---- */
typedef struct dm_batch dm_batch;
typedef struct dm_query {
    int32_t graph;    /* index into the batch's graphs */
    int32_t motif;    /* 0: k-clique count, 1: k-motif histogram (k <= 6) */
    int32_t k;
    int32_t reserved;
} dm_query;
/* Copies the graphs (their degree-rank CSRs) into one batch; the graphs may
 * be freed afterwards. */
int dm_batch_create(dm_graph *const *graphs, size_t ngraphs, dm_batch **out);
/* Result slots the queries need: 1 per clique query, dm_motif_num_classes(k)
 * per motif query. */
int dm_batch_num_results(const dm_batch *b, const dm_query *queries, size_t nq,
                         size_t *nres);
/* Runs every query in one persistent kernel launch, captured in a hipGraph
 * the first time a query list / launch shape is seen and replayed after.
 * results[result_offsets[i] ..) holds query i's count or histogram;
 * graph_latency_ms[g] (nullable) = time from launch start until graph g's
 * last query finished. */
int dm_batch_run(dm_batch *b, const dm_query *queries, size_t nq,
                 const dm_config *cfg, uint64_t *results, size_t results_len,
                 uint64_t *result_offsets, double *graph_latency_ms);
/* Times whole-batch replays for widths 64/32/16 and steal granularities 2/8
 * within about budget_ms and keeps the fastest for later dm_batch_run calls
 * with vwave = 0. */
int dm_batch_autotune(dm_batch *b, const dm_query *queries, size_t nq,
                      double budget_ms, dm_config *best);
int dm_batch_last_stats(const dm_batch *b, dm_stats *st);
void dm_batch_free(dm_batch *b);

/* ---- DFS-wide step API and host-driven load balancing ----
 * The reference's programming model: a user kernel drives one DFS per
 * (virtual) warp through the DuMatoGPU steps start / extend / filter /
 * compact / aggregate / forward / backward (src/main/clique_PA.cu:9-32,
 * src/main/motifs_DM_WCLB.cu:9-36, src/DM_WCV/clique_DM_WCV8.cu:8-34), and
 * the host class DuMatoCPU launches it, polls idle warps, stops the kernel,
 * moves work between warps and relaunches (src/main/DuMatoCPU.cu:316-562,
 * src/main/EnumerationHelper.cpp:90-140, the loop of
 * src/main/motifs_DM_WCLB.cu:85-120).  Here the steps are the header-only
 * HIP API of include/dumato_device.h over per-worker state in device
 * memory, and dm_dfs_run is the host loop.  (The fused kernels behind
 * dm_clique_count / dm_motif_count are the fast path; this is the
 * reference-shaped one.) */

/* Per-worker DFS state in device memory, passed by value to the kernel.
 * Vertex ids are the graph's degree ranks (ascending (degree, id)). */
#define DM_DFS_MAXK 16
typedef struct dm_dfs_data {
    const int64_t *off;       /* symmetric CSR: off[n+1], adj, each list ascending */
    const int32_t *adj;
    const int64_t *ostart;    /* ostart[v]: first index of N(v) above v */
    int32_t n, K, motif, workers;
    int32_t vwave, jobs_per_worker;
    int32_t cap_base;         /* list capacity unit: max degree (motif) or max out-degree */
    int32_t ncls;             /* motif classes (block histogram slots) */
    int64_t ext_per_worker;   /* int32 entries of a worker's level lists */
    int32_t *ext;             /* level lists (workers x ext_per_worker) */
    int8_t *src;              /* motif: first neighbourhood of each list entry */
    int32_t *k;               /* [workers] current level; -1: take the next job */
    int32_t *status;          /* [workers] 1 working, 2 idle (no job left) */
    int32_t *id;              /* [workers x DM_DFS_MAXK] vertex at each level */
    int32_t *next;            /* [workers x DM_DFS_MAXK] extensions left per level; -1: not generated */
    int32_t *fixed;           /* [workers x DM_DFS_MAXK] full size of each level's list */
    uint64_t *ind;            /* [workers x DM_DFS_MAXK] induction (quick bits) per level */
    int32_t *jobs;            /* [workers x jobs_per_worker x DM_DFS_MAXK] job prefixes */
    int32_t *job_level;       /* [workers x jobs_per_worker] */
    uint64_t *job_ind;        /* [workers x jobs_per_worker] induction of the job's last level */
    int32_t *job_valid;       /* [workers] jobs in the queue */
    int32_t *job_cur;         /* [workers] next job to take */
    const uint16_t *lut;      /* motif: quick code -> class */
    unsigned long long *result; /* clique count, or the histogram (ncls slots) */
    const int32_t *stop;      /* host sets 1: every worker leaves its loop at the next step */
} dm_dfs_data;

#define DM_DFS_CLIQUE_PA 0   /* clique_PA.cu: extend_clique / forward / backward_clique */
#define DM_DFS_CLIQUE_WC 1   /* clique_DM_WCV8.cu: extend / filterClique / forward / backward */
#define DM_DFS_MOTIF 2       /* motifs_DM_WCLB.cu: extend(0,k) / canonicalFilter2 / forward_induce */
#define DM_DFS_NEAR_CLIQUE 3 /* q4_PA.cu .. q13_PA.cu: extend_q / forward / backward_q */

/* What a step-API state enumerates (dm_dfs_create). */
#define DM_PATTERN_CLIQUE 0      /* k-cliques: clique_PA / clique_DM_WCV kernels */
#define DM_PATTERN_MOTIF 1       /* k-motif histogram: motifs kernel */
#define DM_PATTERN_NEAR_CLIQUE 2 /* the pattern-aware query qk (4 <= k <= 13): a
                                    (k-2)-clique and two vertices adjacent to all of it,
                                    counted per clique and unordered pair of common
                                    neighbours (K_k minus one edge, not induced) */

typedef struct dm_dfs_config {
    int32_t vwave;           /* lanes per DFS worker: 8, 16, 32, 64 (0 = 32, the reference's warp) */
    int32_t blocks;          /* 0 = as many as stay resident */
    int32_t jobs_per_worker; /* job-queue slots per worker (reference donationsPerWarp); 0 = 16 */
    int32_t threshold;       /* rebalance when more than threshold % of the workers are idle
                                (reference argv threshold); >= 100: no load balancing */
    int32_t poll_us;         /* host poll period of the idle count (reference: 100 ms); 0 = 1000 */
    int32_t stop_every_us;   /* > 0: also stop, snapshot and rebalance this often */
    int32_t max_rebalances;  /* 0 = unlimited */
    int32_t reserved;
    double watchdog_s;       /* host time limit of the whole run (0 = none) */
} dm_dfs_config;

typedef struct dm_dfs_stats {
    double total_ms;         /* run time: launches, polls and rebalancing */
    double lb_ms;            /* time spent stopped: snapshot, donation, copy back */
    int32_t launches, rebalances, workers, vwave;
    uint64_t donated_jobs;
    uint64_t subgraphs;
} dm_dfs_stats;

typedef struct dm_dfs dm_dfs;
int dm_dfs_config_init(dm_dfs_config *cfg);
/* Device state for enumerating the DM_PATTERN_* `pattern` of size k in g with
 * the step API (reference: the DuMatoCPU constructor,
 * src/main/DuMatoCPU.cu:13-30, 204-276). */
int dm_dfs_create(dm_graph *g, int k, int pattern, const dm_dfs_config *cfg, dm_dfs **out);
/* Host stub of a kernel of this library written against the step API
 * (kind = DM_DFS_*) for the given width, for dm_dfs_run. */
int dm_dfs_kernel(int kind, int vwave, const void **kernel);
/* Runs `kernel` (a __global__ void(dm_dfs_data) built against
 * include/dumato_device.h with the state's width) from the static partition
 * (root v to worker v mod workers, by degree rank), with the host-driven load
 * balancing of the config, until every job is done.  out: the clique count
 * (out_len >= 1) or the histogram (out_len >= dm_motif_num_classes(k)).
 * Replaces runKernel / gpuIsIdle / stopKernel / rebalance / waitKernel /
 * outputAggregate* (src/main/DuMatoCPU.cu:316-562). */
int dm_dfs_run(dm_dfs *s, const void *kernel, uint64_t *out, size_t out_len, dm_dfs_stats *st);
void dm_dfs_free(dm_dfs *s);

/* ---- validation against Peregrine output files ---- */
/* Reads "<graph_path>.<k>.aggregateCounter" and sets *match = 1 when the sum
 * of its '['-line values equals result.  Replaces
 * DuMatoCPU::validateAggregateCounter (src/main/DuMatoCPU.cu:564-595).
 * DM_ERR_IO when the file cannot be opened (the reference prints an error and
 * returns).  Values are 64-bit (the reference's atoi wraps above 2^31-1). */
int dm_validate_counter(const char *graph_path, int k, uint64_t result, int *match);
/* Reads "<graph_path>.<k>.aggregatePattern"; *count_match = 1 when its total
 * equals sum(hist), *pattern_match = 1 when the sorted multiset of its positive
 * values equals that of hist's positive slots (class ids are not compared).
 * file_total / file_patterns (nullable) receive the file's total and number of
 * positive patterns.  Replaces DuMatoCPU::validateAggregatePattern
 * (src/main/DuMatoCPU.cu:597-661). */
int dm_validate_pattern(const char *graph_path, int k, const uint64_t *hist, size_t hist_len,
                        int *count_match, int *pattern_match, uint64_t *file_total,
                        size_t *file_patterns);

const char *dm_last_error(void);
int dm_version(void);

#ifdef __cplusplus
}
#endif
#endif
