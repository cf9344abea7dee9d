// Batched queries over many graphs in one persistent launch, captured once in
// a hipGraph and replayed.  No reference counterpart: DuMato runs one graph
// and one k per executable (src/main/clique_PA.cu:34-130).
//
// The graphs are concatenated into one disjoint-union CSR (each relabelled by
// degree rank inside its own id range, so cliques and motifs never cross
// graphs); every (graph, pattern, k) query gets a descriptor with its class
// table and result slots, and static range jobs over its graph's roots.  The
// kernel switches query per job; donation and counting are the single-graph
// kernel's.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <numeric>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

#include "internal.h"
#include "runtime.h"

namespace dm {

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return set_error(DM_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));     \
    } while (0)

struct BatchPlan {
    std::vector<dm_query> queries;
    int vwave = 0, lds = 0, steal = 1, steal_min = 2, steal_period = 2, bitmap = 1, blocks = 0, threads = 256;
    double watchdog_s = 0;
    bool same(const std::vector<dm_query> &q, const dm_config &c, int vw, int l, int sm, int thr) const {
        if (q.size() != queries.size()) return false;
        for (size_t i = 0; i < q.size(); i++)
            if (q[i].graph != queries[i].graph || q[i].motif != queries[i].motif || q[i].k != queries[i].k)
                return false;
        return vw == vwave && l == lds && c.steal == steal && sm == steal_min && c.steal_period == steal_period &&
               c.clique_bitmap == bitmap && c.blocks == blocks && c.watchdog_s == watchdog_s && thr == threads;
    }
};

}  // namespace dm

struct dm_batch {
    int ngraphs = 0;
    std::vector<int32_t> vbase;      // first union id of each graph
    std::vector<int32_t> gn;         // vertices per graph
    dm::HostCSR u;                   // union CSR (degree-rank ids per graph)
    std::vector<int64_t> ostart;
    int32_t max_out = 0;
    // device: the union CSR (and the device context whose stream, queue,
    // epoch and class tables the batch uses), and the workspaces the captured
    // hipGraph points to, which are the batch's own
    dm::DeviceGraph *dev = nullptr;
    char *d_arena = nullptr;
    size_t arena_cap = 0;
    int64_t *d_sq_off = nullptr;
    size_t sq_off_cap = 0;
    int32_t *d_sq_jobs = nullptr;
    size_t sq_jobs_cap = 0;
    int32_t *d_sq_cur = nullptr;
    size_t sq_cur_cap = 0;
    unsigned long long *d_qmask = nullptr;  // non-empty static queues (shared by the phases)
    size_t qmask_cap = 0;
    unsigned long long *d_results = nullptr;
    size_t results_cap = 0;
    dm::QueryDesc *d_qd = nullptr;
    size_t qd_cap = 0;
    unsigned long long *d_done = nullptr;
    size_t done_cap = 0;
    unsigned long long *d_start = nullptr;
    size_t start_cap = 0;
    dm::SchedCounters *d_sc = nullptr;  // one per phase
    size_t sc_cap = 0;
    int nphases = 1;
    int tuned_steal_min = 0;  // from dm_batch_autotune; 0 = not tuned
    // cached launch
    dm::BatchPlan plan;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t nres = 0;
    std::vector<uint64_t> res_off;
    std::vector<int32_t> q_graph;
    dm_stats last{};
};

namespace dm {

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

static void drop_graph(dm_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static int batch_device(dm_batch *b) {
    if (b->dev)
        return hipSetDevice(b->dev->ctx->device) == hipSuccess ? DM_OK : set_error(DM_ERR_HIP, "hipSetDevice");
    DeviceCtx *c = nullptr;
    int r = device_ctx(&c);
    if (r) return r;
    DeviceGraph *d = new DeviceGraph();
    d->ctx = c;
    if ((r = upload_csr(b->u, b->ostart, d))) {
        free_device_graph(d);
        return r;
    }
    b->dev = d;
    return DM_OK;
}

// Range jobs: each query's roots in chunks; a chunk's cost estimate is the
// sum over its roots of (out-degree)^(k-2) (cliques) or degree^(k-2)
// (motifs).  The smallest graphs' chunks first, then the rest heaviest
// first; dealt round-robin.
static void batch_partition(const dm_batch *b, const std::vector<dm_query> &qs, int workers,
                            std::vector<int64_t> &sq_off, std::vector<int32_t> &sq_jobs) {
    struct Piece {
        double cost;
        int32_t r0, r1, q;
    };
    std::vector<Piece> pieces;
    const int chunk = 32;
    for (size_t qi = 0; qi < qs.size(); qi++) {
        const dm_query &Q = qs[qi];
        const int32_t v0 = b->vbase[Q.graph], v1 = v0 + b->gn[Q.graph];
        for (int32_t r = v0; r < v1; r += chunk) {
            const int32_t e = std::min(r + chunk, v1);
            double c = 0;
            for (int32_t v = r; v < e; v++) {
                double d = Q.motif ? (double)(b->u.off[v + 1] - b->u.off[v]) : (double)(b->u.off[v + 1] - b->ostart[v]);
                c += std::pow(d, std::max(1, Q.k - 2));
            }
            if (c > 0) pieces.push_back({c, r, e, (int32_t)qi});
        }
    }
    std::sort(pieces.begin(), pieces.end(), [](const Piece &a, const Piece &c) { return a.cost > c.cost; });
    // Per-graph latency: the smallest graphs, up to kEarlyShare of the
    // estimated work, go first, one graph after another (each heaviest
    // first), then every other piece heaviest first.  Dealt heaviest first
    // alone, every graph's light pieces come last and every graph finishes
    // within 5% of the end; this way half of them finish in the first fifth
    // of the launch, and the heavy pieces start only that small share later
    // (256-graph batch: p50 9.39 -> 1.72 s, makespan unchanged, with a 5%
    // share at 4 waves per SIMD).  With the 6-wave motif kernel a 5% share
    // gave p50 2.42 s, 10-35% 2.04-2.05 s (makespan 8.98 -> 8.94 s), 2% 8.2 s.
    constexpr double kEarlyShare = 0.20;
    {
        std::vector<double> gcost(b->ngraphs, 0.0);
        double total = 0;
        for (const Piece &pc : pieces) {
            gcost[qs[pc.q].graph] += pc.cost;
            total += pc.cost;
        }
        std::vector<int32_t> gord(b->ngraphs);
        std::iota(gord.begin(), gord.end(), 0);
        std::stable_sort(gord.begin(), gord.end(), [&](int32_t a, int32_t c) { return gcost[a] < gcost[c]; });
        std::vector<int32_t> grank(b->ngraphs, -1);  // place among the early graphs, -1: not early
        double acc = 0;
        int32_t ne = 0;
        for (int32_t gi : gord) {
            if (gcost[gi] <= 0) continue;
            if (acc + gcost[gi] > kEarlyShare * total) break;
            acc += gcost[gi];
            grank[gi] = ne++;
        }
        std::stable_sort(pieces.begin(), pieces.end(), [&](const Piece &a, const Piece &c) {
            const int32_t ra = grank[qs[a.q].graph], rc = grank[qs[c.q].graph];
            if ((ra >= 0) != (rc >= 0)) return ra >= 0;
            return ra >= 0 && ra < rc;
        });
    }
    sq_off.assign((size_t)workers + 1, 0);
    const int64_t n = (int64_t)pieces.size();
    for (int w = 0; w < workers; w++) sq_off[w + 1] = sq_off[w] + (n > w ? (n - w + workers - 1) / workers : 0);
    sq_jobs.assign((size_t)n * 4, 0);
    for (int w = 0; w < workers; w++) {
        int64_t o = sq_off[w];
        for (int64_t i = w; i < n; i += workers, o++) {
            sq_jobs[4 * o] = pieces[i].r0;
            sq_jobs[4 * o + 1] = -1;
            sq_jobs[4 * o + 2] = pieces[i].r1;
            sq_jobs[4 * o + 3] = pieces[i].q;
        }
    }
}

// The launch is captured as one hipGraph.  A batch that mixes clique and
// motif queries runs one persistent kernel per family inside it (clique
// queries first), so that each family gets its specialised kernel and
// occupancy; the two share the queue, the result slots and the clock.
static int batch_build(dm_batch *b, const std::vector<dm_query> &qs, const dm_config &cfg) {
    DeviceGraph *dg = b->dev;
    DeviceCtx *d = dg->ctx;
    int kmax_f[2] = {2, 2}, hslots = 0;
    bool has[2] = {false, false};
    for (const dm_query &Q : qs) {
        if (Q.graph < 0 || Q.graph >= b->ngraphs) return set_error(DM_ERR_ARG, "query graph %d out of range", Q.graph);
        if (Q.motif) {
            if (Q.k < 3 || Q.k > 6) return set_error(DM_ERR_UNSUPPORTED, "batch motif k=%d (3..6)", Q.k);
            const PatternTable *t = pattern_table(Q.k);
            if (!t) return last_code();
            hslots = std::max(hslots, t->ncls);
        } else if (Q.k < 2 || Q.k > kMaxK) {
            return set_error(DM_ERR_ARG, "batch clique k=%d (2..%d)", Q.k, kMaxK);
        }
        const int f = Q.motif ? 1 : 0;
        has[f] = true;
        kmax_f[f] = std::max(kmax_f[f], Q.k);
    }
    // launch shape: width / LDS from the config (autotuned values cached in the plan)
    int vw = cfg.vwave ? cfg.vwave : (b->plan.vwave ? b->plan.vwave : 64);
    if (vw != 8 && vw != 16 && vw != 32 && vw != 64) return set_error(DM_ERR_ARG, "vwave must be 8/16/32/64");
    int threads = 256;
    int r = block_threads(cfg, &threads);
    if (r) return r;
    // 0: each phase takes the largest LDS stack its occupancy allows
    const int lds = cfg.lds_bytes > 0 ? (cfg.lds_bytes + 15) & ~15 : 0;
    const int sm = cfg.steal_min ? cfg.steal_min : (b->tuned_steal_min ? b->tuned_steal_min : 2);
    if (b->exec && b->plan.same(qs, cfg, vw, lds, sm, threads)) return DM_OK;
    drop_graph(b);
    // phases: (mode, kmax, per-worker class slots) per family present; a
    // single query keeps the batch's per-worker counters (mode 2)
    struct Phase {
        int mode, kmax, wslots, blocks, workers, lds;
        size_t lds_total;
        std::vector<int64_t> sq_off;
        std::vector<int32_t> sq_jobs;
        std::vector<dm_query> qs;     // the family's queries
        std::vector<int32_t> qidx;    // their indices in qs
    };
    std::vector<Phase> ph;
    for (int f = 0; f < 2; f++) {
        if (!has[f]) continue;
        Phase P;
        P.kmax = kmax_f[f];
        P.mode = qs.size() == 1 ? 2 : f == 1 ? launch_mode(true, P.kmax) : 0;
        P.wslots = f == 1 ? hslots : 0;
        for (size_t i = 0; i < qs.size(); i++)
            if ((qs[i].motif ? 1 : 0) == f) {
                P.qs.push_back(qs[i]);
                P.qidx.push_back((int32_t)i);
            }
        ph.push_back(std::move(P));
    }
    int max_workers = 0;
    size_t sq_total = 0, jobs_total = 0;
    for (Phase &P : ph) {
        // one query: the kernel keeps block-wide class counters (sized by
        // the query's classes) instead of per-worker ones
        const int ncls_single = qs.size() == 1 ? P.wslots : 0;
        const int hs = qs.size() == 1 ? 0 : P.wslots;
        P.lds = lds > 0 ? lds : max_lds_per_worker(vw, threads, P.mode, P.kmax, ncls_single, hs);
        P.lds_total = enum_lds_bytes(vw, threads, P.kmax, ncls_single, hs, P.lds);
        int bpc = 0;
        if ((r = enum_occupancy(vw, P.mode, threads, P.lds_total, &bpc))) return r;
        if (bpc < 1) return set_error(DM_ERR_ARG, "batch launch does not fit one CU (LDS %zu B)", P.lds_total);
        P.blocks = cfg.blocks > 0 ? cfg.blocks : bpc * d->num_cu;
        P.workers = P.blocks * (threads / vw);
        batch_partition(b, P.qs, P.workers, P.sq_off, P.sq_jobs);
        // job records name queries by their index in the whole batch
        for (size_t j = 0; j < P.sq_jobs.size(); j += 4) P.sq_jobs[j + 3] = P.qidx[P.sq_jobs[j + 3]];
        max_workers = std::max(max_workers, P.workers);
        sq_total += P.sq_off.size();
        jobs_total += P.sq_jobs.size();
    }
    // result slots
    b->res_off.assign(qs.size() + 1, 0);
    for (size_t i = 0; i < qs.size(); i++)
        b->res_off[i + 1] = b->res_off[i] + (qs[i].motif ? (uint64_t)pattern_table(qs[i].k)->ncls : 1);
    b->nres = b->res_off.back();
    int64_t arena = 256;
    for (const dm_query &Q : qs)
        arena = std::max(arena, arena_bound_for(b->u.maxdeg, b->max_out, Q.k, Q.motif != 0));
    if ((r = dev_grow(b->d_arena, b->arena_cap, (size_t)arena * max_workers))) return r;
    if ((r = dev_grow(b->d_sq_off, b->sq_off_cap, sq_total))) return r;
    if ((r = dev_grow(b->d_sq_jobs, b->sq_jobs_cap, std::max<size_t>(jobs_total, 4)))) return r;
    if ((r = dev_grow(b->d_sq_cur, b->sq_cur_cap, (size_t)max_workers))) return r;
    if ((r = dev_grow(b->d_qmask, b->qmask_cap, (size_t)qmask_words(max_workers, 1)))) return r;
    if ((r = dev_grow(b->d_results, b->results_cap, b->nres))) return r;
    if ((r = dev_grow(b->d_qd, b->qd_cap, qs.size()))) return r;
    if ((r = dev_grow(b->d_done, b->done_cap, qs.size()))) return r;
    if ((r = dev_grow(b->d_start, b->start_cap, 1))) return r;
    if ((r = dev_grow(b->d_sc, b->sc_cap, 2))) return r;
    std::vector<QueryDesc> qd(qs.size());
    b->q_graph.assign(qs.size(), 0);
    for (size_t i = 0; i < qs.size(); i++) {
        const dm_query &Q = qs[i];
        QueryDesc &D = qd[i];
        memset(&D, 0, sizeof(D));
        D.K = Q.k;
        D.motif = Q.motif ? 1 : 0;
        D.graph = Q.graph;
        b->q_graph[i] = Q.graph;
        if (Q.motif) {
            D.ncls = pattern_table(Q.k)->ncls;
            if ((r = device_lut(d, Q.k, &D.lut))) return r;
        }
        D.out = b->d_results + b->res_off[i];
        D.done_ticks = b->d_done + i;
    }
    HIPCHK(hipMemcpy(b->d_qd, qd.data(), qd.size() * sizeof(QueryDesc), hipMemcpyHostToDevice));
    size_t o_off = 0, o_jobs = 0;
    std::vector<EnumParams> params(ph.size());
    for (size_t pi = 0; pi < ph.size(); pi++) {
        Phase &P = ph[pi];
        HIPCHK(hipMemcpy(b->d_sq_off + o_off, P.sq_off.data(), P.sq_off.size() * 8, hipMemcpyHostToDevice));
        if (!P.sq_jobs.empty())
            HIPCHK(hipMemcpy(b->d_sq_jobs + o_jobs, P.sq_jobs.data(), P.sq_jobs.size() * 4, hipMemcpyHostToDevice));
        EnumParams &p = params[pi];
        memset(&p, 0, sizeof(p));
        p.off = dg->d_off;
        p.adj = dg->d_adj;
        p.ostart = dg->d_ostart;
        p.n = b->u.n;
        p.nq = (int32_t)qs.size();
        p.qd = b->d_qd;
        p.kmax = P.kmax;
        p.hist_slots = P.wslots;
        p.nworkers = P.workers;
        p.sq_off = b->d_sq_off + o_off;
        p.sq_jobs = b->d_sq_jobs + o_jobs;
        p.sq_cur = b->d_sq_cur;
        p.qmask = P.mode == 1 || P.mode == 3 ? b->d_qmask : nullptr;  // motif phases only, as in single-graph calls
        p.q = d->d_q;
        p.q_cap = (uint32_t)d->q_cap_alloc;
        p.epoch = d->d_epoch;
        p.start_ticks = b->d_start;
        p.sc = b->d_sc + pi;
        p.arena = b->d_arena;
        p.arena_per_worker = arena;
        p.lds_per_worker = P.lds;
        p.steal = cfg.steal ? 1 : 0;
        p.steal_min = sm;
        p.poll_ticks = (uint64_t)(cfg.steal_period >= 0 ? cfg.steal_period : 2) * (uint64_t)d->wall_khz / 1000;
        p.bitmap = cfg.clique_bitmap ? 1 : 0;
        p.watchdog_ticks = cfg.watchdog_s > 0 ? (uint64_t)(cfg.watchdog_s * d->wall_khz * 1000.0) : 0;
        o_off += P.sq_off.size();
        o_jobs += P.sq_jobs.size();
    }
    // capture: per phase an init kernel and the persistent kernel; only the
    // first init clears the results, latencies and start clock
    hipStream_t s = d->stream;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rr = DM_OK;
    for (size_t pi = 0; pi < ph.size() && !rr; pi++) {
        const Phase &P = ph[pi];
        const int64_t njobs = (int64_t)P.sq_jobs.size() / 4;
        int queues = 0;
        for (size_t w = 0; w + 1 < P.sq_off.size(); w++) queues += P.sq_off[w + 1] > P.sq_off[w] ? 1 : 0;
        EnumParams ip = params[pi];
        if (pi > 0) ip.start_ticks = nullptr;
        rr = enum_init_launch(ip, (unsigned long long)njobs, queues, d->d_epoch, pi == 0 ? b->d_results : nullptr,
                              pi == 0 ? b->nres : 0, pi == 0 ? b->d_done : nullptr, pi == 0 ? (int)qs.size() : 0, s);
        if (!rr && njobs > 0) rr = enum_launch(params[pi], P.mode, vw, P.blocks, threads, P.lds_total, s);
    }
    hipError_t ce = hipStreamEndCapture(s, &b->graph);
    if (rr) return rr;
    if (ce != hipSuccess) return set_error(DM_ERR_HIP, "graph capture: %s", hipGetErrorString(ce));
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->nphases = (int)ph.size();
    b->plan.queries = qs;
    b->plan.vwave = vw;
    b->plan.lds = lds;
    b->plan.steal = cfg.steal;
    b->plan.steal_min = sm;
    b->plan.steal_period = cfg.steal_period;
    b->plan.bitmap = cfg.clique_bitmap;
    b->plan.blocks = cfg.blocks;
    b->plan.watchdog_s = cfg.watchdog_s;
    b->plan.threads = threads;
    b->last.vwave = vw;
    b->last.workers = max_workers;
    b->last.blocks = ph.back().blocks;
    b->last.lds_bytes = ph.back().lds;
    return DM_OK;
}

static int batch_run(dm_batch *b, const std::vector<dm_query> &qs, const dm_config &cfg, uint64_t *results,
                     size_t results_len, uint64_t *offsets, double *latency_ms) {
    const double t0 = now_ms();
    int r = check_config(cfg);
    if (r) return r;
    if ((r = batch_device(b))) return r;
    DeviceCtx *d = b->dev->ctx;
    std::lock_guard<std::recursive_mutex> lock(d->mu);
    if ((r = batch_build(b, qs, cfg))) return r;
    const double t1 = now_ms();
    hipStream_t s = d->stream;
    if (results && results_len < b->nres) return set_error(DM_ERR_ARG, "results need %zu slots", b->nres);
    HIPCHK(hipEventRecord(d->ev0, s));
    HIPCHK(hipGraphLaunch(b->exec, s));
    HIPCHK(hipEventRecord(d->ev1, s));
    SchedCounters scs[2];
    HIPCHK(hipMemcpyAsync(scs, b->d_sc, sizeof(SchedCounters) * b->nphases, hipMemcpyDeviceToHost, s));
    std::vector<unsigned long long> res(b->nres), done(qs.size());
    unsigned long long start = 0;
    HIPCHK(hipMemcpyAsync(res.data(), b->d_results, b->nres * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(done.data(), b->d_done, qs.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&start, b->d_start, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float kms = 0;
    HIPCHK(hipEventElapsedTime(&kms, d->ev0, d->ev1));
    // both phases' counters: sums, maxima, and the timeline of the last phase
    // measured from the first one's start
    SchedCounters sc1 = scs[0];
    for (int pi = 1; pi < b->nphases; pi++) {
        const SchedCounters &c = scs[pi];
        if (!sc1.error) sc1.error = c.error;
        sc1.alg_bytes += c.alg_bytes;
        sc1.donations += c.donations;
        sc1.jobs += c.jobs;
        sc1.busy_sum += c.busy_sum;
        sc1.busy_max = std::max(sc1.busy_max, c.busy_max);
        sc1.job_max = std::max(sc1.job_max, c.job_max);
        sc1.active_workers = std::max(sc1.active_workers, c.active_workers);
        sc1.t_first_job = c.t_first_job;
        sc1.t_static_done = c.t_static_done;
        sc1.t_last_job = c.t_last_job;
        sc1.t_exit = c.t_exit;
    }
    if (sc1.error) return device_error(sc1.error, "batch kernel");
    uint64_t total = 0;
    for (size_t i = 0; i < b->nres; i++) total += res[i];
    if (results)
        for (size_t i = 0; i < b->nres; i++) results[i] = res[i];
    if (offsets)
        for (size_t i = 0; i <= qs.size(); i++) offsets[i] = b->res_off[i];
    if (latency_ms) {
        for (int gi = 0; gi < b->ngraphs; gi++) latency_ms[gi] = 0;
        for (size_t i = 0; i < qs.size(); i++) {
            double ms = done[i] > start ? (double)(done[i] - start) / d->wall_khz : 0.0;
            latency_ms[b->q_graph[i]] = std::max(latency_ms[b->q_graph[i]], ms);
        }
    }
    dm_stats &st = b->last;
    st.device_ms = kms;
    st.sched_ms = t1 - t0;
    st.wall_ms = now_ms() - t0;
    st.subgraphs = total;
    st.alg_bytes = sc1.alg_bytes;
    st.donations = sc1.donations;
    st.jobs = sc1.jobs;
    st.busy_sum_cycles = sc1.busy_sum;
    st.busy_max_cycles = sc1.busy_max;
    {
        auto rel = [&](unsigned long long t) {
            return t > sc1.t_start && sc1.t_start != ~0ull ? (double)(t - sc1.t_start) / d->wall_khz : 0.0;
        };
        st.first_job_ms = rel(sc1.t_first_job);
        st.static_done_ms = rel(sc1.t_static_done);
        st.last_job_ms = rel(sc1.t_last_job);
        st.exit_ms = rel(sc1.t_exit);
        st.job_max_cycles = sc1.job_max;
    }
    st.active_workers = sc1.active_workers;
    return DM_OK;
}

}  // namespace dm

using namespace dm;

extern "C" {

int dm_batch_create(dm_graph *const *graphs, size_t ngraphs, dm_batch **out) {
    if (!graphs || !out || ngraphs == 0) return set_error(DM_ERR_ARG, "need graphs and out");
    *out = nullptr;
    dm_batch *b = new dm_batch();
    b->ngraphs = (int)ngraphs;
    int64_t nv = 0, ne = 0;
    for (size_t i = 0; i < ngraphs; i++) {
        if (!graphs[i]) {
            delete b;
            return set_error(DM_ERR_ARG, "graph %zu is NULL", i);
        }
        nv += graphs[i]->rcsr.n;
        ne += graphs[i]->rcsr.nnz();
    }
    if (nv > INT_MAX - 1) {
        delete b;
        return set_error(DM_ERR_ARG, "batch has too many vertices");
    }
    b->u.n = (int32_t)nv;
    b->u.off.assign((size_t)nv + 1, 0);
    b->u.adj.resize((size_t)ne);
    b->ostart.assign((size_t)nv, 0);
    int32_t vb = 0;
    int64_t eb = 0;
    for (size_t i = 0; i < ngraphs; i++) {
        const dm_graph *g = graphs[i];
        b->vbase.push_back(vb);
        b->gn.push_back(g->rcsr.n);
        for (int32_t v = 0; v < g->rcsr.n; v++) {
            b->u.off[vb + v + 1] = eb + g->rcsr.off[v + 1];
            b->ostart[vb + v] = eb + g->ostart[v];
        }
        for (int64_t e = 0; e < g->rcsr.nnz(); e++) b->u.adj[eb + e] = vb + g->rcsr.adj[e];
        b->u.maxdeg = std::max(b->u.maxdeg, g->rcsr.maxdeg);
        b->max_out = std::max(b->max_out, g->max_out);
        vb += g->rcsr.n;
        eb += g->rcsr.nnz();
    }
    *out = b;
    return DM_OK;
}

int dm_batch_run(dm_batch *b, const dm_query *queries, size_t nq, const dm_config *cfg, uint64_t *results,
                 size_t results_len, uint64_t *result_offsets, double *graph_latency_ms) {
    if (!b || !queries || nq == 0) return set_error(DM_ERR_ARG, "need batch and queries");
    dm_config c;
    if (cfg) c = *cfg;
    else dm_config_init(&c);
    std::vector<dm_query> qs(queries, queries + nq);
    return batch_run(b, qs, c, results, results_len, result_offsets, graph_latency_ms);
}

// Tunes width and steal granularity on a sample of the queries: every S-th
// query in descending order of estimated cost, S doubled until one trial
// takes at most a twelfth of the budget, then as many of the candidates as
// the budget allows.
int dm_batch_autotune(dm_batch *b, const dm_query *queries, size_t nq, double budget_ms, dm_config *best) {
    if (!b || !queries || nq == 0) return set_error(DM_ERR_ARG, "need batch and queries");
    std::vector<dm_query> all(queries, queries + nq);
    for (const dm_query &Q : all)
        if (Q.graph < 0 || Q.graph >= b->ngraphs) return set_error(DM_ERR_ARG, "query graph %d out of range", Q.graph);
    std::vector<double> est(nq, 0.0);
    for (size_t i = 0; i < nq; i++) {
        const dm_query &Q = all[i];
        const int32_t v0 = b->vbase[Q.graph], v1 = v0 + b->gn[Q.graph];
        for (int32_t v = v0; v < v1; v++) {
            const double d = Q.motif ? (double)(b->u.off[v + 1] - b->u.off[v]) : (double)(b->u.off[v + 1] - b->ostart[v]);
            est[i] += std::pow(d, std::max(1, Q.k - 2));
        }
    }
    std::vector<size_t> ord(nq);
    for (size_t i = 0; i < nq; i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t c) { return est[a] > est[c]; });
    auto sample = [&](size_t S) {
        std::vector<dm_query> q;
        for (size_t i = S / 2; i < nq; i += S) q.push_back(all[ord[i]]);
        if (q.empty()) q.push_back(all[ord[nq - 1]]);
        return q;
    };
    dm_config base;
    dm_config_init(&base);
    base.vwave = 64;
    base.steal_min = 2;
    auto trial = [&](const std::vector<dm_query> &q, const dm_config &c, double &ms) -> int {
        int r = batch_run(b, q, c, nullptr, 0, nullptr, nullptr);  // capture (and a first run)
        if (r) return r;
        r = batch_run(b, q, c, nullptr, 0, nullptr, nullptr);  // replay timing
        ms = b->last.device_ms;
        return r;
    };
    const double t_start = now_ms();
    size_t S = std::max<size_t>(1, (nq + 47) / 48);
    std::vector<dm_query> qs = sample(S);
    double t_best = 0;
    int r = trial(qs, base, t_best);
    if (r) return r;
    while (t_best > budget_ms / 12 && qs.size() > 4 && now_ms() - t_start < budget_ms) {
        S *= 2;
        qs = sample(S);
        if ((r = trial(qs, base, t_best))) return r;
    }
    dm_config bc = base;
    for (int w : {64, 32, 16}) {
        for (int sm : {2, 8}) {
            if (w == 64 && sm == 2) continue;
            if (now_ms() - t_start > budget_ms) break;
            dm_config c = base;
            c.vwave = w;
            c.steal_min = sm;
            double t = 0;
            if ((r = trial(qs, c, t))) return r;
            if (t < t_best) {
                t_best = t;
                bc = c;
            }
        }
    }
    bc.lds_bytes = 0;  // each phase: the largest stack its occupancy allows
    b->plan.vwave = bc.vwave;
    b->tuned_steal_min = bc.steal_min;
    if (best) *best = bc;
    return DM_OK;
}

int dm_batch_last_stats(const dm_batch *b, dm_stats *st) {
    if (!b || !st) return set_error(DM_ERR_ARG, "NULL argument");
    *st = b->last;
    return DM_OK;
}

int dm_batch_num_results(const dm_batch *b, const dm_query *queries, size_t nq, size_t *nres) {
    if (!b || !queries || !nres) return set_error(DM_ERR_ARG, "NULL argument");
    size_t t = 0;
    for (size_t i = 0; i < nq; i++) {
        if (queries[i].motif) {
            const PatternTable *p = pattern_table(queries[i].k);
            if (!p) return last_code();
            t += (size_t)p->ncls;
        } else {
            t += 1;
        }
    }
    *nres = t;
    return DM_OK;
}

void dm_batch_free(dm_batch *b) {
    if (!b) return;
    drop_graph(b);
    void *bufs[] = {b->d_arena, b->d_sq_off, b->d_sq_jobs, b->d_sq_cur, b->d_qmask, b->d_results,
                    b->d_qd, b->d_done, b->d_start, b->d_sc};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    free_device_graph(b->dev);
    delete b;
}

}  // extern "C"
