// Host runtime and C ABI.
//
// Stands in for the reference's host class DuMatoCPU (src/main/DuMatoCPU.{h,cu}):
// graph load -> CSR (:94-108) -> root partition (:121-135) -> device buffers
// (:204-276) -> launch (:316-319) -> result reduction (:360-384, :419-423).
// The host-driven stop/snapshot/donate load balancer of the reference
// (:386-562) has no host part here: donation runs inside the kernel.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <tuple>

#include "internal.h"
#include "runtime.h"

namespace dm {

static thread_local int t_code = DM_OK;
static thread_local char t_msg[1024] = "";

int set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_msg, sizeof(t_msg), fmt, ap);
    va_end(ap);
    t_code = code;
    return code;
}
int last_code() { return t_code; }

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) return set_error(DM_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
int graph_finish(dm_graph *g) {
    HostCSR &c = g->csr;
    const int32_t n = c.n;
    g->perm.assign(n, 0);
    g->inv.assign(n, 0);
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        int64_t da = c.off[a + 1] - c.off[a], db = c.off[b + 1] - c.off[b];
        return da != db ? da < db : a < b;
    });
    for (int32_t r = 0; r < n; r++) {
        g->perm[order[r]] = r;
        g->inv[r] = order[r];
    }
    HostCSR &rc = g->rcsr;
    rc.n = n;
    rc.maxdeg = c.maxdeg;
    rc.off.assign((size_t)n + 1, 0);
    rc.adj.resize(c.adj.size());
    g->ostart.assign(n, 0);
    int32_t max_out = 0;
    for (int32_t r = 0; r < n; r++) {
        int32_t v = order[r];
        int64_t b = rc.off[r];
        int64_t d = c.off[v + 1] - c.off[v];
        for (int64_t i = 0; i < d; i++) rc.adj[b + i] = g->perm[c.adj[c.off[v] + i]];
        std::sort(rc.adj.begin() + b, rc.adj.begin() + b + d);
        rc.off[r + 1] = b + d;
        const int64_t up = std::upper_bound(rc.adj.begin() + b, rc.adj.begin() + b + d, r) - rc.adj.begin();
        g->ostart[r] = up;
        max_out = std::max<int32_t>(max_out, (int32_t)(b + d - up));
    }
    g->max_out = max_out;
    return DM_OK;
}

// ---------------------------------------------------------------------------
// device contexts (one per GPU, kept until the process ends)
static std::mutex g_ctx_mu;
static std::map<int, std::unique_ptr<DeviceCtx>> g_ctx;

int device_ctx(DeviceCtx **out) {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) return set_error(DM_ERR_HIP, "no HIP device available");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    auto it = g_ctx.find(dev);
    if (it != g_ctx.end()) {
        *out = it->second.get();
        return DM_OK;
    }
    std::unique_ptr<DeviceCtx> c(new DeviceCtx());
    c->device = dev;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    c->num_cu = prop.multiProcessorCount;
    int rate = 0;
    if (hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, dev) != hipSuccess || rate <= 0) rate = 100000;
    c->wall_khz = rate;
    snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    // donated-job queue: 2^20 slots; epoch 0 never matches a launch
    const size_t qcap = (size_t)1 << 20;
    int r = dev_grow(c->d_q, c->q_cap_alloc, qcap);
    if (r) return r;
    HIPCHK(hipMemset(c->d_q, 0, qcap * sizeof(Job)));
    size_t cap = 0;
    if ((r = dev_grow(c->d_epoch, cap, 1))) return r;
    HIPCHK(hipMemset(c->d_epoch, 0, sizeof(int32_t)));
    cap = 0;
    if ((r = dev_grow(c->d_sc, cap, 1))) return r;
    *out = c.get();
    g_ctx[dev] = std::move(c);
    return DM_OK;
}

int device_lut(DeviceCtx *c, int K, const uint16_t **out) {
    const PatternTable *tab = pattern_table(K);
    if (!tab) return last_code();
    if (!c->d_lut[K]) {
        uint16_t *p = nullptr;
        size_t cap = 0;
        int r = dev_grow(p, cap, tab->q2c.size());
        if (r) return r;
        hipError_t e = hipMemcpy(p, tab->q2c.data(), tab->q2c.size() * 2, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(p);
            return set_error(DM_ERR_HIP, "class table upload: %s", hipGetErrorString(e));
        }
        c->d_lut[K] = p;
    }
    *out = c->d_lut[K];
    return DM_OK;
}

void free_device_graph(DeviceGraph *d) {
    if (!d) return;
    auto F = [](void *p) {
        if (p) (void)hipFree(p);
    };
    F(d->d_off);
    F(d->d_adj);
    F(d->d_ostart);
    F(d->d_root_cost);
    F(d->d_root_count);
    delete d;
}

int upload_csr(const HostCSR &rc, const std::vector<int64_t> &ostart, DeviceGraph *d) {
    size_t cap = 0;
    int r = dev_grow(d->d_off, cap, (size_t)rc.n + 1);
    if (r) return r;
    cap = 0;
    if ((r = dev_grow(d->d_adj, cap, std::max<size_t>(rc.adj.size(), 1)))) return r;
    cap = 0;
    if ((r = dev_grow(d->d_ostart, cap, std::max<size_t>(ostart.size(), 1)))) return r;
    HIPCHK(hipMemcpy(d->d_off, rc.off.data(), rc.off.size() * 8, hipMemcpyHostToDevice));
    if (!rc.adj.empty()) HIPCHK(hipMemcpy(d->d_adj, rc.adj.data(), rc.adj.size() * 4, hipMemcpyHostToDevice));
    if (!ostart.empty()) HIPCHK(hipMemcpy(d->d_ostart, ostart.data(), ostart.size() * 8, hipMemcpyHostToDevice));
    return DM_OK;
}

int ensure_device(dm_graph *g) {
    if (g->dev) return hipSetDevice(g->dev->ctx->device) == hipSuccess ? DM_OK : set_error(DM_ERR_HIP, "hipSetDevice");
    DeviceCtx *c = nullptr;
    int r = device_ctx(&c);
    if (r) return r;
    DeviceGraph *d = new DeviceGraph();
    d->ctx = c;
    if ((r = upload_csr(g->rcsr, g->ostart, d))) {
        free_device_graph(d);
        return r;
    }
    g->dev = d;
    return DM_OK;
}

// ---------------------------------------------------------------------------
// launch shape
int block_threads(const dm_config &cfg, int *threads) {
    const int w = cfg.waves_per_block == 0 ? 4 : cfg.waves_per_block;
    if (w != 1 && w != 2 && w != 4) return set_error(DM_ERR_ARG, "waves_per_block must be 1, 2 or 4 (got %d)", w);
    *threads = 64 * w;
    return DM_OK;
}

static int max_lds_uncached(int vwave, int threads, int mode, int K, int ncls_single, int hist_slots);

// Largest per-worker LDS stack (a multiple of 16 B, at most 16 KiB) that
// keeps the occupancy the kernel's registers allow: the LDS of a CU (160
// KiB) divided over that many blocks, less each block's fixed records.
int max_lds_per_worker(int vwave, int threads, int mode, int K, int ncls_single, int hist_slots) {
    // the answer depends only on the arguments (and the one device): memoised
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, int, int, int>, int> memo;
    const auto key = std::make_tuple(vwave, threads, mode, K, ncls_single, hist_slots);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = memo.find(key);
        if (it != memo.end()) return it->second;
    }
    const int lds = max_lds_uncached(vwave, threads, mode, K, ncls_single, hist_slots);
    std::lock_guard<std::mutex> lk(mu);
    memo[key] = lds;
    return lds;
}

static int max_lds_uncached(int vwave, int threads, int mode, int K, int ncls_single, int hist_slots) {
    const int wpb = threads / vwave;
    int bpc = 0;
    const size_t fixed = enum_lds_bytes(vwave, threads, K, ncls_single, hist_slots, 0);
    if (enum_occupancy(vwave, mode, threads, fixed, &bpc) != DM_OK || bpc < 1) return 32768 / wpb;
    const size_t per_block = (size_t)160 * 1024 / (size_t)bpc;
    if (per_block <= fixed + 256) return std::min(32768 / wpb, 16384);
    // 16-byte steps: the bitmap path's rows and build hash at out-degree 256
    // need all but a few bytes of a worker's share at 3 blocks per CU
    int lds = std::min((int)(((per_block - fixed) / (size_t)wpb) & ~(size_t)15), 16384);
    // the LDS allocation granularity may cost a block: step down until the
    // occupancy holds
    for (; lds > 256; lds -= 16) {
        int b2 = 0;
        const size_t total = enum_lds_bytes(vwave, threads, K, ncls_single, hist_slots, lds);
        if (enum_occupancy(vwave, mode, threads, total, &b2) == DM_OK && b2 >= bpc) break;
    }
    return lds;
}

int pick_shape(dm_graph *g, int K, bool motif, const dm_config &cfg, Shape &sh) {
    DeviceCtx *c = g->dev->ctx;
    int vw = cfg.vwave;
    if (vw == 0) vw = autotuned_vwave(g, K, motif);
    if (vw != 8 && vw != 16 && vw != 32 && vw != 64)
        return set_error(DM_ERR_ARG, "vwave must be 8, 16, 32 or 64 (got %d)", vw);
    int threads = 256;
    int r = block_threads(cfg, &threads);
    if (r) return r;
    int ncls = 0;
    if (motif) ncls = pattern_table(K)->ncls;
    // a tuned LDS size belongs to the full-size block it was tuned with
    const int dl = max_lds_per_worker(vw, threads, launch_mode(motif, K), K, ncls, 0);
    int lds = cfg.lds_bytes > 0 ? cfg.lds_bytes : threads == 256 ? tuned_lds(g, K, motif, vw, dl) : dl;
    lds = (lds + 15) & ~15;
    size_t lds_total = enum_lds_bytes(vw, threads, K, ncls, 0, lds);
    int bpc = 0;
    if ((r = enum_occupancy(vw, launch_mode(motif, K), threads, lds_total, &bpc))) return r;
    if (bpc < 1) return set_error(DM_ERR_ARG, "launch does not fit one CU (LDS %zu B)", lds_total);
    sh.vwave = vw;
    sh.threads = threads;
    sh.lds_per_worker = lds;
    sh.lds_total = lds_total;
    sh.blocks = cfg.blocks > 0 ? cfg.blocks : bpc * c->num_cu;
    sh.workers = sh.blocks * (threads / vw);
    return DM_OK;
}

// per-worker global arena bound (see the allocator in enum_kernel.hip)
int64_t arena_bound_for(int32_t maxdeg, int32_t max_out_, int K, bool motif) {
    int64_t md = std::max<int64_t>(maxdeg, 1);
    int64_t b = 256;
    if (motif) {
        for (int l = 1; l <= K - 3; l++) b += 10 * (int64_t)(l + 1) * md + 64;
        b += 4 * md + (int64_t)(K - 2) * md + 64;
        // leaf-phase scratch (at most 64 slots, 512 counters, a 128-entry
        // lookup queue)
        b += 64 * 8 + 512 * 4 + 64 * 12 + 4 + 16 * 64 + 16;
    } else {
        int64_t mo = std::max<int64_t>(max_out_, 1);
        for (int l = 1; l <= K - 3; l++) b += 4 * mo + 16;
        // local bitmap of a root (out-degree <= 256): rows, out-degrees,
        // prefix, then the build's hash (512 keys + positions) and list starts
        b += 256 * 4 * 8 + 2 * 257 * 4 + 512 * 5 + 256 * 8 + 6 * 16;
        // k = 6: a round's pair staging (64 lanes x 4 words, 65 prefix entries)
        if (K == 6) b += 64 * 4 * 8 + 65 * 4 + 2 * 16;
    }
    return (b + 255) & ~(int64_t)255;
}

static int64_t arena_bound(const dm_graph *g, int K, bool motif) {
    return arena_bound_for(g->rcsr.maxdeg, g->max_out, K, motif);
}

int device_error(int code, const char *what) {
    const char *why = code == 1 ? "watchdog expired" : code == 2 ? "stack arena overflow"
                    : code == 4 ? "stopped" : "invalid job";
    return set_error(DM_ERR_DEVICE, "%s aborted: %s", what, why);
}

int check_config(const dm_config &cfg) {
    if (cfg.steal_min == 1 || cfg.steal_min < 0 || cfg.steal_period < 0)
        return set_error(DM_ERR_ARG, "steal_min must be 0 (tuned) or >= 2; steal_period >= 0");
    int t = 0;
    return block_threads(cfg, &t);
}

// ---------------------------------------------------------------------------
// static partition
// A root without a neighbour above it has no subgraph and no DFS node below
// it.  Roots with fewer than k-1 such neighbours still run, so that the
// algorithmic-byte count of cliques covers every DFS node.
bool root_has_work(const dm_graph *g, int32_t v, int K, bool motif) {
    (void)K;
    (void)motif;
    return g->rcsr.off[v + 1] - g->ostart[v] > 0;
}

// Plain: one job per root, in descending order of a work estimate, dealt
// round-robin to the workers like the reference's v -> warp v % W partition
// (src/main/DuMatoCPU.cu:121-135), heaviest first.  The estimate is the
// out-degree for cliques.  For motifs it is the sum of the degrees of the
// root's higher-ranked neighbours, on a log scale (16 steps per octave): a
// root's subgraphs grow out through those neighbours, and their degrees set
// the leaf work, so a modest root beside hubs is heavy.  (Chung-Lu 18K k = 4:
// rank correlation with measured per-root cycles 0.61 for the degree, 0.88
// for this sum; ordered by degree such roots started late and ran into the
// launch's last fifth.)  Grouping light roots into range jobs measured
// slower: fewer static jobs end the static phase early and push the work
// through donations.
void plain_partition(const dm_graph *g, const std::vector<int32_t> &roots, int K, bool motif,
                     int workers, std::vector<int64_t> &sq_off, std::vector<int32_t> &sq_jobs) {
    (void)K;
    // descending key by a counting sort (equal keys keep their ascending
    // rank), then job t to worker t mod workers
    const size_t nr = roots.size(), W = (size_t)std::max(workers, 1);
    // scratch keeps its capacity between calls (a fresh allocation's
    // first-touch page faults cost as much as the sort)
    static thread_local std::vector<int32_t> key, order;
    static thread_local std::vector<int64_t> start;
    key.resize(nr);
    int32_t kmax = 0;
    const int64_t *off = g->rcsr.off.data();
    const int32_t *adj = g->rcsr.adj.data();
    for (size_t i = 0; i < nr; i++) {
        const int32_t v = roots[i];
        if (motif) {
            int64_t s = 0;
            for (int64_t e = g->ostart[v]; e < off[v + 1]; e++) s += off[adj[e] + 1] - off[adj[e]];
            key[i] = (int32_t)(16.0 * std::log2(1.0 + (double)s));
        } else {
            key[i] = (int32_t)(off[v + 1] - g->ostart[v]);
        }
        kmax = std::max(kmax, key[i]);
    }
    start.assign((size_t)kmax + 2, 0);
    for (size_t i = 0; i < nr; i++) start[(size_t)(kmax - key[i]) + 1]++;
    for (size_t b = 1; b < start.size(); b++) start[b] += start[b - 1];
    order.resize(nr);
    for (size_t i = 0; i < nr; i++) order[start[(size_t)(kmax - key[i])]++] = (int32_t)i;
    sq_off.assign(W + 1, 0);
    for (size_t w = 0; w < W; w++) sq_off[w + 1] = sq_off[w] + (int64_t)(nr > w ? (nr - 1 - w) / W + 1 : 0);
    sq_jobs.resize(nr * 4);
    // job t goes to worker t % W, slot t / W: kept as counters, no division
    size_t w = 0, slot = 0;
    for (size_t t = 0; t < nr; t++) {
        int32_t *jp = &sq_jobs[4 * (size_t)(sq_off[w] + (int64_t)slot)];
        jp[0] = roots[order[t]];
        jp[1] = 0;
        jp[2] = INT_MAX;
        jp[3] = 0;
        if (++w == W) {
            w = 0;
            slot++;
        }
    }
}

// ---------------------------------------------------------------------------
int run_enum(dm_graph *g, int K, bool motif, const dm_config &cfg, uint64_t *hist, size_t hist_len,
             uint64_t *count_out, int sample_mod) {
    const double t_wall0 = now_ms();
    const PatternTable *tab = nullptr;
    if (motif) {
        if (K < 3 || K > kMaxMotifK)
            return set_error(DM_ERR_UNSUPPORTED, "motif size k=%d unsupported (3..%d)", K, kMaxMotifK);
        tab = pattern_table(K);
        if (!tab) return last_code();
        if (hist && hist_len < (size_t)tab->ncls)
            return set_error(DM_ERR_ARG, "histogram needs %d slots (got %zu)", tab->ncls, hist_len);
    } else if (K < 2 || K > kMaxK) {
        return set_error(DM_ERR_ARG, "clique size k=%d unsupported (2..%d)", K, kMaxK);
    }
    int r = check_config(cfg);
    if (r) return r;
    if ((r = ensure_device(g))) return r;
    DeviceGraph *dg = g->dev;
    DeviceCtx *d = dg->ctx;
    std::lock_guard<std::recursive_mutex> lock(d->mu);
    const int32_t n = g->rcsr.n;

    Shape sh;
    if ((r = pick_shape(g, K, motif, cfg, sh))) return r;

    // roots and their static partition (reused while nothing it depends on
    // changes)
    const double t_s0 = now_ms();
    g->last_infer_ms = g->last_bin_ms = 0;
    int32_t rb = std::max(cfg.root_begin, 0), re = cfg.root_end > 0 ? std::min(cfg.root_end, n) : n;
    const int learned = cfg.sched == DM_SCHED_LEARNED ? 1 : 0;
    // the bitmap switch is part of the key: a learned partition gives heavy
    // bitmap roots shared-cursor jobs, which need their prebuilt bitmaps
    const int key[8] = {K, motif ? 1 : 0, learned, sh.workers, rb, re, sample_mod, cfg.clique_bitmap ? 1 : 0};
    dm_graph::Partition *part = nullptr;
    for (size_t i = 0; i < g->partitions.size() && !cfg.repartition; i++) {
        dm_graph::Partition &c = g->partitions[i];
        if (std::equal(key, key + 8, c.key) && (!learned || c.model_gen == g->model_gen)) {
            if (i) std::swap(g->partitions[0], c);
            part = &g->partitions[0];
            break;
        }
    }
    if (!part) {
        const int rkey[4] = {rb, re, sample_mod, motif ? 1 : 0};
        std::shared_ptr<const dm_graph::RootList> rl;
        for (const auto &c : g->root_lists)
            if (std::equal(rkey, rkey + 4, c->key)) rl = c;
        if (!rl) {
            auto nl = std::make_shared<dm_graph::RootList>();
            std::copy(rkey, rkey + 4, nl->key);
            nl->roots.reserve(n);
            nl->skipped_bytes = 0;  // motif root nodes with an empty extension set
            for (int32_t v = 0; v < n; v++) {
                int32_t orig = g->inv[v];
                if (orig < rb || orig >= re) continue;
                if (sample_mod > 1 && v % sample_mod != 0) continue;
                if (root_has_work(g, v, K, motif)) nl->roots.push_back(v);
                else if (motif) nl->skipped_bytes += 4ull * (uint64_t)(g->rcsr.off[v + 1] - g->rcsr.off[v]);
            }
            if (g->root_lists.size() >= 4) g->root_lists.pop_back();
            g->root_lists.insert(g->root_lists.begin(), nl);
            rl = nl;
        }
        const std::vector<int32_t> &roots = rl->roots;
        const uint64_t skipped_bytes = rl->skipped_bytes;
        dm_graph::Partition np;
        // a partition recomputed for the same key takes over the old one's
        // arrays: their capacity spares the new ones first-touch page faults
        for (dm_graph::Partition &c : g->partitions)
            if (std::equal(key, key + 8, c.key)) {
                np.sq_off.swap(c.sq_off);
                np.sq_jobs.swap(c.sq_jobs);
                np.pre_roots.swap(c.pre_roots);
                np.pre_offs.swap(c.pre_offs);
                break;
            }
        std::copy(key, key + 8, np.key);
        static std::atomic<uint64_t> part_serial{0};
        np.model_gen = g->model_gen;
        np.serial = ++part_serial;
        np.skipped_bytes = skipped_bytes;
        if (learned) {
            if ((r = learned_partition(g, roots, K, motif, cfg, sh, np))) return r;
        } else {
            plain_partition(g, roots, K, motif, sh.workers, np.sq_off, np.sq_jobs);
        }
        g->partitions.erase(std::remove_if(g->partitions.begin(), g->partitions.end(),
                                           [&](const dm_graph::Partition &c) {
                                               return std::equal(key, key + 8, c.key);
                                           }),
                            g->partitions.end());
        if (g->partitions.size() >= 8) g->partitions.pop_back();
        g->partitions.insert(g->partitions.begin(), std::move(np));
        part = &g->partitions[0];
    }
    const std::vector<int64_t> &sq_off = part->sq_off;
    const std::vector<int32_t> &sq_jobs = part->sq_jobs;
    const uint64_t skipped_bytes = part->skipped_bytes;
    const int npre = cfg.clique_bitmap ? (int)part->pre_roots.size() : 0;
    const double t_s1 = now_ms();
    const int64_t njobs = (int64_t)sq_jobs.size() / 4;

    // device workspaces
    size_t need_arena = (size_t)arena_bound(g, K, motif) * (size_t)sh.workers;
    if ((r = dev_grow(d->d_arena, d->arena_cap, need_arena))) return r;
    const void *const old_bufs[4] = {d->d_sq_off, d->d_sq_jobs, d->d_pre_roots, d->d_pre_offs};
    if ((r = dev_grow(d->d_sq_off, d->sq_off_cap, sq_off.size()))) return r;
    if ((r = dev_grow(d->d_sq_jobs, d->sq_jobs_cap, std::max<size_t>(sq_jobs.size(), 4)))) return r;
    if ((r = dev_grow(d->d_sq_cur, d->sq_cur_cap, (size_t)sh.workers))) return r;
    if ((r = dev_grow(d->d_qmask, d->qmask_cap, (size_t)qmask_words(sh.workers, part->group)))) return r;
    const size_t ncls = tab ? (size_t)tab->ncls : 0;
    const size_t nres = motif ? ncls : 1;
    if ((r = dev_grow(d->d_results, d->results_cap, std::max<size_t>(nres, 1)))) return r;
    const uint16_t *lut = nullptr;
    if (tab && (r = device_lut(d, K, &lut))) return r;
    const bool want_cost = cfg.record_costs != 0, want_cnt = cfg.per_root != 0;
    if (want_cost && (r = dev_grow(dg->d_root_cost, dg->root_cost_cap, (size_t)n))) return r;
    if (want_cnt && (r = dev_grow(dg->d_root_count, dg->root_count_cap, (size_t)n))) return r;
    QueryDesc qd;
    memset(&qd, 0, sizeof(qd));
    qd.K = K;
    qd.motif = motif ? 1 : 0;
    qd.ncls = (int32_t)ncls;
    qd.lut = lut;
    qd.out = d->d_results;
    if ((r = dev_grow(d->d_qd, d->qd_cap, 1))) return r;
    if (npre > 0) {
        if ((r = dev_grow(d->d_pre_roots, d->pre_roots_cap, (size_t)npre)) ||
            (r = dev_grow(d->d_pre_offs, d->pre_offs_cap, (size_t)npre)) ||
            (r = dev_grow(d->d_pre_cursor, d->pre_cursor_cap, (size_t)npre)) ||
            (r = dev_grow(d->d_pre, d->pre_cap, (size_t)part->pre_bytes)))
            return r;
    }

    // pinned staging for the counters and results read back after the launch
    const size_t stage_need = sizeof(SchedCounters) + std::max<size_t>(nres, 1) * 8;
    if (stage_need > d->stage_cap) {
        if (d->h_stage) (void)hipHostFree(d->h_stage);
        d->h_stage = nullptr;
        d->stage_cap = 0;
        HIPCHK(hipHostMalloc((void **)&d->h_stage, stage_need, hipHostMallocDefault));
        d->stage_cap = stage_need;
    }

    hipStream_t s = d->stream;
    // the queues (and prebuilt-root list) go up once per partition: a
    // repeated query finds them in place unless another partition, or a
    // regrown buffer, has replaced them since
    const void *const new_bufs[4] = {d->d_sq_off, d->d_sq_jobs, d->d_pre_roots, d->d_pre_offs};
    if (!std::equal(old_bufs, old_bufs + 4, new_bufs)) d->up_serial = 0;
    if (d->up_serial != part->serial) {
        d->up_serial = 0;
        if (npre > 0) {
            HIPCHK(hipMemcpyAsync(d->d_pre_roots, part->pre_roots.data(), (size_t)npre * 4, hipMemcpyHostToDevice,
                                  s));
            HIPCHK(hipMemcpyAsync(d->d_pre_offs, part->pre_offs.data(), (size_t)npre * 8, hipMemcpyHostToDevice, s));
        }
        HIPCHK(hipMemcpyAsync(d->d_sq_off, sq_off.data(), sq_off.size() * 8, hipMemcpyHostToDevice, s));
        if (!sq_jobs.empty())
            HIPCHK(hipMemcpyAsync(d->d_sq_jobs, sq_jobs.data(), sq_jobs.size() * 4, hipMemcpyHostToDevice, s));
        d->up_serial = part->serial;
    }
    if (npre > 0 && (r = fill_u32_launch(d->d_pre_cursor, 0, (size_t)npre, s))) return r;
    HIPCHK(hipMemcpyAsync(d->d_qd, &qd, sizeof(qd), hipMemcpyHostToDevice, s));
    if (want_cost && (r = fill_u32_launch(dg->d_root_cost, 0, (size_t)n * 2, s))) return r;
    if (want_cnt && (r = fill_u32_launch(dg->d_root_count, 0, (size_t)n * 2, s))) return r;

    EnumParams p;
    memset(&p, 0, sizeof(p));
    p.off = dg->d_off;
    p.adj = dg->d_adj;
    p.ostart = dg->d_ostart;
    p.n = n;
    p.nq = 1;
    p.qd = d->d_qd;
    p.kmax = K;
    p.hist_slots = 0;
    p.nworkers = sh.workers;
    p.sq_off = d->d_sq_off;
    p.sq_jobs = d->d_sq_jobs;
    p.sq_cur = d->d_sq_cur;
    p.q = d->d_q;
    p.q_cap = (uint32_t)d->q_cap_alloc;
    p.epoch = d->d_epoch;
    p.start_ticks = nullptr;
    p.sc = d->d_sc;
    p.arena = (char *)d->d_arena;
    p.arena_per_worker = arena_bound(g, K, motif);
    p.lds_per_worker = sh.lds_per_worker;
    p.root_cost = want_cost ? dg->d_root_cost : nullptr;
    p.root_count = want_cnt ? dg->d_root_count : nullptr;
    p.steal = cfg.steal ? 1 : 0;
    p.steal_min = cfg.steal_min ? cfg.steal_min : tuned_steal_min(g, K, motif);
    p.poll_ticks = (uint64_t)cfg.steal_period * (uint64_t)d->wall_khz / 1000;
    p.bitmap = cfg.clique_bitmap ? 1 : 0;
    p.watchdog_ticks = cfg.watchdog_s > 0 ? (uint64_t)(cfg.watchdog_s * d->wall_khz * 1000.0) : 0;
    p.npre = npre;
    p.pre_roots = npre ? d->d_pre_roots : nullptr;
    p.pre_offs = npre ? d->d_pre_offs : nullptr;
    p.pre = npre ? d->d_pre : nullptr;
    p.pre_cursor = npre ? d->d_pre_cursor : nullptr;
    p.q_group = part->group;
    // mask-guided steals for motifs only (motif k=4 on Chung-Lu 18K: 4.74 ->
    // 4.65 ms).  A clique launch's makespan is a few long jobs, which run
    // faster beside idle workers that sleep than beside thieves that always
    // find a light job (R-MAT s17 k=5: 1.43 -> 1.54 ms with the mask).
    p.qmask = motif ? d->d_qmask : nullptr;
    int queues = 0;
    for (size_t w = 0; w + 1 < sq_off.size(); w++) queues += sq_off[w + 1] > sq_off[w] ? 1 : 0;
    if ((r = enum_init_launch(p, (unsigned long long)njobs, queues, d->d_epoch, d->d_results, nres, nullptr, 0, s)))
        return r;

    // the kernel time includes the bitmap pre-pass of a learned partition
    HIPCHK(hipEventRecord(d->ev0, s));
    if (npre > 0 && (r = prebuild_launch(p, s))) return r;
    if (njobs > 0 && (r = enum_launch(p, launch_mode(motif, K), sh.vwave, sh.blocks, sh.threads, sh.lds_total, s)))
        return r;
    HIPCHK(hipEventRecord(d->ev1, s));
    HIPCHK(hipMemcpyAsync(d->h_stage, d->d_sc, sizeof(SchedCounters), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(d->h_stage + sizeof(SchedCounters), d->d_results, nres * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    SchedCounters sc1;
    memcpy(&sc1, d->h_stage, sizeof(sc1));
    std::vector<unsigned long long> h(nres);
    if (nres) memcpy(h.data(), d->h_stage + sizeof(SchedCounters), nres * 8);
    float kms = 0;
    HIPCHK(hipEventElapsedTime(&kms, d->ev0, d->ev1));
    if (sc1.error) return device_error(sc1.error, "enumeration kernel");
    uint64_t total = 0;
    if (motif) {
        for (size_t i = 0; i < ncls; i++) total += h[i];
        if (hist)
            for (size_t i = 0; i < ncls; i++) hist[i] = h[i];
    } else {
        total = h[0];
    }
    if (count_out) *count_out = total;

    // per-root records back in original ids
    if (want_cost) {
        std::vector<unsigned long long> tmp(n);
        HIPCHK(hipMemcpy(tmp.data(), dg->d_root_cost, (size_t)n * 8, hipMemcpyDeviceToHost));
        g->last_costs.assign(n, 0);
        for (int32_t v = 0; v < n; v++) g->last_costs[g->inv[v]] = tmp[v];
    } else {
        g->last_costs.clear();
    }
    if (want_cnt) {
        std::vector<unsigned long long> tmp(n);
        HIPCHK(hipMemcpy(tmp.data(), dg->d_root_count, (size_t)n * 8, hipMemcpyDeviceToHost));
        g->last_counts.assign(n, 0);
        for (int32_t v = 0; v < n; v++) g->last_counts[g->inv[v]] = tmp[v];
    } else {
        g->last_counts.clear();
    }

    dm_stats &st = g->last;
    memset(&st, 0, sizeof(st));
    st.device_ms = kms;
    st.sched_ms = t_s1 - t_s0;
    st.infer_ms = g->last_infer_ms;
    st.bin_ms = g->last_bin_ms;
    st.wall_ms = now_ms() - t_wall0;
    st.subgraphs = total;
    st.alg_bytes = sc1.alg_bytes + skipped_bytes;
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
    st.vwave = sh.vwave;
    st.workers = sh.workers;
    st.blocks = sh.blocks;
    st.lds_bytes = sh.lds_per_worker;
    return DM_OK;
}

}  // namespace dm

using namespace dm;

// ===========================================================================
// C ABI
extern "C" {

int dm_version(void) { return 2; }

const char *dm_last_error(void) { return t_msg; }

int dm_config_init(dm_config *cfg) {
    if (!cfg) return set_error(DM_ERR_ARG, "cfg is NULL");
    memset(cfg, 0, sizeof(*cfg));
    cfg->sched = DM_SCHED_PLAIN;
    cfg->vwave = 0;
    cfg->steal = 1;
    cfg->steal_min = 0;
    cfg->steal_period = 2;
    cfg->lds_bytes = 0;
    cfg->waves_per_block = 4;
    cfg->blocks = 0;
    cfg->clique_bitmap = 1;
    cfg->watchdog_s = 120.0;
    return DM_OK;
}

static int make_graph(std::vector<int32_t> &s, std::vector<int32_t> &d, int32_t n, dm_graph **out) {
    if (!out) return set_error(DM_ERR_ARG, "out is NULL");
    *out = nullptr;
    dm_graph *g = new dm_graph();
    int r = build_csr(n, s.data(), d.data(), s.size(), g->csr);
    if (r == DM_OK) r = graph_finish(g);
    if (r) {
        delete g;
        return r;
    }
    *out = g;
    return DM_OK;
}

int dm_graph_load_edgelist(const char *path, dm_graph **out) {
    if (!path) return set_error(DM_ERR_ARG, "path is NULL");
    std::vector<int32_t> s, d;
    int r = parse_edgelist(path, s, d);
    if (r) return r;
    return make_graph(s, d, 0, out);
}

int dm_graph_from_edges(const int32_t *src, const int32_t *dst, size_t m, int32_t n, dm_graph **out) {
    if (m && (!src || !dst)) return set_error(DM_ERR_ARG, "edge arrays are NULL");
    std::vector<int32_t> s(src, src + m), d(dst, dst + m);
    return make_graph(s, d, n, out);
}

int dm_graph_info(const dm_graph *g, int32_t *n, int64_t *m, int32_t *maxdeg) {
    if (!g) return set_error(DM_ERR_ARG, "graph is NULL");
    if (n) *n = g->csr.n;
    if (m) *m = g->csr.nnz() / 2;
    if (maxdeg) *maxdeg = g->csr.maxdeg;
    return DM_OK;
}

int dm_graph_csr(const dm_graph *g, int64_t *off, size_t off_len, int32_t *adj, size_t adj_len) {
    if (!g) return set_error(DM_ERR_ARG, "graph is NULL");
    if (off) {
        if (off_len < g->csr.off.size()) return set_error(DM_ERR_ARG, "off needs %zu entries", g->csr.off.size());
        memcpy(off, g->csr.off.data(), g->csr.off.size() * 8);
    }
    if (adj) {
        if (adj_len < g->csr.adj.size()) return set_error(DM_ERR_ARG, "adj needs %zu entries", g->csr.adj.size());
        memcpy(adj, g->csr.adj.data(), g->csr.adj.size() * 4);
    }
    return DM_OK;
}

void dm_graph_free(dm_graph *g) {
    if (!g) return;
    free_device_graph(g->dev);
    free_model(g);
    delete g;
}

int dm_gen_erdos_renyi(int32_t n, int64_t m, uint64_t seed, dm_graph **out) {
    std::vector<int32_t> s, d;
    int r = gen_erdos_renyi(n, m, seed, s, d);
    return r ? r : make_graph(s, d, n, out);
}

int dm_gen_chung_lu(int32_t n, int64_t m_target, double gamma, double wmax, uint64_t seed, dm_graph **out) {
    std::vector<int32_t> s, d;
    int r = gen_chung_lu(n, m_target, gamma, wmax, seed, s, d);
    return r ? r : make_graph(s, d, n, out);
}

int dm_gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed, dm_graph **out) {
    std::vector<int32_t> s, d;
    int r = gen_rmat(scale, ef, a, b, c, seed, s, d);
    return r ? r : make_graph(s, d, (int32_t)1 << scale, out);
}

int dm_gen_barabasi_albert(int32_t n, int32_t m, int32_t planted, int32_t kmin, int32_t kmax, uint64_t seed,
                           dm_graph **out) {
    std::vector<int32_t> s, d;
    int r = gen_barabasi_albert(n, m, planted, kmin, kmax, seed, s, d);
    return r ? r : make_graph(s, d, n, out);
}

int dm_clique_count(dm_graph *g, int k, const dm_config *cfg, uint64_t *out) {
    if (!g || !out) return set_error(DM_ERR_ARG, "graph or out is NULL");
    dm_config c;
    if (cfg) c = *cfg;
    else dm_config_init(&c);
    return run_enum(g, k, false, c, nullptr, 0, out, 1);
}

int dm_motif_count(dm_graph *g, int k, const dm_config *cfg, uint64_t *hist, size_t hist_len, uint64_t *total) {
    if (!g || !hist) return set_error(DM_ERR_ARG, "graph or hist is NULL");
    dm_config c;
    if (cfg) c = *cfg;
    else dm_config_init(&c);
    return run_enum(g, k, true, c, hist, hist_len, total, 1);
}

int dm_last_stats(const dm_graph *g, dm_stats *st) {
    if (!g || !st) return set_error(DM_ERR_ARG, "NULL argument");
    *st = g->last;
    return DM_OK;
}

int dm_last_root_costs(const dm_graph *g, uint64_t *cycles, size_t n) {
    if (!g || !cycles) return set_error(DM_ERR_ARG, "NULL argument");
    if (g->last_costs.empty()) return set_error(DM_ERR_ARG, "last call did not record costs");
    if (n < g->last_costs.size()) return set_error(DM_ERR_ARG, "need %zu entries", g->last_costs.size());
    memcpy(cycles, g->last_costs.data(), g->last_costs.size() * 8);
    return DM_OK;
}

int dm_last_root_counts(const dm_graph *g, uint64_t *counts, size_t n) {
    if (!g || !counts) return set_error(DM_ERR_ARG, "NULL argument");
    if (g->last_counts.empty()) return set_error(DM_ERR_ARG, "last call did not record per-root counts");
    if (n < g->last_counts.size()) return set_error(DM_ERR_ARG, "need %zu entries", g->last_counts.size());
    memcpy(counts, g->last_counts.data(), g->last_counts.size() * 8);
    return DM_OK;
}

}  // extern "C"
