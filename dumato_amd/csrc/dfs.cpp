// Host side of the DFS-wide step API (include/dumato_device.h): per-worker
// state in device memory, the static partition, and the host-driven load
// balancer.
//
// Reference: DuMatoCPU (src/main/DuMatoCPU.cu): constructor and buffers
// (:13-30, 204-276), static partition v -> warp v % W (:121-135), runKernel /
// waitKernel, the idle poll gpuIsIdle (:386-405), stopKernel (:407-417),
// rebalance / organizeThreadStatus / donate (:419-562) with
// EnumerationHelper::getTargetLevel / getWeight / popLastExtension
// (src/main/EnumerationHelper.cpp:90-140), and the driver loop of
// src/main/motifs_DM_WCLB.cu:85-120.
//
// One deliberate difference: when a stop finds nothing to donate (every busy
// worker is on its last level, or only static jobs are left), the reference
// ends the run without relaunching (motifs_DM_WCLB.cu:95-105) and drops the
// stopped warps' remaining work; here the kernel is relaunched and the run
// continues without further balancing, so every subgraph is counted.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <queue>
#include <thread>
#include <vector>

#include "internal.h"
#include "runtime.h"

struct dm_dfs {
    dm_graph *g = nullptr;
    int K = 0;
    bool motif = false;
    dm_dfs_config cfg{};
    int vw = 32, blocks = 0, workers = 0, J = 16, ncls = 0;
    dm::DeviceCtx *ctx = nullptr;
    hipStream_t poll = nullptr;  // idle-count polls and stop requests while the kernel runs
    int32_t *pin = nullptr;      // pinned: [0] stop value to send, [1..workers] status snapshot
    dm_dfs_data d{};
    int32_t *d_stop = nullptr;
    std::vector<void *> bufs;    // device allocations
    // host snapshot
    std::vector<int32_t> k, status, id, next, jobs, job_level, job_valid, job_cur, ext;
    std::vector<uint64_t> ind, job_ind;
    dm_dfs_stats last{};
};

namespace dm {

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return set_error(DM_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));     \
    } while (0)

static double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

template <class T>
static int dfs_alloc(dm_dfs *s, T *&p, size_t n) {
    p = nullptr;
    hipError_t e = hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return set_error(DM_ERR_NOMEM, "step-API state (%zu B): %s", n * sizeof(T), hipGetErrorString(e));
    s->bufs.push_back((void *)p);
    return DM_OK;
}

static void dfs_release(dm_dfs *s) {
    for (void *p : s->bufs) (void)hipFree(p);
    s->bufs.clear();
    if (s->pin) (void)hipHostFree(s->pin);
    if (s->poll) (void)hipStreamDestroy(s->poll);
    s->pin = nullptr;
    s->poll = nullptr;
}

static size_t list_off(const dm_dfs *s, int l) {
    return s->motif ? (size_t)s->d.cap_base * (size_t)(l * (l + 1) / 2) : (size_t)s->d.cap_base * (size_t)l;
}

static bool host_adjacent(const HostCSR &c, int32_t u, int32_t x) {
    const int32_t *a = c.adj.data() + c.off[u];
    const int32_t *e = c.adj.data() + c.off[u + 1];
    const int32_t *p = std::lower_bound(a, e, x);
    return p != e && *p == x;
}

static int dfs_create(dm_graph *g, int K, int pattern, const dm_dfs_config &cfg, dm_dfs **out) {
    const bool motif = pattern == DM_PATTERN_MOTIF, nearq = pattern == DM_PATTERN_NEAR_CLIQUE;
    if (pattern < DM_PATTERN_CLIQUE || pattern > DM_PATTERN_NEAR_CLIQUE)
        return set_error(DM_ERR_ARG, "unknown pattern %d", pattern);
    if (motif) {
        if (K < 3 || K > kMaxMotifK) return set_error(DM_ERR_UNSUPPORTED, "motif size k=%d unsupported (3..%d)", K, kMaxMotifK);
    } else if (nearq) {
        if (K < 4 || K > 13) return set_error(DM_ERR_ARG, "near-clique query size k=%d unsupported (4..13)", K);
    } else if (K < 2 || K > DM_DFS_MAXK - 1) {
        return set_error(DM_ERR_ARG, "clique size k=%d unsupported (2..%d)", K, DM_DFS_MAXK - 1);
    }
    const int vw = cfg.vwave ? cfg.vwave : 32;
    if (vw != 8 && vw != 16 && vw != 32 && vw != 64) return set_error(DM_ERR_ARG, "vwave must be 8, 16, 32 or 64");
    if (cfg.jobs_per_worker < 0 || cfg.poll_us < 0 || cfg.stop_every_us < 0 || cfg.max_rebalances < 0 || cfg.blocks < 0)
        return set_error(DM_ERR_ARG, "negative step-API setting");
    const PatternTable *tab = motif ? pattern_table(K) : nullptr;
    if (motif && !tab) return last_code();
    int r = ensure_device(g);
    if (r) return r;
    dm_dfs *s = new dm_dfs();
    s->g = g;
    s->K = K;
    s->motif = motif;
    s->cfg = cfg;
    s->vw = vw;
    s->ctx = g->dev->ctx;
    // four 256-thread blocks per CU: 16 waves, within what these kernels'
    // registers allow; more blocks than stay resident would only wait
    s->blocks = cfg.blocks ? cfg.blocks : 4 * s->ctx->num_cu;
    s->workers = s->blocks * (256 / vw);
    const int32_t n = g->rcsr.n;
    s->J = std::max<int>(cfg.jobs_per_worker ? cfg.jobs_per_worker : 16, (n + s->workers - 1) / s->workers);
    s->ncls = motif ? tab->ncls : 0;
    // list capacity per level: a union of up to l+1 neighbourhoods (motif), a
    // neighbourhood (near-clique: the common neighbours of the clique), or
    // an out-neighbourhood (clique)
    const int32_t cap = motif || nearq ? std::max(g->rcsr.maxdeg, 1) : std::max(g->max_out, 1);
    const int64_t per = motif ? (int64_t)cap * K * (K - 1) / 2 : (int64_t)cap * (K - 1);
    const size_t W = (size_t)s->workers, WM = W * DM_DFS_MAXK, WJ = W * (size_t)s->J;
    dm_dfs_data &d = s->d;
    d.off = g->dev->d_off;
    d.adj = g->dev->d_adj;
    d.ostart = g->dev->d_ostart;
    d.n = n;
    d.K = K;
    d.motif = motif ? 1 : 0;
    d.workers = s->workers;
    d.vwave = vw;
    d.jobs_per_worker = s->J;
    d.cap_base = cap;
    d.ncls = s->ncls;
    d.ext_per_worker = per;
    if ((r = dfs_alloc(s, d.ext, W * (size_t)per)) || (motif && (r = dfs_alloc(s, d.src, W * (size_t)per))) ||
        (r = dfs_alloc(s, d.k, W)) || (r = dfs_alloc(s, d.status, W)) || (r = dfs_alloc(s, d.id, WM)) ||
        (r = dfs_alloc(s, d.next, WM)) || (r = dfs_alloc(s, d.fixed, WM)) || (r = dfs_alloc(s, d.ind, WM)) ||
        (r = dfs_alloc(s, d.jobs, WJ * DM_DFS_MAXK)) || (r = dfs_alloc(s, d.job_level, WJ)) ||
        (r = dfs_alloc(s, d.job_ind, WJ)) || (r = dfs_alloc(s, d.job_valid, W)) || (r = dfs_alloc(s, d.job_cur, W)) ||
        (r = dfs_alloc(s, d.result, (size_t)std::max(s->ncls, 1))) || (r = dfs_alloc(s, s->d_stop, 1))) {
        dfs_release(s);
        delete s;
        return r;
    }
    d.stop = s->d_stop;
    if (motif && (r = device_lut(s->ctx, K, &d.lut))) {
        dfs_release(s);
        delete s;
        return r;
    }
    if (hipStreamCreateWithFlags(&s->poll, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void **)&s->pin, (W + 1) * sizeof(int32_t), hipHostMallocDefault) != hipSuccess) {
        dfs_release(s);
        delete s;
        return set_error(DM_ERR_HIP, "step-API poll stream or pinned buffer");
    }
    *out = s;
    return DM_OK;
}

// Static partition: roots with a neighbour above them, heaviest (highest
// degree rank) first, dealt round-robin (the reference's v -> warp v % W).
static int dfs_reset(dm_dfs *s) {
    const size_t W = (size_t)s->workers, WM = W * DM_DFS_MAXK, J = (size_t)s->J;
    const dm_graph *g = s->g;
    s->k.assign(W, -1);
    s->status.assign(W, 1);
    s->next.assign(WM, 0);
    s->jobs.assign(W * J * DM_DFS_MAXK, -1);
    s->job_level.assign(W * J, 0);
    s->job_ind.assign(W * J, 0);
    s->job_valid.assign(W, 0);
    s->job_cur.assign(W, 0);
    size_t t = 0;
    for (int32_t v = g->rcsr.n - 1; v >= 0; v--) {
        if (g->rcsr.off[v + 1] - g->ostart[v] <= 0) continue;
        const size_t w = t % W, j = (size_t)s->job_valid[w]++;
        s->jobs[(w * J + j) * DM_DFS_MAXK] = v;
        t++;
    }
    const dm_dfs_data &d = s->d;
    HIPCHK(hipMemcpy(d.k, s->k.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.status, s->status.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.next, s->next.data(), WM * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.fixed, s->next.data(), WM * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.jobs, s->jobs.data(), s->jobs.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.job_level, s->job_level.data(), W * J * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.job_ind, s->job_ind.data(), W * J * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.job_valid, s->job_valid.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.job_cur, s->job_cur.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(d.result, 0, (size_t)std::max(s->ncls, 1) * 8));
    HIPCHK(hipMemset(s->d_stop, 0, 4));
    return DM_OK;
}

static int dfs_launch(dm_dfs *s, const void *kernel) {
    dm_dfs_data d = s->d;
    void *args[] = {&d};
    const size_t lds = s->motif ? (size_t)s->ncls * 8 : 0;
    hipError_t e = hipLaunchKernel(kernel, dim3(s->blocks), dim3(256), args, lds, s->ctx->stream);
    if (e != hipSuccess) return set_error(DM_ERR_HIP, "step-API kernel launch: %s", hipGetErrorString(e));
    s->last.launches++;
    return DM_OK;
}

// Lowest level below the last one with extensions left (getTargetLevel).
static int target_level(const dm_dfs *s, size_t w) {
    const int kw = s->k[w];
    for (int l = 0; l <= kw && l < s->K - 2; l++)
        if (s->next[w * DM_DFS_MAXK + l] > 0) return l;
    return -1;
}

// Extensions left from the target level up (getWeight).
static int weight_of(const dm_dfs *s, size_t w, int t) {
    int sum = 0;
    for (int l = t; l <= s->k[w]; l++) sum += std::max(s->next[w * DM_DFS_MAXK + l], 0);
    return sum;
}

// Moves work from the heaviest busy workers to the idle ones: each donation
// is the last extension of the donor's target level, as a job whose prefix
// is the donor's ids up to that level plus that vertex (donate).  Returns the
// number of jobs moved.
static uint64_t donate(dm_dfs *s) {
    const size_t W = (size_t)s->workers, J = (size_t)s->J;
    struct Donor {
        int weight;
        size_t w;
        int t;
        bool operator<(const Donor &o) const { return weight < o.weight; }
    };
    std::vector<size_t> idles;
    std::priority_queue<Donor> donors;
    long long total = 0;
    for (size_t w = 0; w < W; w++) {
        if (s->status[w] == 2 && s->job_cur[w] >= s->job_valid[w]) {
            idles.push_back(w);
        } else if (s->status[w] == 1 && s->k[w] >= 0) {
            const int t = target_level(s, w);
            if (t >= 0) {
                const int wt = weight_of(s, w, t);
                donors.push({wt, w, t});
                total += wt;
            }
        }
    }
    if (donors.empty() || idles.empty()) return 0;
    const long long ideal = (total + (long long)(donors.size() + idles.size()) - 1) /
                            (long long)(donors.size() + idles.size());
    const int per = (int)std::max<long long>(1, std::min<long long>(ideal, (long long)J));
    const HostCSR &c = s->g->rcsr;
    uint64_t moved = 0;
    // an idle worker's spent queue restarts at its first donated job, so at
    // most `per` <= J jobs are ever appended to it
    std::vector<char> fresh(idles.size(), 1);
    for (int job = 0; job < per && !donors.empty(); job++) {
        for (size_t ii = 0; ii < idles.size() && !donors.empty(); ii++) {
            const size_t i = idles[ii];
            Donor dn = donors.top();
            donors.pop();
            if (fresh[ii]) {
                fresh[ii] = 0;
                s->job_valid[i] = 0;
                s->job_cur[i] = 0;
                s->k[i] = -1;
            }
            const size_t w = dn.w;
            const int t = dn.t;
            int32_t &left = s->next[w * DM_DFS_MAXK + t];
            const int32_t x = s->ext[w * (size_t)s->d.ext_per_worker + list_off(s, t) + (size_t)(left - 1)];
            left--;
            const size_t slot = i * J + (size_t)s->job_valid[i]++;
            int32_t *jp = &s->jobs[slot * DM_DFS_MAXK];
            for (int l = 0; l < DM_DFS_MAXK; l++) jp[l] = l <= t ? s->id[w * DM_DFS_MAXK + l] : -1;
            jp[t + 1] = x;
            s->job_level[slot] = t + 1;
            uint64_t q = 0;
            if (s->motif && t + 1 >= 2) {
                const int o = ((t + 1 - 2) * (t + 1 + 1)) / 2;
                for (int l = 0; l <= t; l++)
                    if (host_adjacent(c, s->id[w * DM_DFS_MAXK + l], x)) q |= 1ull << (o + l);
                q += s->ind[w * DM_DFS_MAXK + t];
            }
            s->job_ind[slot] = q;
            moved++;
            dn.weight--;
            const int t2 = target_level(s, w);
            if (t2 != -1 && dn.weight >= per) {
                dn.t = t2;
                donors.push(dn);
            }
        }
    }
    return moved;
}

// Stopped kernel: state to the host, donation, state back.  Returns the
// number of jobs moved (0: nothing to donate).
static int rebalance(dm_dfs *s, uint64_t &moved) {
    const size_t W = (size_t)s->workers, WM = W * DM_DFS_MAXK, J = (size_t)s->J;
    const dm_dfs_data &d = s->d;
    s->id.resize(WM);
    s->ind.resize(WM);
    s->ext.resize(W * (size_t)d.ext_per_worker);
    HIPCHK(hipMemcpy(s->k.data(), d.k, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s->status.data(), d.status, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s->id.data(), d.id, WM * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s->next.data(), d.next, WM * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s->ind.data(), d.ind, WM * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s->ext.data(), d.ext, s->ext.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s->jobs.data(), d.jobs, s->jobs.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s->job_level.data(), d.job_level, W * J * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s->job_ind.data(), d.job_ind, W * J * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s->job_valid.data(), d.job_valid, W * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(s->job_cur.data(), d.job_cur, W * 4, hipMemcpyDeviceToHost));
    moved = donate(s);
    if (moved == 0) return DM_OK;
    HIPCHK(hipMemcpy(d.k, s->k.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.next, s->next.data(), WM * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.jobs, s->jobs.data(), s->jobs.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.job_level, s->job_level.data(), W * J * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.job_ind, s->job_ind.data(), W * J * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.job_valid, s->job_valid.data(), W * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d.job_cur, s->job_cur.data(), W * 4, hipMemcpyHostToDevice));
    return DM_OK;
}

static int set_stop(dm_dfs *s, int v) {
    s->pin[0] = v;
    HIPCHK(hipMemcpyAsync(s->d_stop, s->pin, 4, hipMemcpyHostToDevice, s->poll));
    HIPCHK(hipStreamSynchronize(s->poll));
    return DM_OK;
}

static int dfs_run(dm_dfs *s, const void *kernel, uint64_t *out, size_t out_len, dm_dfs_stats *st_out) {
    const size_t nres = s->motif ? (size_t)s->ncls : 1;
    if (!kernel || !out) return set_error(DM_ERR_ARG, "kernel or out is NULL");
    if (out_len < nres) return set_error(DM_ERR_ARG, "output needs %zu slots (got %zu)", nres, out_len);
    if (hipSetDevice(s->ctx->device) != hipSuccess) return set_error(DM_ERR_HIP, "hipSetDevice");
    std::lock_guard<std::recursive_mutex> lock(s->ctx->mu);
    const double t0 = now_ms();
    s->last = dm_dfs_stats{};
    int r = dfs_reset(s);
    if (r) return r;
    if ((r = dfs_launch(s, kernel))) return r;
    const dm_dfs_config &cfg = s->cfg;
    bool lb = cfg.threshold < 100 || cfg.stop_every_us > 0;
    const int poll_us = cfg.poll_us ? cfg.poll_us : 1000;
    double last_stop = now_ms();
    while (lb) {
        std::this_thread::sleep_for(std::chrono::microseconds(poll_us));
        hipError_t q = hipStreamQuery(s->ctx->stream);
        if (q == hipSuccess) break;  // every worker ran out of work
        if (q != hipErrorNotReady) return set_error(DM_ERR_HIP, "step-API kernel: %s", hipGetErrorString(q));
        if (cfg.watchdog_s > 0 && now_ms() - t0 > cfg.watchdog_s * 1e3) {
            if (set_stop(s, 1) == DM_OK) (void)hipStreamSynchronize(s->ctx->stream);
            return set_error(DM_ERR_DEVICE, "step-API run exceeded its watchdog");
        }
        // idle share (gpuIsIdle), read while the kernel runs
        HIPCHK(hipMemcpyAsync(s->pin + 1, s->d.status, (size_t)s->workers * 4, hipMemcpyDeviceToHost, s->poll));
        HIPCHK(hipStreamSynchronize(s->poll));
        int idle = 0;
        for (int w = 0; w < s->workers; w++) idle += s->pin[1 + w] == 2 ? 1 : 0;
        const double pct = 100.0 * idle / s->workers;
        const bool forced = cfg.stop_every_us > 0 && (now_ms() - last_stop) * 1e3 >= cfg.stop_every_us;
        const bool idle_trigger = cfg.threshold < 100 && (pct >= 100.0 || pct > cfg.threshold);
        if (!forced && !idle_trigger) continue;
        if (cfg.max_rebalances && s->last.rebalances >= cfg.max_rebalances) break;
        // stopKernel: every worker leaves its loop at the next step and saves
        // its level
        const double t1 = now_ms();
        if ((r = set_stop(s, 1))) return r;
        HIPCHK(hipStreamSynchronize(s->ctx->stream));
        uint64_t moved = 0;
        if ((r = rebalance(s, moved))) return r;
        if ((r = set_stop(s, 0))) return r;
        s->last.rebalances++;
        s->last.donated_jobs += moved;
        if ((r = dfs_launch(s, kernel))) return r;
        s->last.lb_ms += now_ms() - t1;
        last_stop = now_ms();
        // nothing to donate: balancing cannot help any more
        if (moved == 0 && cfg.stop_every_us == 0) lb = false;
    }
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
    std::vector<unsigned long long> h(nres);
    HIPCHK(hipMemcpy(h.data(), s->d.result, nres * 8, hipMemcpyDeviceToHost));
    uint64_t total = 0;
    for (size_t i = 0; i < nres; i++) {
        out[i] = h[i];
        total += h[i];
    }
    s->last.total_ms = now_ms() - t0;
    s->last.workers = s->workers;
    s->last.vwave = s->vw;
    s->last.subgraphs = total;
    if (st_out) *st_out = s->last;
    return DM_OK;
}

}  // namespace dm

using namespace dm;

extern "C" {

int dm_dfs_config_init(dm_dfs_config *cfg) {
    if (!cfg) return set_error(DM_ERR_ARG, "cfg is NULL");
    memset(cfg, 0, sizeof(*cfg));
    cfg->vwave = 32;
    cfg->jobs_per_worker = 16;
    cfg->threshold = 30;
    cfg->poll_us = 1000;
    cfg->watchdog_s = 120.0;
    return DM_OK;
}

int dm_dfs_create(dm_graph *g, int k, int pattern, const dm_dfs_config *cfg, dm_dfs **out) {
    if (!g || !out) return set_error(DM_ERR_ARG, "graph or out is NULL");
    *out = nullptr;
    dm_dfs_config c;
    if (cfg) c = *cfg;
    else dm_dfs_config_init(&c);
    return dfs_create(g, k, pattern, c, out);
}

int dm_dfs_run(dm_dfs *s, const void *kernel, uint64_t *out, size_t out_len, dm_dfs_stats *st) {
    if (!s) return set_error(DM_ERR_ARG, "state is NULL");
    return dfs_run(s, kernel, out, out_len, st);
}

void dm_dfs_free(dm_dfs *s) {
    if (!s) return;
    dfs_release(s);
    delete s;
}

}  // extern "C"
