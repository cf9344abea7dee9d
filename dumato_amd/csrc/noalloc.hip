// Counting without device allocation: dm_count_noalloc.
//
// No reference counterpart.  It is the counting call of dm_clique_count /
// dm_motif_count for callers that must not allocate device memory (code under
// stream capture, serving loops, harnesses that forbid hipMalloc).  Every
// device buffer the enumeration kernel needs beyond the graph (its input) and
// the result (its output) -- scheduler counters, launch epoch, query
// descriptor, result slots, class table, static-queue cursors, donated-job
// queue and the per-worker stack arenas -- is carved from one file-scope
// device array of kScratchBytes.  Static work uses the kernel's strided
// partition (root r goes to worker (n-1-r) mod workers, highest degree rank
// first), which needs no job arrays.  The number of workers is capped so that
// their arenas fit; a graph whose arena does not fit one block is rejected.
//
// The array is one per GPU (each device holds its own copy at its own
// address) and keeps the last call's contents: the init kernel clears the
// counters, cursors and result slots of every launch, and queue slots carry
// the launch epoch, which only grows, so old slots never match.  Calls are
// serialised per device: the call holds the device lock and runs on the
// device's one stream from its first copy to its last.  Work that another
// program or stream puts on the same array while a call runs would race with
// it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "internal.h"
#include "runtime.h"

namespace dm {

constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

// layout of the array (byte offsets)
constexpr size_t kSoSc = 0;                           // SchedCounters
constexpr size_t kSoEpoch = 1024;                     // int32 launch epoch
constexpr size_t kSoQd = 1280;                        // QueryDesc
constexpr size_t kSoRes = 2048;                       // result slots
constexpr size_t kResSlots = 2048;
constexpr size_t kSoLut = 32 << 10;                   // class table (uint16)
constexpr size_t kLutBytes = (size_t)2 << 20;         // 2^20 entries: k <= 7
constexpr size_t kSoCur = kSoLut + kLutBytes;         // static-queue cursors
constexpr size_t kMaxWorkers = 65536;
constexpr size_t kSoQmask = kSoRes + kResSlots * 8;    // non-empty static queues (1 bit per worker)
constexpr size_t kSoQueue = kSoCur + kMaxWorkers * 4;  // donated-job queue
constexpr uint32_t kQueueSlots = 16384;
constexpr size_t kSoArena = kSoQueue + (size_t)kQueueSlots * sizeof(Job);
constexpr size_t kArenaBytes = kScratchBytes - kSoArena;
static_assert(sizeof(SchedCounters) <= kSoEpoch - kSoSc, "counters overlap the epoch");
static_assert(sizeof(QueryDesc) <= kSoRes - kSoQd, "query descriptor overlaps the results");
static_assert(kSoQmask + kMaxWorkers / 8 <= kSoLut, "queue mask overlaps the class table");
static_assert(kSoQueue % 256 == 0 && kSoArena % 256 == 0, "misaligned scratch regions");

// Device address of `scratch` per device, taken once per device with the
// device current.  (In host code the array's name only stands for it.)
static std::mutex g_mu;
constexpr int kMaxDevices = 64;
static char *g_ptr[kMaxDevices] = {nullptr};
static int g_lut_k[kMaxDevices] = {0};  // k whose class table the array holds

static int scratch_on(DeviceCtx *c, char **out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (c->device < 0 || c->device >= kMaxDevices) return set_error(DM_ERR_HIP, "device id %d", c->device);
    if (!g_ptr[c->device]) {
        char *p = nullptr;
        hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(DM_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        // queue slots and the epoch start at zero (epoch 0 never matches a
        // launch); cleared with a kernel, not a memset
        int r = fill_u32_launch(p + kSoEpoch, 0, 1, c->stream);
        if (!r) r = fill_u32_launch(p + kSoQueue, 0, (size_t)kQueueSlots * sizeof(Job) / 4, c->stream);
        if (r) return r;
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return set_error(DM_ERR_HIP, "scratch clear: %s", hipGetErrorString(e));
        g_ptr[c->device] = p;
        g_lut_k[c->device] = 0;
    }
    *out = g_ptr[c->device];
    return DM_OK;
}

static int noalloc_count(dm_graph *g, bool motif, int K, const dm_config &cfg, uint64_t *out, size_t out_len,
                         uint64_t *total_out) {
    // ---- arguments and sizes (host arithmetic only) ----
    const PatternTable *tab = nullptr;
    if (motif) {
        if (K < 3 || K > kMaxMotifK)
            return set_error(DM_ERR_UNSUPPORTED, "motif size k=%d unsupported (3..%d)", K, kMaxMotifK);
        if (!(tab = pattern_table(K))) return last_code();
    } else if (K < 2 || K > kMaxK) {
        return set_error(DM_ERR_ARG, "clique size k=%d unsupported (2..%d)", K, kMaxK);
    }
    const size_t nres = motif ? (size_t)tab->ncls : 1;
    if (out_len < nres) return set_error(DM_ERR_ARG, "output needs %zu slots (got %zu)", nres, out_len);
    if (cfg.sched != DM_SCHED_PLAIN) return set_error(DM_ERR_ARG, "dm_count_noalloc: plain scheduler only");
    if (cfg.record_costs || cfg.per_root) return set_error(DM_ERR_ARG, "dm_count_noalloc: no per-root records");
    if (cfg.root_begin > 0 || cfg.root_end > 0) return set_error(DM_ERR_ARG, "dm_count_noalloc: no root ranges");
    int r = check_config(cfg);
    if (r) return r;
    if (nres > kResSlots || (tab && tab->q2c.size() * 2 > kLutBytes))
        return set_error(DM_ERR_ARG, "dm_count_noalloc: k=%d needs more result or table space than the %zu-byte "
                                     "static scratch holds", K, kScratchBytes);
    const int64_t arena = arena_bound_for(g->rcsr.maxdeg, g->max_out, K, motif);
    const int vw_min = cfg.vwave > 0 ? cfg.vwave : 64;
    int threads = 256;
    block_threads(cfg, &threads);
    if (arena * (threads / std::max(vw_min, 8)) > (int64_t)kArenaBytes)
        return set_error(DM_ERR_ARG, "dm_count_noalloc: this graph needs %lld bytes of stack arena per block at "
                                     "k=%d, beyond the %zu bytes of the static scratch (max degree %d)",
                         (long long)(arena * (threads / std::max(vw_min, 8))), K, kArenaBytes, g->rcsr.maxdeg);
    if (!g->dev)
        return set_error(DM_ERR_ARG, "dm_count_noalloc: the graph is not on the GPU yet (its first counting or "
                                     "tuning call uploads it)");
    if (hipSetDevice(g->dev->ctx->device) != hipSuccess) return set_error(DM_ERR_HIP, "hipSetDevice");

    DeviceGraph *dg = g->dev;
    DeviceCtx *d = dg->ctx;
    std::lock_guard<std::recursive_mutex> lock(d->mu);
    char *sp = nullptr;
    if ((r = scratch_on(d, &sp))) return r;
    Shape sh;
    if ((r = pick_shape(g, K, motif, cfg, sh))) return r;
    const int wpb = sh.threads / sh.vwave;
    const int64_t fit_blocks = (int64_t)kArenaBytes / (arena * wpb);
    const int64_t fit_workers = std::min<int64_t>(fit_blocks * wpb, (int64_t)kMaxWorkers);
    if (fit_workers < wpb)
        return set_error(DM_ERR_ARG, "dm_count_noalloc: one block of %d workers needs %lld bytes of stack arena, "
                                     "beyond the %zu bytes of the static scratch", wpb, (long long)(arena * wpb),
                         kArenaBytes);
    if (sh.workers > fit_workers) {
        if (cfg.blocks > 0)
            return set_error(DM_ERR_ARG, "dm_count_noalloc: %d blocks need more stack arena than the static "
                                         "scratch holds (at most %lld blocks)", cfg.blocks,
                             (long long)(fit_workers / wpb));
        sh.blocks = (int)(fit_workers / wpb);
        sh.workers = sh.blocks * wpb;
    }
    const int32_t n = g->rcsr.n;
    hipStream_t s = d->stream;
    unsigned long long *res = (unsigned long long *)(sp + kSoRes);
    uint16_t *lut = (uint16_t *)(sp + kSoLut);
    if (tab && g_lut_k[d->device] != K) {
        hipError_t e = hipMemcpyAsync(lut, tab->q2c.data(), tab->q2c.size() * 2, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return set_error(DM_ERR_HIP, "class table copy: %s", hipGetErrorString(e));
        g_lut_k[d->device] = K;
    }
    QueryDesc qd;
    memset(&qd, 0, sizeof(qd));
    qd.K = K;
    qd.motif = motif ? 1 : 0;
    qd.ncls = motif ? tab->ncls : 0;
    qd.lut = tab ? lut : nullptr;
    qd.out = res;
    hipError_t e = hipMemcpyAsync(sp + kSoQd, &qd, sizeof(qd), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return set_error(DM_ERR_HIP, "query copy: %s", hipGetErrorString(e));

    EnumParams p;
    memset(&p, 0, sizeof(p));
    p.off = dg->d_off;
    p.adj = dg->d_adj;
    p.ostart = dg->d_ostart;
    p.n = n;
    p.nq = 1;
    p.qd = (const QueryDesc *)(sp + kSoQd);
    p.kmax = K;
    p.nworkers = sh.workers;
    p.sq_off = nullptr;
    p.sq_jobs = nullptr;  // strided static partition
    p.sq_cur = (int32_t *)(sp + kSoCur);
    p.qmask = motif ? (unsigned long long *)(sp + kSoQmask) : nullptr;  // motifs only, as in dm_count
    p.q = (Job *)(sp + kSoQueue);
    p.q_cap = kQueueSlots;
    p.epoch = (const int32_t *)(sp + kSoEpoch);
    p.sc = (SchedCounters *)(sp + kSoSc);
    p.arena = sp + kSoArena;
    p.arena_per_worker = arena;
    p.lds_per_worker = sh.lds_per_worker;
    p.steal = cfg.steal ? 1 : 0;
    p.steal_min = cfg.steal_min ? cfg.steal_min : tuned_steal_min(g, K, motif);
    p.poll_ticks = (uint64_t)cfg.steal_period * (uint64_t)d->wall_khz / 1000;
    p.bitmap = cfg.clique_bitmap ? 1 : 0;
    p.watchdog_ticks = cfg.watchdog_s > 0 ? (uint64_t)(cfg.watchdog_s * d->wall_khz * 1000.0) : 0;
    const unsigned long long njobs = (unsigned long long)n;
    const int queues = (int)std::min<int64_t>(n, sh.workers);
    if ((r = enum_init_launch(p, njobs, queues, (int32_t *)(sp + kSoEpoch), res, nres, nullptr, 0, s))) return r;
    if ((e = hipEventRecord(d->ev0, s)) != hipSuccess) return set_error(DM_ERR_HIP, "event: %s", hipGetErrorString(e));
    if (njobs > 0 && (r = enum_launch(p, launch_mode(motif, K), sh.vwave, sh.blocks, sh.threads, sh.lds_total, s)))
        return r;
    if ((e = hipEventRecord(d->ev1, s)) != hipSuccess) return set_error(DM_ERR_HIP, "event: %s", hipGetErrorString(e));
    SchedCounters sc1;
    e = hipMemcpyAsync(&sc1, sp + kSoSc, sizeof(sc1), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(out, res, nres * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return set_error(DM_ERR_HIP, "result copy: %s", hipGetErrorString(e));
    float kms = 0;
    (void)hipEventElapsedTime(&kms, d->ev0, d->ev1);
    if (sc1.error) return device_error(sc1.error, "enumeration kernel");
    uint64_t total = 0;
    for (size_t i = 0; i < nres; i++) total += out[i];
    if (total_out) *total_out = total;
    dm_stats &st = g->last;
    memset(&st, 0, sizeof(st));
    st.device_ms = kms;
    st.subgraphs = total;
    st.alg_bytes = sc1.alg_bytes;
    st.donations = sc1.donations;
    st.jobs = sc1.jobs;
    st.busy_sum_cycles = sc1.busy_sum;
    st.busy_max_cycles = sc1.busy_max;
    st.job_max_cycles = sc1.job_max;
    st.active_workers = sc1.active_workers;
    st.vwave = sh.vwave;
    st.workers = sh.workers;
    st.blocks = sh.blocks;
    st.lds_bytes = sh.lds_per_worker;
    return DM_OK;
}

}  // namespace dm

extern "C" int dm_count_noalloc(dm_graph *g, int motif, int k, const dm_config *cfg, uint64_t *out, size_t out_len,
                                uint64_t *total) {
    using namespace dm;
    if (!g || !out) return set_error(DM_ERR_ARG, "graph or out is NULL");
    dm_config c;
    if (cfg) c = *cfg;
    else dm_config_init(&c);
    return noalloc_count(g, motif != 0, k, c, out, out_len, total);
}
