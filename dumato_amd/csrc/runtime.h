// Host-side state of the runtime: one context per GPU, shared by every graph
// handle and batch on it, and the per-handle graph data.
#ifndef DM_RUNTIME_H
#define DM_RUNTIME_H

#include <hip/hip_runtime_api.h>

#include <memory>
#include <mutex>
#include <vector>

#include "internal.h"
#include "mlp.h"

namespace dm {

// State of one GPU, created on first use and kept until the process ends.
// Every counting launch on the device goes to its one stream, so launches run
// one after another on the GPU; that is what lets them share the donated-job
// queue (slots are tagged with the launch epoch) and the workspaces.  `mu` is
// held for the whole of a call that uses the stream or the workspaces (it is
// recursive: a learned partition runs a k=3 count inside a counting call).
struct DeviceCtx {
    int device = 0;
    int num_cu = 0;
    int wall_khz = 100000;
    char arch[256] = "";
    std::recursive_mutex mu;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // donated-job queue (fixed size, never reallocated: batches capture it)
    Job *d_q = nullptr;
    size_t q_cap_alloc = 0;
    int32_t *d_epoch = nullptr;
    // class tables per k (uploaded once, never reallocated)
    uint16_t *d_lut[kMaxMotifK + 1] = {nullptr};
    // grow-only workspaces of single-graph calls
    char *d_arena = nullptr;
    size_t arena_cap = 0;
    int64_t *d_sq_off = nullptr;
    size_t sq_off_cap = 0;
    int32_t *d_sq_jobs = nullptr;
    size_t sq_jobs_cap = 0;
    int32_t *d_sq_cur = nullptr;
    size_t sq_cur_cap = 0;
    unsigned long long *d_qmask = nullptr;  // non-empty static queues
    size_t qmask_cap = 0;
    SchedCounters *d_sc = nullptr;
    unsigned long long *d_results = nullptr;  // clique count or motif histogram
    size_t results_cap = 0;
    QueryDesc *d_qd = nullptr;
    size_t qd_cap = 0;
    // prebuilt local bitmaps of a learned partition's heavy roots
    int32_t *d_pre_roots = nullptr;
    size_t pre_roots_cap = 0;
    int64_t *d_pre_offs = nullptr;
    size_t pre_offs_cap = 0;
    int32_t *d_pre_cursor = nullptr;
    size_t pre_cursor_cap = 0;
    char *d_pre = nullptr;
    size_t pre_cap = 0;
    // serial of the partition whose queues (and prebuilt-root list) the
    // buffers above hold: a repeated query skips the upload
    uint64_t up_serial = 0;
    // pinned staging of a launch's counters and results (copied back per call)
    char *h_stage = nullptr;
    size_t stage_cap = 0;
};

// The current device's context (created, with its stream, events and queue,
// on the first call).
int device_ctx(DeviceCtx **out);

// Device copy of one graph handle.
struct DeviceGraph {
    DeviceCtx *ctx = nullptr;
    int64_t *d_off = nullptr;
    int32_t *d_adj = nullptr;
    int64_t *d_ostart = nullptr;
    unsigned long long *d_root_cost = nullptr;
    size_t root_cost_cap = 0;
    unsigned long long *d_root_count = nullptr;
    size_t root_count_cap = 0;
};

struct Shape {
    int vwave = 64;
    int threads = 256;
    int lds_per_worker = 0;
    size_t lds_total = 0;
    int blocks = 0;
    int workers = 0;
};

}  // namespace dm

struct dm_graph {
    dm::HostCSR csr;                 // original ids
    dm::HostCSR rcsr;                // degree-rank ids (ascending (degree, id))
    std::vector<int32_t> perm, inv;  // perm[orig] = rank, inv[rank] = orig
    std::vector<int64_t> ostart;     // per rank: index of the first neighbour above it
    int32_t max_out = 0;
    dm::DeviceGraph *dev = nullptr;
    dm_stats last{};
    std::vector<uint64_t> last_costs, last_counts;  // original ids
    dm::CostModel *model = nullptr;
    std::vector<float> features;                    // kMlpIn per degree-rank vertex
    std::vector<float> root_density;                // per degree-rank vertex: oriented triangles / C(|N+(v)|, 2)
    // per degree-rank vertex: log clique growth from size growth_from to
    // growth_to (the learned partition's extrapolation; cached per sizes)
    std::vector<float> growth;
    int growth_from = 0, growth_to = 0;
    // autotuned launch parameters per [motif][k]; 0 = not tuned
    int tuned_vwave[2][dm::kMaxK + 1] = {{0}};
    int tuned_lds[2][dm::kMaxK + 1] = {{0}};
    int tuned_steal_min[2][dm::kMaxK + 1] = {{0}};
    // static partitions of recent calls (the roots, costs and model they
    // derive from do not change between calls); model_gen counts trainings
    struct Partition {
        int key[8];
        uint64_t model_gen;
        uint64_t serial = 0;  // unique per computed partition (device upload cache)
        std::vector<int64_t> sq_off;
        std::vector<int32_t> sq_jobs;
        uint64_t skipped_bytes;
        int group = 1;  // workers per static queue (sq_off has one entry per queue, + 1)
        // roots whose local bitmaps a pre-pass builds once for their pieces
        // (ascending), their offsets in the prebuilt buffer, its size
        std::vector<int32_t> pre_roots;
        std::vector<int64_t> pre_offs;
        int64_t pre_bytes = 0;
    };
    std::vector<Partition> partitions;  // most recent first
    // roots with work (and the node bytes of motif roots without) per
    // (root range, sample step, motif): graph-derived, so a repartition
    // reuses them
    struct RootList {
        int key[4];
        std::vector<int32_t> roots;
        uint64_t skipped_bytes;
    };
    // most recent first; a call holds its entry, so a nested call (the
    // learned partition's feature count) may evict it safely
    std::vector<std::shared_ptr<const RootList>> root_lists;
    uint64_t model_gen = 0;
    double last_infer_ms = 0, last_bin_ms = 0;  // phases of the last learned partition
};

namespace dm {
// grow-only device buffer
template <class T>
inline int dev_grow(T *&ptr, size_t &cap, size_t need) {
    if (need <= cap && ptr) return DM_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    size_t alloc = need > 0 ? need : 1;
    hipError_t e = hipMalloc((void **)&ptr, alloc * sizeof(T));
    if (e != hipSuccess)
        return set_error(DM_ERR_NOMEM, "hipMalloc(%zu B): %s", alloc * sizeof(T), hipGetErrorString(e));
    cap = alloc;
    return DM_OK;
}
int64_t arena_bound_for(int32_t maxdeg, int32_t max_out, int K, bool motif);
int graph_finish(dm_graph *g);
// Uploads g's CSR to the current device (once) and binds g to its context.
int ensure_device(dm_graph *g);
// Uploads a device copy of a degree-rank CSR (off, adj, ostart).
int upload_csr(const HostCSR &rc, const std::vector<int64_t> &ostart, DeviceGraph *d);
void free_device_graph(DeviceGraph *d);
// The class table of k on the context's device (uploaded on first use).
int device_lut(DeviceCtx *c, int K, const uint16_t **out);
// Threads per block for cfg.waves_per_block (1, 2 or 4 waves; 0 = 4).
int block_threads(const dm_config &cfg, int *threads);
int pick_shape(dm_graph *g, int K, bool motif, const dm_config &cfg, Shape &sh);
bool root_has_work(const dm_graph *g, int32_t v, int K, bool motif);
void plain_partition(const dm_graph *g, const std::vector<int32_t> &roots, int K, bool motif, int workers,
                     std::vector<int64_t> &sq_off, std::vector<int32_t> &sq_jobs);
int run_enum(dm_graph *g, int K, bool motif, const dm_config &cfg, uint64_t *hist, size_t hist_len,
             uint64_t *count_out, int sample_mod = 1);
int check_config(const dm_config &cfg);
int tuned_steal_min(const dm_graph *g, int K, bool motif);
int tuned_lds(const dm_graph *g, int K, bool motif, int vwave, int fallback);
int max_lds_per_worker(int vwave, int threads, int mode, int K, int ncls_single, int hist_slots);
// Kernel abort reason of a launch's error word as a DM_ERR_DEVICE failure.
int device_error(int code, const char *what);

// sched.cpp
int autotuned_vwave(dm_graph *g, int K, bool motif);
int learned_partition(dm_graph *g, const std::vector<int32_t> &roots, int K, bool motif, const dm_config &cfg,
                      const Shape &sh, dm_graph::Partition &part);
void free_model(dm_graph *g);

// graph.cpp
int gen_erdos_renyi(int32_t n, int64_t m, uint64_t seed, std::vector<int32_t> &s, std::vector<int32_t> &d);
int gen_chung_lu(int32_t n, int64_t m_target, double gamma, double wmax, uint64_t seed, std::vector<int32_t> &s,
                 std::vector<int32_t> &d);
int gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed, std::vector<int32_t> &s,
             std::vector<int32_t> &d);
int gen_barabasi_albert(int32_t n, int32_t m, int32_t planted, int32_t kmin, int32_t kmax, uint64_t seed,
                        std::vector<int32_t> &s, std::vector<int32_t> &d);
}  // namespace dm

#endif
