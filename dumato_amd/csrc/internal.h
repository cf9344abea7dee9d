// Internal declarations shared by the host runtime (C++) and the HIP kernels.
#ifndef DM_INTERNAL_H
#define DM_INTERNAL_H

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dumato.h"

namespace dm {

// ---- limits ---------------------------------------------------------------
constexpr int kMaxK = 12;              // largest clique size; motifs stop at 7
constexpr int kMaxMotifK = 7;
constexpr int kMaxLevels = kMaxK - 2;  // stack levels 0..K-3 (+1 spare)
constexpr int kJobIds = 12;

// ---- errors ---------------------------------------------------------------
int set_error(int code, const char *fmt, ...);
int last_code();

// ---- host graph -------------------------------------------------------------
struct HostCSR {
    int32_t n = 0;
    std::vector<int64_t> off;  // n + 1
    std::vector<int32_t> adj;  // sorted neighbour lists
    int32_t maxdeg = 0;
    int64_t nnz() const { return off.empty() ? 0 : off.back(); }
};

int parse_edgelist(const char *path, std::vector<int32_t> &src, std::vector<int32_t> &dst);
int build_csr(int32_t n_hint, const int32_t *src, const int32_t *dst, size_t m, HostCSR &out);

// ---- pattern tables -----------------------------------------------------------
struct PatternTable {
    int k = 0;
    int ncls = 0;
    std::vector<uint16_t> q2c;       // quick code -> class id
    std::vector<uint64_t> cls_code;  // class id -> canonical code
};
const PatternTable *pattern_table(int k);

// ---- random numbers (xoshiro256**, seeded by splitmix64) ---------------------
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; i++) {
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    inline double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    inline uint64_t below(uint64_t n) { return (uint64_t)(uniform() * (double)n) % (n ? n : 1); }
};

// ---- device-side structures (identical layout on host and device) ----------
// A donated job: children [lo, hi) of the node with prefix ids[0..lvl] in
// query q.  Word j holds (launch epoch << 32 | value): words 0..3 = lvl, lo,
// hi, q; words 4.. = ids.
constexpr int kJobWords = 4 + kJobIds;
struct Job {
    uint64_t w[kJobWords];
};

// Scheduler counters.  Fields that idle and busy workers poll or update
// concurrently each sit on their own 128-byte line, so traffic on one never
// stalls the others.
struct SchedCounters {
    unsigned long long outstanding;  // unfinished jobs (static + donated)
    char pad0[120];
    int32_t static_left;             // worker queues whose last job is not yet claimed
    char pad2[124];
    // donated-job queue: next ticket (high 32 bits) and next free slot (low
    // 32 bits) in one word, so one load tells whether tickets wait for jobs
    unsigned long long q_ht;
    char pad3[120];
    int32_t error;                   // 1 watchdog, 2 arena overflow, 3 bad job
    char pad5[124];
    unsigned long long donations;
    unsigned long long jobs;
    unsigned long long alg_bytes;
    unsigned long long count;        // clique count / motif subgraph total
    unsigned long long busy_sum;     // sum over workers of cycles spent in jobs
    unsigned long long busy_max;     // largest per-worker busy cycles
    // timeline (wall-clock ticks): first block start, then the latest worker
    // to start its first job, to find static work exhausted, to end its last
    // job and to exit
    unsigned long long t_start, t_first_job, t_static_done, t_last_job, t_exit;
    unsigned long long job_max;  // cycles of the longest single job
    unsigned long long active_workers;  // workers that ran at least one job
};

// One counting query of a launch.  A single-graph call has one; a batch
// has one per (graph, pattern, k).
struct QueryDesc {
    int32_t K;
    int32_t motif;
    int32_t ncls;
    int32_t graph;                    // graph index in a batch (latency grouping)
    const uint16_t *lut;              // quick code -> class (motif)
    unsigned long long *out;          // clique: out[0]; motif: out[0 .. ncls)
    unsigned long long *done_ticks;   // nullable: wall clock of the query's last job end
};

struct EnumParams {
    // graph: symmetric CSR in degree-rank ids (a batch concatenates its graphs,
    // each relabelled inside its own id range); ostart[v] = first index of
    // N(v) above v
    const int64_t *off;
    const int32_t *adj;
    const int64_t *ostart;
    int32_t n;
    int32_t nq;                       // queries
    const QueryDesc *qd;
    int32_t kmax;                     // largest k of the queries (LDS record size)
    int32_t hist_slots;               // per-worker class counters (batch launches)
    int32_t nworkers;
    int32_t pad1;
    // static partition: worker w owns records sq_jobs[4*sq_off[w] .. 4*sq_off[w+1]):
    //   (root, lo, hi, q)   children [lo, hi) of root (q = query), or
    //   (r0, -1, r1, q)     every root of [r0, r1), all children
    // sq_jobs == nullptr: the strided partition of one query, with no arrays:
    // worker w's j-th job is root r = n-1 - (w + j*nworkers) as the range
    // (r, -1, r+1, 0) -- the reference's v -> warp v % W deal, heaviest
    // (highest degree rank) first
    const int64_t *sq_off;
    const int32_t *sq_jobs;
    int32_t *sq_cur;
    // donated jobs
    Job *q;
    uint32_t q_cap;
    int32_t pad2;
    const int32_t *epoch;             // launch epoch in device memory (set by the init kernel)
    unsigned long long *start_ticks;  // wall clock of the first block start
    SchedCounters *sc;
    // per-worker stacks: LDS slice of lds_per_worker bytes, then global arena
    char *arena;
    int64_t arena_per_worker;
    int32_t lds_per_worker;
    int32_t pad3;
    unsigned long long *root_cost;    // nullable, per vertex
    unsigned long long *root_count;   // nullable
    int32_t steal, steal_min, bitmap, pad4;
    uint64_t poll_ticks;              // wall-clock ticks between a worker's donation checks
    uint64_t watchdog_ticks;          // wall-clock ticks
    // local bitmaps prebuilt for some roots (cliques, k = 4..6): pre_roots
    // ascending (npre entries), root pre_roots[i]'s rows (d x NW uint64) and
    // out-degrees (d int32) at pre + pre_offs[i]
    const int32_t *pre_roots;
    const int64_t *pre_offs;
    const char *pre;
    // per prebuilt root: the next child its shared-cursor jobs claim (zeroed
    // before each launch); a static job (root, -2, P, q) is one of the P
    // jobs that share root's cursor
    int32_t *pre_cursor;
    int32_t npre;
    // workers per static queue (0 or 1: one queue per worker): worker w
    // claims from queue w / q_group of the (nworkers + q_group - 1) / q_group
    // queues, and an idle worker probes the other queues
    int32_t q_group;
    int32_t pad6;
    // nullable: one bit per static queue, set while the queue holds an
    // unclaimed job (set by the init kernel, cleared by the claimer of the
    // queue's last job); idle workers steal from a set bit's queue
    unsigned long long *qmask;
};

// 64-bit words of the non-empty-queue mask of a launch with these workers
inline int qmask_words(int nworkers, int q_group) {
    const int g = q_group > 1 ? q_group : 1;
    return ((nworkers + g - 1) / g + 63) / 64;
}

// A static or donated job's query word: bits 0..15 the query index, bits
// 16..23 the job's own donation threshold (0: the launch's steal_min).
constexpr int kQueryBits = 16;
constexpr int job_query(int32_t q) { return q & ((1 << kQueryBits) - 1); }
constexpr int job_steal_min(int32_t q) { return (q >> kQueryBits) & 0xff; }
constexpr int32_t make_job_query(int q, int steal_min) { return q | (steal_min << kQueryBits); }

// kernels (enum_kernel.hip)
// mode: 0 cliques only, 1 motifs only, 2 mixed (batch), 3 motifs of size >= 5
// (a motif kernel budgeted for one more wave per SIMD)
inline int launch_mode(bool motif, int K) { return motif ? (K >= 5 ? 3 : 1) : 0; }
size_t enum_lds_bytes(int vwave, int threads, int kmax, int ncls_single, int hist_slots, int lds_per_worker);
int enum_occupancy(int vwave, int mode, int threads, size_t lds, int *blocks_per_cu);
int enum_launch(const EnumParams &p, int mode, int vwave, int blocks, int threads, size_t lds,
                void *stream);
// resets the scheduler state of a launch (graph-capturable: kernels only)
// queues: worker queues holding at least one static job
int enum_init_launch(const EnumParams &p, unsigned long long njobs, int queues, int32_t *epoch,
                     unsigned long long *results, size_t nresults, unsigned long long *done_ticks, int nq,
                     void *stream);
int fill_u32_launch(void *ptr, uint32_t value, size_t count, void *stream);
// builds the local bitmaps of p.pre_roots into p.pre (one block per root)
int prebuild_launch(const EnumParams &p, void *stream);
// bytes of a prebuilt bitmap of a root with out-degree d (0: not on the bitmap path)
constexpr int64_t prebuilt_bytes(int32_t d) {
    const int nw = d <= 64 ? 1 : d <= 128 ? 2 : d <= 256 ? 4 : 0;
    return nw ? (((int64_t)d * nw * 8 + (int64_t)d * 4 + 15) & ~(int64_t)15) : 0;
}

}  // namespace dm

#endif
