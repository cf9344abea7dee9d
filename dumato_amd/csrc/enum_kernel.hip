// DFS-wide subgraph enumeration on gfx950: k-clique counting and k-motif
// histograms.
//
// Reference semantics (what is computed):
//   clique  src/main/clique_PA.cu:9-32 with extend_clique / forward / backward /
//           aggregate_counter (src/main/DuMatoGPU.cu:1194-1220, 836-844, 745-750,
//           1769-1771): C_{l+1} = N+(last) & C_l, the last level adds |C|.
//   motif   src/main/motifs_DM_WCLB.cu:9-36 with extend(0,k) (:1702-1721),
//           canonicalFilter2 (:989-1012), forward_induce (:846-869) and
//           aggregate_pattern (:930-967).
//
// How it runs here (MI355X-first, not a translation):
//   * A worker is a virtual wave of W = 8/16/32/64 lanes of a wave64.  All
//     cross-lane steps (compaction, counts) use 64-bit ballots masked to the
//     worker's lanes, mbcnt-style prefix popcounts and shuffles of width W.
//   * Graph: symmetric CSR in degree-rank ids.  Cliques use the oriented view
//     N+(v) = the suffix of N(v) above v (ostart[v]), so "u > last" is the
//     degree orientation.
//   * Stack levels l = 0..K-3 live in a per-worker LDS slice and spill to a
//     per-worker global arena when a level does not fit.
//   * Motif extend keeps, per level, U_l = sorted union of N(id[0..l]) above
//     id[0], each element with the mask of the lists containing it (bit 7 =
//     passes the canonical rule).  The valid elements E_l are exactly the
//     reference's extensions after canonicalFilter2; the mask bits are the
//     reference's induction bits (forward_induce / aggregate_pattern).
//   * The last level is never materialised: for each child c of level K-3,
//     one pass over N(c) (an LDS membership bitmap of U_{K-3} first; the
//     elements it does not rule out are searched in U_{K-3}) gives, per mask
//     value, how many leaves are adjacent to c, and the class of every leaf
//     follows from (mask, adjacency to c) through the class table.
//   * Work: static per-worker queues of (root, lo, hi) jobs, stealing of
//     other workers' static jobs, and donation of the lowest level's
//     remaining children to a device queue when workers are idle (the
//     reference's donate(), src/main/DuMatoCPU.cu:463-539, done on the device
//     instead of by the host).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "internal.h"

namespace dm {

#define DEV __device__ __forceinline__

// Largest block (4 waves); a launch may use 1, 2 or 4 waves per block.
constexpr int kThreads = 256;
// Waves per SIMD each instantiation is compiled for: 3 (168 VGPRs) for the
// clique and mixed kernels, 4 (128 VGPRs) for the full-wave motif kernel.
// Motif sizes k >= 5 get a second full-wave motif kernel at 6 waves (80
// VGPRs, more of them spilled): their deeper searches wait on more loads, and
// the extra waves hide them (Barabasi-Albert k = 5: 5.16 s at 4 waves, 4.73 s
// at 5, 4.53 s at 6, 4.54 s at 7, 5.52 s at 8), while k = 4 is faster at 4
// (Chung-Lu 18K: 4.66 ms at 4, 4.96 ms at 5).  The clique kernel is slower
// at 4 than at 3 (R-MAT k = 5: 1.42 -> 1.66 ms).
template <int W, int MODE>
constexpr int min_waves() { return W == 64 && MODE == 1 ? 4 : 3; }
constexpr int kDeepMotifWaves = 6;
constexpr int kMaxPm = 32;  // 2^(K-2) mask values for K <= 7

// ---------------------------------------------------------------------------
// virtual-wave primitives
template <int W>
struct Grp {
    int gl;  // lane within the worker
    int sh;  // first wave lane of the worker
    DEV Grp() {
        int l = __lane_id();
        gl = l & (W - 1);
        sh = l & ~(W - 1);
    }
    DEV uint64_t ballot(bool p) const {
        uint64_t b = __ballot(p);
        if constexpr (W == 64) return b;
        else return (b >> sh) & ((1ull << W) - 1);
    }
    DEV int below(uint64_t m) const { return __popcll(m & ((1ull << gl) - 1)); }
    DEV int isum(int x) const {
#pragma unroll
        for (int o = W / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, W);
        return x;
    }
    DEV unsigned long long lsum(unsigned long long x) const {
#pragma unroll
        for (int o = W / 2; o > 0; o >>= 1) x += __shfl_xor(x, o, W);
        return x;
    }
    // src is the same for every lane of the worker at every call site; for a
    // full wave the value is then read into a scalar register
    DEV int bcast(int x, int src) const {
        if constexpr (W == 64) return __builtin_amdgcn_readlane(x, src);
        else return __shfl(x, src, W);
    }
    DEV int iscan(int x) const {  // inclusive prefix sum over the worker's lanes
#pragma unroll
        for (int o = 1; o < W; o <<= 1) {
            int y = __shfl_up(x, o, W);
            if (gl >= o) x += y;
        }
        return x;
    }
    DEV unsigned long long bcast64(unsigned long long x, int src) const {
        if constexpr (W == 64) {
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, src);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), src);
            return ((unsigned long long)hi << 32) | lo;
        } else {
            return __shfl(x, src, W);
        }
    }
};

template <int W>
DEV int uni(int x) {
    if constexpr (W == 64) return __builtin_amdgcn_readfirstlane(x);
    else return x;
}

DEV void wsync() { __threadfence_block(); }

DEV int lower_bound(const int32_t *a, int n, int32_t x) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

DEV int quick_off(int s) { return ((s - 2) * (s + 1)) >> 1; }

DEV uint64_t now_ticks() { return wall_clock64(); }

// ---------------------------------------------------------------------------
// per-worker state in LDS
struct Lvl {
    int32_t *uid;     // U_l ids (motif) / C_l (clique)
    uint8_t *upm;     // U_l masks; nullptr = level 0 (mask 1, valid)
    int32_t *eid;     // E_l ids (motif)
    uint8_t *epm;     // E_l masks; nullptr = level 0
    char *mark_lds;   // allocator tops before this level's buffers
    char *mark_glb;
    uint64_t ind;     // induction (quick bits) of the prefix id[0..l]
    int32_t nu, ne;   // |U_l|, |E_l|
    int32_t cur, end; // children cursor / end
    int32_t vtx;      // vertex placed at level l
    int32_t sdeg;     // sum of degrees of id[0..l] (algorithmic bytes)
};

// Per-worker LDS record, at fixed offsets from one base so that a worker
// holds a single pointer to it: uint64 st[kStats] (the worker's statistics),
// int32 jid[kJobIds] (the prefix of the current job), uint32 n_gt[kMaxPm]
// (motif leaf-phase counters), then Lvl lv[nlv], nlv = max(K-2, 1).
enum { kStBusy, kStJobs, kStFirst, kStStatic, kStLast, kStQCount, kStDonated, kStT0, kStJobMax, kStats };
constexpr int kJidOff = 80;
constexpr int kNgtOff = kJidOff + kJobIds * 4;
constexpr int kLvOff = (kNgtOff + kMaxPm * 4 + 15) & ~15;
static_assert(kStats * 8 <= kJidOff, "worker statistics overlap the job prefix");
struct WorkerLDS {
    char *b;
    DEV unsigned long long *st() const { return (unsigned long long *)b; }
    DEV int32_t *jid() const { return (int32_t *)(b + kJidOff); }
    DEV uint32_t *n_gt() const { return (uint32_t *)(b + kNgtOff); }
    DEV Lvl *lv() const { return (Lvl *)(b + kLvOff); }
};
DEV int nlevels(int K) { return K - 2 > 1 ? K - 2 : 1; }
__host__ __device__ inline size_t worker_lds_bytes(int K) {
    int nlv = K - 2 > 1 ? K - 2 : 1;
    return ((size_t)kLvOff + (size_t)nlv * sizeof(Lvl) + 15) & ~(size_t)15;
}
DEV WorkerLDS worker_lds(char *base, int K) {
    (void)K;
    WorkerLDS w;
    w.b = base;
    return w;
}

template <int W, int MODE>
struct Worker {
    Grp<W> g;
    const EnumParams &p;
    WorkerLDS ws;
    unsigned long long *shist;       // class counters the leaf phase adds to
    const uint16_t *lut;
    int qid;
    int32_t qword;  // the current job's query word (query index | donation threshold)
    int job_sm;     // the current job's donation threshold
    bool qmotif;
    uint32_t epoch;
    char *lds_lo, *lds_top;
    char *g_lo, *g_top;
    int K;
    int32_t v0;
    int job_lvl;
    int steps;
    unsigned long long cnt;    // per-lane partial count of the current job
    unsigned long long bytes;  // per-lane partial algorithmic bytes
    // finished jobs not yet subtracted from the shared outstanding count
    // (lane 0's copy is the one that counts); flushed when the worker runs
    // out of work, so the count never drops below the unfinished jobs
    uint32_t pending;
    int range_hi;  // end of the current range job's roots (v0 + 1 otherwise)
    uint64_t t_start;
    uint64_t last_poll;  // wall clock of the last donation check
    uint32_t polls;      // donation checks so far (lane 0's count)
    bool abort;
    // motif leaf-phase scratch, reserved first in the LDS slice of each job:
    // cs int64[B] | ic u32[B][npm] | nn u32[B] | cc i32[B] | pre i32[B + 1]
    // | qx i32[2W] | qa i32[2W] (lookups waiting for the search of U)
    char *lf;
    // children per leaf batch: a full worker, capped so that the per-child
    // mask counters (slots x 2^(K-2)) stay within 512 words
    DEV static int leaf_slots(int K) {
        const int cap = 512 >> (K - 2);
        return W < cap ? W : cap;
    }
    DEV static int64_t leaf_scratch_bytes(int K) {
        const int B = leaf_slots(K);
        return (int64_t)B * 8 + (int64_t)B * (1 << (K - 2)) * 4 + (int64_t)B * 12 + 4 + 16 * W;
    }

    DEV Worker(const EnumParams &p_, WorkerLDS ws_, unsigned long long *sh_, char *lds_slice,
               int wid)
        : p(p_), ws(ws_), shist(sh_) {
        lds_lo = lds_top = lds_slice;
        g_lo = g_top = p.arena + (int64_t)wid * p.arena_per_worker;
        K = 0;
        qid = -1;
        qword = 0;
        job_sm = p.steal_min;
        qmotif = MODE == 1;
        lut = nullptr;
        epoch = (uint32_t)*p.epoch;
        steps = 0;
        cnt = bytes = 0;
        pending = 0;
        range_hi = 0;
        abort = false;
        t_start = now_ticks();
        last_poll = t_start;
        polls = 0;
    }

    // ---- helpers -----------------------------------------------------------
    DEV int deg(int32_t v) const { return (int)(p.off[v + 1] - p.off[v]); }
    DEV const int32_t *nbr(int32_t v) const { return p.adj + p.off[v]; }

    DEV void fail(int code) {
        if (g.gl == 0) atomicCAS(&p.sc->error, 0, code);
        abort = true;
    }

    DEV char *alloc(int64_t bytes_) {
        int64_t b = (bytes_ + 15) & ~(int64_t)15;
        if (lds_top + b <= lds_lo + p.lds_per_worker) {
            char *r = lds_top;
            lds_top += b;
            return r;
        }
        if (g_top + b <= g_lo + p.arena_per_worker) {
            char *r = g_top;
            g_top += b;
            return r;
        }
        fail(2);
        return nullptr;
    }
    DEV char *alloc_glb(int64_t bytes_) {
        int64_t b = (bytes_ + 15) & ~(int64_t)15;
        if (g_top + b <= g_lo + p.arena_per_worker) {
            char *r = g_top;
            g_top += b;
            return r;
        }
        fail(2);
        return nullptr;
    }

    DEV bool watchdog_expired() {
        return p.watchdog_ticks && (now_ticks() - t_start > p.watchdog_ticks);
    }

    DEV bool check_abort() {
        if ((++steps & 255) == 0) {
            int e = __hip_atomic_load(&p.sc->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (e || watchdog_expired()) {
                if (!e) fail(1);
                abort = true;
            }
        }
        return abort;
    }

    // ---- donation --------------------------------------------------------------
    // Donated jobs are published without fences: each of a job's 16 words is
    // one 64-bit agent-scope atomic store tagged with the launch epoch in the
    // high half; a reader accepts the slot once all 16 tags match.
    // Boundary j (0 <= j <= parts) when [cur, end) is cut into `parts`
    // pieces of equal estimated work.  A clique child i of a list of n
    // candidates roots about (n - i)^e subtrees, e = K - 3 - l; motif
    // children and the leaf level count as equal.  Every piece keeps at least
    // one child (end - cur >= parts).
    DEV int split_at(int l, int cur, int end, int j, int parts) const {
        if (j <= 0) return cur;
        if (j >= parts) return end;
        const bool clique = MODE == 0 || (MODE == 2 && !qmotif);
        const int e = K - 3 - l;
        int m;
        if (!clique || e < 1) {
            m = cur + (int)(((int64_t)(end - cur) * j) / parts);
        } else {
            const float n = (float)ws.lv()[l].nu;
            const float ep = (float)(e + 1);
            const float a = powf(n - (float)cur, ep);
            const float b = powf(n - (float)end, ep);
            const float fr = (float)j / (float)parts;
            m = (int)(n - powf(a - fr * (a - b), 1.0f / ep));
        }
        if (m < cur + j) m = cur + j;
        if (m > end - (parts - j)) m = end - (parts - j);
        return m;
    }

    // Publishes up to P jobs for the children [cur, end) of level lvl: the
    // range is cut into np + 1 pieces of equal estimated work, np <= P as the
    // queue allows, and all but one are queued.  The worker keeps the first
    // piece (upper == true; returns its new end) or the last one (returns its
    // new start).  Returns end (start) unchanged when nothing was queued.
    DEV int push_pieces(int lvl, int cur, int end, int P, bool upper) {
        uint32_t slot = 0;
        const int np = reserve_slots(P, slot);
        if (np == 0) return upper ? end : cur;
        const int first = upper ? 1 : 0;  // first queued piece
        const uint64_t tag = (uint64_t)epoch << 32;
        for (int idx = g.gl; idx < np * kJobWords; idx += W) {
            const int r = idx / kJobWords, j = idx - r * kJobWords;
            int32_t v;
            if (j == 0) v = lvl;
            else if (j == 1) v = split_at(lvl, cur, end, first + r, np + 1);
            else if (j == 2) v = split_at(lvl, cur, end, first + r + 1, np + 1);
            else if (j == 3) v = qword;
            else v = (j - 4 <= lvl) ? ws.lv()[j - 4].vtx : -1;
            __hip_atomic_store(p.q[slot + r].w + j, tag | (uint32_t)v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        if (g.gl == 0) ws.st()[kStDonated] += (unsigned long long)np;
        return split_at(lvl, cur, end, upper ? 1 : np, np + 1);
    }

    // Queues the roots [lo, hi) of this worker's range job as a range job
    // of their own; returns false when the queue is full.
    DEV bool push_range(int lo, int hi) {
        uint32_t slot = 0;
        if (reserve_slots(1, slot) == 0) return false;
        const uint64_t tag = (uint64_t)epoch << 32;
        for (int j = g.gl; j < kJobWords; j += W) {
            const int32_t v = j == 1 ? -1 : j == 2 ? hi : j == 3 ? qword : j == 4 ? lo : j == 0 ? 0 : -1;
            __hip_atomic_store(p.q[slot].w + j, tag | (uint32_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (g.gl == 0) ws.st()[kStDonated]++;
        return true;
    }

    // Reserves up to P consecutive queue slots (as many as the queue has)
    // and counts the jobs that will fill them as outstanding before they
    // become visible, so that outstanding cannot reach zero while they wait;
    // finished jobs not yet subtracted stand in for some of them.  Returns
    // the number reserved (the same on every lane) and the first slot.
    DEV int reserve_slots(int P, uint32_t &slot_out) {
        int np = 0;
        uint32_t slot = 0;
        if (g.gl == 0) {
            const uint32_t t = (uint32_t)__hip_atomic_load(&p.sc->q_ht, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (t < p.q_cap) {
                const uint32_t t0 = (uint32_t)atomicAdd(&p.sc->q_ht, (unsigned long long)P);
                if (t0 < p.q_cap) {
                    np = (int)(p.q_cap - t0) < P ? (int)(p.q_cap - t0) : P;
                    slot = t0;
                    const uint32_t own = pending < (uint32_t)np ? pending : (uint32_t)np;
                    pending -= own;
                    if ((uint32_t)np > own) atomicAdd(&p.sc->outstanding, (unsigned long long)(np - own));
                }
            }
        }
        np = g.bcast(np, 0);
        slot_out = (uint32_t)g.bcast((int)slot, 0);
        return np;
    }

    // Donate remaining children when tickets wait for jobs: the shallowest
    // level of the job with enough children left is cut into one piece per
    // waiting ticket (at most kMaxPieces) plus the piece this worker keeps.
    // leaf_k / leaf_lo / leaf_i describe the descending leaf-phase loop of
    // level leaf_k (remaining children [lo, i]); leaf_k < 0 when not in it.
    static constexpr int kMaxPieces = 8;
    // Waiting tickets (at most kMaxPieces), checked at most every poll
    // period; 0 when stealing is off.  A worker takes a ticket once no
    // static job is left.  The check is a hint: most checks are a
    // workgroup-scope load that may be served from this CU's cache and lag,
    // which only delays a donation; the push itself is exact.
    DEV int donation_demand() {
        if (!p.steal) return 0;
        const uint64_t now = now_ticks();
        if (now - last_poll < p.poll_ticks) return 0;
        last_poll = now;
        int want = 0;
        if (g.gl == 0) {
            // every 4th check reads through to the device scope: a worker
            // whose loop touches little global memory keeps the line in its
            // CU's cache, and a cached copy could hide demand indefinitely
            const unsigned long long ht =
                (++polls & 3) == 0 ? __hip_atomic_load(&p.sc->q_ht, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                   : __hip_atomic_load(&p.sc->q_ht, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const uint32_t h = (uint32_t)(ht >> 32), t = (uint32_t)ht;
            want = h > t ? (int)(h - t < (uint32_t)kMaxPieces ? h - t : (uint32_t)kMaxPieces) : 0;
        }
        return g.bcast(want, 0);
    }

    // min_first > 0: the job's own level may be cut down to pieces of that
    // many children (heavy children, where steal_min would leave too much)
    DEV void maybe_donate(int top, int leaf_k, int &leaf_lo, int leaf_i, int min_first = 0) {
        const int want = donation_demand();
        if (want == 0) return;
        // untouched roots of a range job go first: they need no prefix
        if (range_hi - v0 - 1 >= 2) {
            const int mid = v0 + 1 + ((range_hi - v0 - 1) >> 1);
            if (push_range(mid, range_hi)) range_hi = mid;
            return;
        }
        // the last level's children are cheap: split it only in bigger pieces
        const int leaf_min = 4 * job_sm;
        for (int l = job_lvl; l < top; l++) {
            const int cur = ws.lv()[l].cur, end = ws.lv()[l].end;
            const int minc = (min_first > 0 && l == job_lvl) ? min_first : l == K - 3 ? leaf_min : job_sm;
            if (end - cur >= minc) {
                int P = (end - cur) / minc - 1;
                if (P < 1) P = 1;
                if (P > want) P = want;
                if (P > end - cur - 1) P = end - cur - 1;
                if (P >= 1) ws.lv()[l].end = push_pieces(l, cur, end, P, true);
                return;
            }
        }
        const int left = leaf_i + 1 - leaf_lo;
        if (leaf_k >= 0 && left >= leaf_min) {
            // the leaf loop runs downward: queue the lower pieces
            int P = left / leaf_min - 1;
            if (P < 1) P = 1;
            if (P > want) P = want;
            if (P > left - 1) P = left - 1;
            if (P >= 1) leaf_lo = push_pieces(leaf_k, leaf_lo, leaf_i + 1, P, false);
        }
    }

    // ---- motif levels ----------------------------------------------------------
    DEV void motif_root_level(int32_t v) {
        Lvl &L = ws.lv()[0];
        const int32_t *a = p.adj + p.ostart[v];
        int n = (int)(p.off[v + 1] - p.ostart[v]);
        L.uid = const_cast<int32_t *>(a);
        L.upm = nullptr;
        L.eid = const_cast<int32_t *>(a);
        L.epm = nullptr;
        L.nu = n;
        L.ne = n;
        L.vtx = v;
        L.ind = 0;
        L.sdeg = deg(v);
        L.mark_lds = lds_top;
        L.mark_glb = g_top;
    }

    DEV uint8_t upm_at(const Lvl &L, int j) const { return L.upm ? L.upm[j] : (uint8_t)0x81; }
    DEV int epm_at(const Lvl &L, int j) const { return L.epm ? (int)L.epm[j] : 1; }

    // U_{l+1}, E_{l+1} from level l and child c (pmc = c's mask in U_l).
    DEV void motif_forward(int l, int32_t c, int pmc) {
        Lvl &P = ws.lv()[l];
        Lvl &N = ws.lv()[l + 1];
        const int l1 = l + 1;
        const int32_t *nc = nbr(c);
        const int dc = deg(c);
        const int s0 = lower_bound(nc, dc, v0 + 1);
        const int q = dc - s0;
        const int pnu = P.nu;
        N.mark_lds = lds_top;
        N.mark_glb = g_top;
        const int cap = pnu + q;
        // At k >= 5 the leaf level's lists (the union of three or more
        // neighbourhoods) go to global memory when they would take more than
        // half the free LDS: the leaf phase reads them rarely and needs the
        // LDS for its membership filter.  (At k = 4 they stay in LDS: short
        // lists, and moving them measured slower.)
        const bool lg = l1 == K - 3 && l1 >= 2 &&
                        (int64_t)cap * 10 > (lds_lo + p.lds_per_worker - lds_top) / 2;
        int32_t *uid = (int32_t *)(lg ? alloc_glb((int64_t)cap * 4) : alloc((int64_t)cap * 4));
        uint8_t *upm = (uint8_t *)(lg ? alloc_glb(cap) : alloc(cap));
        int32_t *eid = (int32_t *)(lg ? alloc_glb((int64_t)cap * 4) : alloc((int64_t)cap * 4));
        uint8_t *epm = (uint8_t *)(lg ? alloc_glb(cap) : alloc(cap));
        char *tmark_l = lds_top, *tmark_g = g_top;
        int32_t *nw = (int32_t *)alloc((int64_t)(q > 0 ? q : 1) * 4);
        uint8_t *hit = (uint8_t *)alloc(pnu > 0 ? pnu : 1);
        if (abort) return;
        for (int j = g.gl; j < pnu; j += W) hit[j] = 0;
        wsync();
        // scan N(c) above v0: mark hits in U_l, collect new vertices (sorted)
        int nnew = 0;
        for (int base = 0; base < q; base += W) {
            int t = base + g.gl;
            bool in = t < q;
            int32_t x = in ? nc[s0 + t] : 0;
            bool found = false;
            if (in) {
                int pos = lower_bound(P.uid, pnu, x);
                found = pos < pnu && P.uid[pos] == x;
                if (found) hit[pos] = 1;
            }
            uint64_t m = g.ballot(in && !found);
            if (in && !found) nw[nnew + g.below(m)] = x;
            nnew += __popcll(m);
        }
        wsync();
        // merge by ranks: old elements and new ones interleave in id order
        const uint8_t bit = (uint8_t)(1u << l1);
        for (int j = g.gl; j < pnu; j += W) {
            int32_t x = P.uid[j];
            int o = j + lower_bound(nw, nnew, x);
            uint8_t pm = upm_at(P, j);
            if (hit[j]) pm |= bit;
            bool valid = (pm & 0x80) && x > c;
            uid[o] = x;
            upm[o] = (uint8_t)((pm & 0x7f) | (valid ? 0x80 : 0));
        }
        for (int t = g.gl; t < nnew; t += W) {
            int32_t x = nw[t];
            int o = t + lower_bound(P.uid, pnu, x);
            uid[o] = x;
            upm[o] = (uint8_t)(bit | 0x80);
        }
        const int nu = pnu + nnew;
        wsync();
        int ne = 0;
        for (int base = 0; base < nu; base += W) {
            int j = base + g.gl;
            bool in = j < nu;
            uint8_t pm = in ? upm[j] : 0;
            bool v = in && (pm & 0x80);
            uint64_t m = g.ballot(v);
            if (v) {
                int o = ne + g.below(m);
                eid[o] = uid[j];
                epm[o] = (uint8_t)(pm & 0x7f);
            }
            ne += __popcll(m);
        }
        wsync();
        // release the temporaries
        lds_top = tmark_l;
        g_top = tmark_g;
        N.uid = uid;
        N.upm = upm;
        N.eid = eid;
        N.epm = epm;
        N.nu = nu;
        N.ne = ne;
        N.vtx = c;
        N.sdeg = P.sdeg + dc;
        N.ind = (l1 >= 2) ? P.ind + ((uint64_t)pmc << quick_off(l1)) : 0;
    }

    // children [lo, hi) of level l = K-3, each a node of level K-2 whose
    // leaves are counted by class without being materialised.  Children are
    // taken in descending order, in batches of up to W; the batch's
    // (child, neighbour) pairs are flattened across the lanes so short
    // neighbour lists do not leave lanes idle.
    DEV void motif_leaf_phase(int l, int lo, int hi) {
        Lvl &P = ws.lv()[l];
        const int npm = 1 << (l + 1);  // mask values over the lists 0..l
        const int top = npm;           // bit of the child's own list
        const int s = quick_off(l + 2);
        const int o1 = quick_off(l + 1);
        const int B = leaf_slots(K);
        int64_t *cs = (int64_t *)lf;
        uint32_t *ic = (uint32_t *)(cs + B);
        uint32_t *nn = ic + B * npm;
        int32_t *cc = (int32_t *)(nn + B);
        int32_t *pre = cc + B;
        int32_t *qx = pre + B + 1;
        int32_t *qa = qx + 2 * W;
        // ws.n_gt()[m]: valid elements of E above the current batch, by mask
        for (int j = g.gl; j < npm; j += W) ws.n_gt()[j] = 0;
        for (int j = g.gl; j < B * npm; j += W) ic[j] = 0;
        for (int j = g.gl; j < B; j += W) nn[j] = 0;
        wsync();
        for (int j = hi + g.gl; j < P.ne; j += W) atomicAdd(&ws.n_gt()[epm_at(P, j)], 1u);
        wsync();
        const int pnu = P.nu;
        // Fences of U, one per lane: fence(f) = U[f * pnu / W].  A lookup
        // finds its fence interval by a search over the lanes' registers
        // (cross-lane permutes, free of LDS bank conflicts), then searches
        // only that interval of U.
        const bool use_f = pnu >= 2 * W;
        const int32_t fence = use_f ? P.uid[(int)(((int64_t)g.gl * pnu) / W)] : 0;
        const int32_t flast = __shfl(fence, W - 1, W);  // read by every lane
        // Membership filter of U in LDS: bit (x mod 32*words) is set for every
        // x in U, with at least 8 bits per element where LDS allows.  Most
        // neighbours of a child are not in U; a clear bit proves it with one
        // LDS read, and only the rest go through the search of U.
        uint32_t *bm = nullptr;
        int32_t bmask = 0;
        {
            int words = 32;
            while (words < 1024 && words * 4 < pnu) words <<= 1;
            while (words >= 32 && lds_top + words * 4 > lds_lo + p.lds_per_worker) words >>= 1;
            if (words >= 32) {
                bm = (uint32_t *)alloc((int64_t)words * 4);
                bmask = words * 32 - 1;
                for (int j = g.gl; j < words; j += W) bm[j] = 0;
                wsync();
                for (int j = g.gl; j < pnu; j += W) {
                    const int32_t u = P.uid[j] & bmask;
                    atomicOr(&bm[u >> 5], 1u << (u & 31));
                }
                wsync();
            }
        }
        // search of U for queued lookups (all lanes take part in the fence
        // permutes; `act` lanes hold an entry): a hit adds to the child's
        // class counter, a miss is a new element of the child
        auto resolve = [&](int32_t x, int a, bool act) {
            int flo = 0, fhi = pnu;  // interval of U holding lower_bound(x)
            if (use_f) {
                int c = 0;  // number of fences below x
#pragma unroll
                for (int st = W / 2; st >= 1; st >>= 1)
                    if (__shfl(fence, c + st - 1, W) < x) c += st;
                if (c == W - 1 && flast < x) c = W;
                if (c == 0) {
                    fhi = 0;
                } else {
                    flo = (int)(((int64_t)(c - 1) * pnu) / W) + 1;
                    fhi = c < W ? (int)(((int64_t)c * pnu) / W) : pnu;
                }
            }
            if (act) {
                const int pos = flo + lower_bound(P.uid + flo, fhi - flo, x);
                if (pos < pnu && P.uid[pos] == x) {
                    const uint8_t pm = upm_at(P, pos);
                    if ((pm & 0x80) && x > cc[a]) atomicAdd(&ic[a * npm + (pm & 0x7f)], 1u);
                } else {
                    atomicAdd(&nn[a], 1u);
                }
            }
        };
        // a batch takes children (at most B) until their neighbour lists hold
        // about `target` elements: few, large batches amortise the per-batch
        // class bookkeeping; the bound keeps the leaf donation checks, one
        // per batch, frequent when the children are hubs
        const int target = 256 * W;
        int i = hi - 1;
        while (i >= lo) {
            if (check_abort()) return;
            // leaf children are cheap and a donated piece rebuilds its whole
            // prefix: offer them only from long leaf loops
            if (i - lo + 1 >= 512) maybe_donate(l, l, lo, i);
            if (i < lo) break;
            const int avail = i - lo + 1 < B ? i - lo + 1 : B;
            int len = 0, pm_a = -1;
            if (g.gl < avail) {
                const int32_t c = P.eid[i - g.gl];
                const int dc = deg(c);
                const int s0 = lower_bound(nbr(c), dc, v0 + 1);
                len = dc - s0;
                cc[g.gl] = c;
                cs[g.gl] = p.off[c] + s0;
                pm_a = epm_at(P, i - g.gl);
            }
            const int sc = g.iscan(len);
            int take = __popcll(g.ballot(g.gl < avail && sc - len < target));
            take = take < 1 ? 1 : take;
            if (g.gl < take) pre[g.gl + 1] = sc;
            if (g.gl == 0) pre[0] = 0;
            const int tot = g.bcast(sc, take - 1);
            wsync();
            int a = 0;
            int qn = 0;  // queued lookups (uniform)
            for (int t0 = 0; t0 < tot; t0 += W) {
                const int t = t0 + g.gl;
                int32_t x = 0x7fffffff;
                if (t < tot) {
                    while (pre[a + 1] <= t) a++;
                    x = p.adj[cs[a] + (t - pre[a])];
                }
                bool need = t < tot;
                if (bm && need) need = (bm[(x & bmask) >> 5] >> (x & 31)) & 1u;
                const bool fresh = t < tot && !need;  // proven outside U
                // new elements are the common case: count them per child with
                // one ballot, the child's first lane in this round adding its
                // segment's popcount (no same-address atomics)
                const uint64_t mf = g.ballot(fresh);
                if (t < tot) {
                    const int s0 = pre[a] > t0 ? pre[a] - t0 : 0;
                    if (g.gl == s0) {
                        const int s1 = pre[a + 1] - t0 < W ? pre[a + 1] - t0 : W;
                        const uint64_t seg = (s1 - s0 >= 64 ? ~0ull : ((1ull << (s1 - s0)) - 1)) << s0;
                        nn[a] += (uint32_t)__popcll(mf & seg);
                    }
                }
                // the rest wait in the queue until a full wave of them can be
                // searched together
                const uint64_t mq = g.ballot(need);
                if (need) {
                    const int o = qn + g.below(mq);
                    qx[o] = x;
                    qa[o] = a;
                }
                qn += __popcll(mq);
                if (qn >= W) {
                    wsync();
                    qn -= W;
                    resolve(qx[qn + g.gl], qa[qn + g.gl], true);
                    wsync();
                }
            }
            if (qn > 0) {
                wsync();
                const bool act = g.gl < qn;
                resolve(act ? qx[g.gl] : 0x7fffffff, act ? qa[g.gl] : 0, act);
            }
            wsync();
            // per child: slot a = lane a; valid elements above it = the
            // batch base plus earlier slots (larger E indices) with that mask
            const bool mine = g.gl < take;
            const uint64_t indc = P.ind + ((mine && l + 1 >= 2) ? ((uint64_t)pm_a << o1) : 0);
            unsigned long long leaves = 0;
            for (int j = 0; j < npm; j++) {
                const uint64_t mj = g.ballot(mine && pm_a == j);
                if (mine) {
                    const uint32_t gtj = ws.n_gt()[j] + (uint32_t)g.below(mj);
                    const uint32_t ai = ic[g.gl * npm + j];
                    const uint32_t bi = gtj - ai;
                    if (ai) atomicAdd(&shist[lut[indc + ((uint64_t)(j | top) << s)]], (unsigned long long)ai);
                    if (bi) atomicAdd(&shist[lut[indc + ((uint64_t)j << s)]], (unsigned long long)bi);
                    ic[g.gl * npm + j] = 0;
                    leaves += gtj;
                }
                wsync();
                if (g.gl == 0) ws.n_gt()[j] += (uint32_t)__popcll(mj);
            }
            if (mine) {
                const uint32_t nw = nn[g.gl];
                if (nw) atomicAdd(&shist[lut[indc + ((uint64_t)top << s)]], (unsigned long long)nw);
                nn[g.gl] = 0;
                leaves += nw;
                cnt += leaves;
                bytes += 4ull * (unsigned long long)(P.sdeg + deg(cc[g.gl])) + 4ull * (K - 1) * leaves;
            }
            wsync();
            i -= take;
        }
    }

    // ---- clique levels ----------------------------------------------------------
    DEV void clique_root_level(int32_t v) {
        Lvl &L = ws.lv()[0];
        L.uid = const_cast<int32_t *>(p.adj + p.ostart[v]);
        L.nu = (int)(p.off[v + 1] - p.ostart[v]);
        L.vtx = v;
        L.mark_lds = lds_top;
        L.mark_glb = g_top;
    }

    DEV const int32_t *onbr(int32_t v) const { return p.adj + p.ostart[v]; }
    DEV int odeg(int32_t v) const { return (int)(p.off[v + 1] - p.ostart[v]); }

    // Intersection of two sorted lists: stride the shorter one, binary-search
    // the longer.  WRITE: compact survivors (sorted) into out and return the
    // count; otherwise return this lane's partial count.
    template <bool WRITE>
    DEV int intersect(const int32_t *A, int na, const int32_t *B, int nb, int32_t *out) {
        const int32_t *S = A, *L = B;
        int ns = na, nl = nb;
        if (na > nb) {
            S = B; L = A; ns = nb; nl = na;
        }
        if (ns == 0 || nl == 0) return 0;
        if (S[ns - 1] < L[0] || L[nl - 1] < S[0]) return 0;
        int w = 0;
        for (int base = 0; base < ns; base += W) {
            int t = base + g.gl;
            bool in = t < ns;
            bool found = false;
            int32_t x = 0;
            if (in) {
                x = S[t];
                int pos = lower_bound(L, nl, x);
                found = pos < nl && L[pos] == x;
            }
            if constexpr (WRITE) {
                uint64_t m = g.ballot(found);
                if (found) out[w + g.below(m)] = x;
                w += __popcll(m);
            } else {
                w += found ? 1 : 0;
            }
        }
        return w;
    }

    DEV void clique_forward(int l, int i) {
        Lvl &P = ws.lv()[l];
        Lvl &N = ws.lv()[l + 1];
        const int32_t c = P.uid[i];
        const int32_t *B = P.uid + i + 1;
        const int nb = P.nu - i - 1;
        const int na = odeg(c);
        N.mark_lds = lds_top;
        N.mark_glb = g_top;
        const int cap = na < nb ? na : nb;
        int32_t *out = (int32_t *)alloc((int64_t)(cap > 0 ? cap : 1) * 4);
        if (abort) return;
        int w = intersect<true>(onbr(c), na, B, nb, out);
        wsync();
        N.uid = out;
        N.nu = w;
        N.vtx = c;
    }

    DEV void clique_leaf_phase(int l, int lo, int hi) {
        Lvl &P = ws.lv()[l];
        for (int i = lo; i < hi; i++) {
            if (check_abort()) return;
            int dummy_lo = 0;
            (void)dummy_lo;
            // donation of this level's tail happens through lv[l].cur/end
            P.cur = i + 1;
            maybe_donate(l + 1, -1, dummy_lo, 0);
            hi = P.end;
            if (i >= hi) break;
            const int32_t c = P.uid[i];
            const int na = odeg(c);
            int part = intersect<false>(onbr(c), na, P.uid + i + 1, P.nu - i - 1, nullptr);
            cnt += (unsigned long long)part;
            if (g.gl == 0) bytes += 4ull * (unsigned long long)na;
        }
    }

    // ---- clique, local bitmap of the root's out-neighbourhood -------------
    // L = N+(v), d = |L| <= 64*NW.  rows[i*NW + w] bit b: L[64w+b] in N+(L[i]).
    // Each lane takes first-level children i and walks the deeper levels as
    // AND/popcount over rows; the loops visit exactly the DFS children of
    // every node above the last level, so the algorithmic bytes (sum of
    // out-degrees of the children of nodes with 1..k-2 vertices) are summed
    // in the same loops from dp[] = out-degrees.
    // Work of one second-level pair (i, j), j in row i: the (k-2)-level
    // candidate set S2 = row_i & row_j, then the deeper levels by AND/popcount.
    template <int NW>
    DEV void bitmap_pair(const uint64_t *rows, const int32_t *dp, int i, int j, unsigned long long &c,
                         unsigned long long &by) const {
        uint64_t S2[NW];
#pragma unroll
        for (int w = 0; w < NW; w++) S2[w] = rows[i * NW + w] & rows[j * NW + w];
        if (K == 4) {
            int pc = 0;
#pragma unroll
            for (int w = 0; w < NW; w++) pc += __popcll(S2[w]);
            c += (unsigned long long)pc;
            return;
        }
#pragma unroll
        for (int w2 = 0; w2 < NW; w2++) {
            uint64_t m2 = S2[w2];
            while (m2) {
                const int k = w2 * 64 + __builtin_ctzll(m2);
                m2 &= m2 - 1;
                by += (unsigned long long)dp[k];
                if (K == 5) {
                    int pc = 0;
#pragma unroll
                    for (int w = 0; w < NW; w++) pc += __popcll(S2[w] & rows[k * NW + w]);
                    c += (unsigned long long)pc;
                    continue;
                }
                uint64_t S3[NW];
#pragma unroll
                for (int w = 0; w < NW; w++) S3[w] = S2[w] & rows[k * NW + w];
                int pc = 0;
#pragma unroll
                for (int w3 = 0; w3 < NW; w3++) {
                    uint64_t m3 = S3[w3];
                    while (m3) {
                        const int q = w3 * 64 + __builtin_ctzll(m3);
                        m3 &= m3 - 1;
                        by += (unsigned long long)dp[q];
#pragma unroll
                        for (int w = 0; w < NW; w++) pc += __popcll(S3[w] & rows[q * NW + w]);
                    }
                }
                c += (unsigned long long)pc;
            }
        }
    }

    // position of the r-th set bit (0-based) of m; m has more than r bits
    DEV static int select64(uint64_t m, int r) {
        int pos = 0;
#pragma unroll
        for (int sh = 32; sh >= 1; sh >>= 1) {
            const int c = __popcll(m & ((1ull << sh) - 1));
            if (r >= c) {
                r -= c;
                m >>= sh;
                pos += sh;
            }
        }
        return pos;
    }

    template <int NW>
    DEV void clique_bitmap_root(int32_t v, int lo, int hi, bool first_piece, bool shared) {
        const int d = odeg(v);
        const int32_t *L = onbr(v);
        Lvl &R = ws.lv()[0];
        R.uid = const_cast<int32_t *>(L);
        R.nu = d;
        R.vtx = v;
        R.mark_lds = lds_top;
        R.mark_glb = g_top;
        const int jobs_on_root = hi;  // shared: the root's job count
        if (shared) {
            lo = 0;
            hi = d;
        }
        if (hi > d) hi = d;
        if (lo >= hi) return;
        // pair-phase data first, so that it gets the LDS
        uint64_t *rows = (uint64_t *)alloc((int64_t)d * NW * 8);
        int32_t *dp = (int32_t *)alloc((int64_t)d * 4);
        int32_t *pre = (int32_t *)alloc((int64_t)(d + 1) * 4);
        if (abort) return;
        // a root whose bitmap the launch prebuilt (a heavy root cut into
        // many pieces): copy its rows instead of building them
        int slot = -1;
        if (p.npre > 0) {
            int a = 0, b = p.npre;
            while (a < b) {
                const int mid = (a + b) >> 1;
                if (p.pre_roots[mid] < v) a = mid + 1;
                else b = mid;
            }
            if (a < p.npre && p.pre_roots[a] == v) slot = a;
        }
        if (slot >= 0) {
            const uint64_t *gr = (const uint64_t *)(p.pre + p.pre_offs[slot]);
            const int32_t *gd = (const int32_t *)(gr + (int64_t)d * NW);
            for (int t = lo * NW + g.gl; t < d * NW; t += W) rows[t] = gr[t];
            for (int i = g.gl; i < d; i += W) dp[i] = gd[i];
            wsync();
        } else if (shared) {
            fail(3);  // the host gives a shared cursor only to prebuilt roots
            return;
        } else {
            build_rows<NW>(v, lo, d, L, rows, dp, pre);
            if (abort) return;
        }
        // k = 6: a round's pair staging, allocated after the build so that
        // the build's hash gets the LDS first (both fit there together only
        // for smaller roots)
        uint64_t *s2 = nullptr;
        int32_t *sp = nullptr;
        if (K == 6) {
            s2 = (uint64_t *)alloc((int64_t)W * NW * 8);
            sp = (int32_t *)alloc((int64_t)(W + 1) * 4);
            if (abort) return;
        }
        if (shared) {
            // guided self-scheduling with the root's other jobs: each claim
            // takes about 1 / (2 x jobs on the root) of the children left, at
            // least one, so the root's last children are spread thinly
            // whatever their weights; the claim that starts at child 0 counts
            // the root's bytes.  (Claims sized by estimated work, pop^(k-3)
            // per child with a compare-and-swap cursor, balanced better but
            // their extra claims cost more: R-MAT k = 6 4.78 -> 5.0 ms.)
            int32_t *cursor = p.pre_cursor + slot;
            const int share = 2 * (jobs_on_root > 1 ? jobs_on_root : 1);
            while (!abort) {
                int a = 0, n = 1;
                if (g.gl == 0) {
                    const int cur = __hip_atomic_load(cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    n = d > cur ? (d - cur + share - 1) / share : 1;
                    a = atomicAdd(cursor, n);
                }
                a = g.bcast(a, 0);
                n = g.bcast(n, 0);
                if (a >= d) break;
                if (a == 0 && g.gl == 0) bytes += 4ull * (unsigned long long)d;
                pair_phase<NW>(R, rows, dp, pre, s2, sp, a, a + n < d ? a + n : d);
            }
        } else {
            if (first_piece && g.gl == 0) bytes += 4ull * (unsigned long long)d;
            pair_phase<NW>(R, rows, dp, pre, s2, sp, lo, hi);
        }
        lds_top = R.mark_lds;
        g_top = R.mark_glb;
    }

    // Rows lo..d-1 of the local bitmap of root v (L = N+(v), d = |L|): row i
    // bit j set when L[j] is in N+(L[i]); dp[i] = out-degree of L[i].
    template <int NW>
    DEV void build_rows(int32_t v, int lo, int d, const int32_t *L, uint64_t *rows, int32_t *dp, int32_t *pre) {
        (void)v;
        // build scratch, released after the build: a hash of N+(v) (id ->
        // position, load factor <= 1/2) and each child's list start
        char *const build_lds = lds_top;
        char *const build_glb = g_top;
        constexpr int HB = NW == 1 ? 7 : (NW == 2 ? 8 : 9);
        constexpr int HS = 1 << HB;
        int32_t *hkey = (int32_t *)alloc(HS * 4);
        uint8_t *hval = (uint8_t *)alloc(HS);
        int64_t *rst = (int64_t *)alloc((int64_t)d * 8);
        if (abort) return;
        for (int t = g.gl; t < HS; t += W) hkey[t] = -1;
        for (int i = g.gl; i < d; i += W) {
            const int32_t u = L[i];
            const int64_t os = p.ostart[u];
            rst[i] = os;
            dp[i] = (int)(p.off[u + 1] - os);
        }
        for (int t = lo * NW + g.gl; t < d * NW; t += W) rows[t] = 0;
        wsync();
        // only children above lo can be set in rows lo.. (lists are sorted)
        for (int j = lo + g.gl; j < d; j += W) {
            const int32_t key = L[j];
            uint32_t h = ((uint32_t)key * 2654435761u) >> (32 - HB);
            while (true) {
                if (atomicCAS(&hkey[h], -1, key) == -1) {
                    hval[h] = (uint8_t)j;
                    break;
                }
                h = (h + 1) & (HS - 1);
            }
        }
        // flatten (i, x in N+(L[i])) for i in [lo, d) and set the bits
        int carry = 0;
        for (int base = lo; base < d; base += W) {
            const int i = base + g.gl;
            const int x = i < d ? dp[i] : 0;
            const int sc = g.iscan(x);
            if (i < d) pre[i - lo + 1] = carry + sc;
            carry += g.bcast(sc, W - 1);
        }
        if (g.gl == 0) pre[0] = 0;
        wsync();
        const int T = carry;
        int a = 0;  // row of this lane's element; t only grows
        constexpr int U = 4;
        for (int t0 = 0; t0 < T; t0 += U * W) {
            int32_t xs[U];
            int ra[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = t0 + u * W + g.gl;
                ra[u] = -1;
                xs[u] = 0;
                if (t < T) {
                    while (pre[a + 1] <= t) a++;
                    ra[u] = a;
                    xs[u] = p.adj[rst[lo + a] + (t - pre[a])];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (ra[u] < 0) continue;
                const int32_t x = xs[u];
                uint32_t h = ((uint32_t)x * 2654435761u) >> (32 - HB);
                int k;
                while ((k = hkey[h]) != x && k != -1) h = (h + 1) & (HS - 1);
                if (k == x) {
                    const int j = hval[h];
                    atomicOr(&rows[(lo + ra[u]) * NW + (j >> 6)], 1ull << (j & 63));
                }
            }
        }
        wsync();
        lds_top = build_lds;
        g_top = build_glb;
    }

    // first-level children [lo, hi) in batches of W; each batch is
    // flattened into its (i, j) pairs, which the lanes stride over
    template <int NW>
    // s2 / sp (k = 6 only): LDS staging of a round's W pairs, W x NW words
    // and W + 1 prefix entries
    DEV void pair_phase(Lvl &R, const uint64_t *rows, const int32_t *dp, int32_t *pre, uint64_t *s2, int32_t *sp,
                        int lo, int hi) {
        R.cur = lo;
        R.end = hi;
        unsigned long long by = 0;
        while (true) {
            if (check_abort()) break;
            int dummy = 0;
            // at k >= 6 a front child with many pairs (its work grows with
            // their cube) is heavy enough to travel alone
            int front = 0;
            if (K >= 6 && g.gl == 0 && R.cur < R.end) {
#pragma unroll
                for (int w = 0; w < NW; w++) front += __popcll(rows[R.cur * NW + w]);
            }
            front = g.bcast(front, 0);
            maybe_donate(1, -1, dummy, 0, front >= W / 2 ? 2 : 0);
            const int cur = R.cur, end = R.end;
            if (cur >= end) break;
            // a batch holds children until it has about `target` pairs (at
            // least one child), so heavy roots reach a donation point often
            const int avail = end - cur < W ? end - cur : W;
            int np = 0;
            if (g.gl < avail) {
#pragma unroll
                for (int w = 0; w < NW; w++) np += __popcll(rows[(cur + g.gl) * NW + w]);
            }
            const int sc = g.iscan(np);
            const int target = K == 4 ? 16 * W : 2 * W;
            int take = __popcll(g.ballot(g.gl < avail && sc - np < target));
            take = take < 1 ? 1 : take;
            R.cur = cur + take;
            if (g.gl < take) {
                by += (unsigned long long)dp[cur + g.gl];
                pre[g.gl + 1] = sc;
            }
            if (g.gl == 0) pre[0] = 0;
            const int P = g.bcast(sc, take - 1);
            wsync();
            if (K == 6) {
                pair_round_k6<NW>(rows, dp, pre, s2, sp, cur, P, by);
                continue;
            }
            int a = 0;
            for (int t = g.gl; t < P; t += W) {
                while (pre[a + 1] <= t) a++;
                const int i = cur + a;
                const int j = pair_second<NW>(rows, i, t - pre[a]);
                by += (unsigned long long)dp[j];
                bitmap_pair<NW>(rows, dp, i, j, cnt, by);
            }
            wsync();
        }
        bytes += 4ull * by;
    }

    // the second child of pair r of child i: the r-th set bit of row i
    template <int NW>
    DEV static int pair_second(const uint64_t *rows, int i, int r) {
        int w = 0;
        uint64_t m = rows[i * NW];
        while (true) {
            const int c = __popcll(m);
            if (r < c || w == NW - 1) break;
            r -= c;
            w++;
            m = rows[i * NW + w];
        }
        return w * 64 + select64(m, r);
    }

    // k = 6: the pairs of a batch, W at a time.  Each lane stages its pair's
    // common candidates S2 = row i & row j in LDS, then the lanes stride over
    // the round's (pair, k in S2) triples, each counting the 2-cliques of
    // S2 & row k above k: one level finer than a lane per pair, whose cost
    // (|S2| x |S3|) varies too much across the lanes of a wave.
    template <int NW>
    DEV void pair_round_k6(const uint64_t *rows, const int32_t *dp, const int32_t *pre, uint64_t *s2, int32_t *sp,
                           int cur, int P, unsigned long long &by) {
        int a = 0;
        for (int t0 = 0; t0 < P; t0 += W) {
            const int t = t0 + g.gl;
            int s = 0;
            if (t < P) {
                while (pre[a + 1] <= t) a++;
                const int i = cur + a;
                const int j = pair_second<NW>(rows, i, t - pre[a]);
                by += (unsigned long long)dp[j];
#pragma unroll
                for (int w = 0; w < NW; w++) {
                    const uint64_t x = rows[i * NW + w] & rows[j * NW + w];
                    s2[g.gl * NW + w] = x;
                    s += __popcll(x);
                }
            }
            const int ps = g.iscan(s);
            sp[g.gl + 1] = ps;
            if (g.gl == 0) sp[0] = 0;
            const int Ts = g.bcast(ps, W - 1);
            wsync();
            int pa = 0;
            unsigned long long c = 0;
            for (int u = g.gl; u < Ts; u += W) {
                while (sp[pa + 1] <= u) pa++;
                const uint64_t *S2 = s2 + pa * NW;
                const int k = pair_second<NW>(S2, 0, u - sp[pa]);
                by += (unsigned long long)dp[k];
                uint64_t S3[NW];
#pragma unroll
                for (int w = 0; w < NW; w++) S3[w] = S2[w] & rows[k * NW + w];
                int pc = 0;
#pragma unroll
                for (int w3 = 0; w3 < NW; w3++) {
                    uint64_t m3 = S3[w3];
                    while (m3) {
                        const int q = w3 * 64 + __builtin_ctzll(m3);
                        m3 &= m3 - 1;
                        by += (unsigned long long)dp[q];
#pragma unroll
                        for (int w = 0; w < NW; w++) pc += __popcll(S3[w] & rows[q * NW + w]);
                    }
                }
                c += (unsigned long long)pc;
            }
            cnt += c;
            wsync();
        }
    }

    // ---- one job ------------------------------------------------------------
    // children [lo, hi) of the node whose prefix ws.jid()[0..lvl] is in LDS;
    // first_piece: a static job that starts at the root's first child (the
    // root node's own bytes are counted once, by that job); shared: one of
    // `hi` jobs on a prebuilt bitmap root that claim its children from the
    // root's shared cursor
    DEV void run_job(int lvl, int lo, int hi, bool first_piece, bool shared) {
        v0 = ws.jid()[0];
        job_lvl = lvl;
        lds_top = lds_lo;
        g_top = g_lo;
        cnt = 0;
        const bool motif = MODE == 2 ? qmotif : MODE == 1;
        if (motif) {
            lf = alloc(leaf_scratch_bytes(K));
            if (abort) return;
        }
        if (!motif && K == 2) {
            if (g.gl == 0) {
                int d = odeg(v0);
                cnt += (unsigned long long)d;
                bytes += 4ull * (unsigned long long)d;
            }
            return;
        }
        if (!motif && p.bitmap && lvl == 0 && K >= 4 && K <= 6) {
            const int d = odeg(v0);
            if (d <= 64) {
                clique_bitmap_root<1>(v0, lo, hi, first_piece, shared);
                return;
            }
            if (d <= 128) {
                clique_bitmap_root<2>(v0, lo, hi, first_piece, shared);
                return;
            }
            if (d <= 256) {
                clique_bitmap_root<4>(v0, lo, hi, first_piece, shared);
                return;
            }
        }
        if (shared) {  // only prebuilt bitmap roots have a shared cursor
            fail(3);
            return;
        }
        // rebuild the prefix levels 0..lvl
        if (motif) motif_root_level(v0);
        else clique_root_level(v0);
        if (first_piece && g.gl == 0)
            bytes += 4ull * (unsigned long long)(motif ? deg(v0) : odeg(v0));
        for (int l = 1; l <= lvl && !abort; l++) {
            const int32_t c = ws.jid()[l];
            Lvl &P = ws.lv()[l - 1];
            int pos = lower_bound(P.uid, P.nu, c);
            if (pos >= P.nu || P.uid[pos] != c) {
                fail(3);
                return;
            }
            if (motif) motif_forward(l - 1, c, upm_at(P, pos) & 0x7f);
            else clique_forward(l - 1, pos);
        }
        if (abort) return;
        const int leafp = K - 3;
        {
            Lvl &L = ws.lv()[lvl];
            int ne = motif ? L.ne : L.nu;
            L.cur = lo < ne ? lo : ne;
            L.end = hi < ne ? hi : ne;
        }
        int k = lvl;
        while (k >= lvl) {
            if (check_abort()) return;
            Lvl &L = ws.lv()[k];
            if (k == leafp) {
                if (motif) motif_leaf_phase(k, L.cur, L.end);
                else clique_leaf_phase(k, L.cur, L.end);
                if (abort) return;
                lds_top = L.mark_lds;
                g_top = L.mark_glb;
                k--;
                continue;
            }
            int dummy = 0;
            maybe_donate(k + 1, -1, dummy, 0);
            if (L.cur < L.end) {
                const int i = L.cur;
                L.cur = i + 1;
                if (motif) {
                    const int32_t c = L.eid[i];
                    const int pmc = epm_at(L, i);
                    motif_forward(k, c, pmc);
                    if (abort) return;
                    Lvl &N = ws.lv()[k + 1];
                    if (g.gl == 0) bytes += 4ull * (unsigned long long)N.sdeg;
                    N.cur = 0;
                    N.end = N.ne;
                } else {
                    if (g.gl == 0) bytes += 4ull * (unsigned long long)odeg(L.uid[i]);
                    clique_forward(k, i);
                    if (abort) return;
                    Lvl &N = ws.lv()[k + 1];
                    N.cur = 0;
                    N.end = N.nu;
                }
                k++;
            } else {
                lds_top = L.mark_lds;
                g_top = L.mark_glb;
                k--;
            }
        }
    }
};

// ---------------------------------------------------------------------------
// work acquisition.  A job head comes back in registers; its prefix ids go to
// the worker's LDS record (jid).
struct JobHead {
    int lvl, lo, hi, q;
    int range_end;  // > 0: a static range job over roots [jid[0], range_end)
    bool donated;
    bool shared;    // one of hi jobs that claim the root's children from its shared cursor
};

// static queues: one per q_group workers (one per worker when q_group <= 1)
DEV int num_queues(const EnumParams &p) {
    const int gq = p.q_group > 1 ? p.q_group : 1;
    return (p.nworkers + gq - 1) / gq;
}
DEV int own_queue(const EnumParams &p, int wid) { return p.q_group > 1 ? wid / p.q_group : wid; }

// Claims the next job of static queue w.
template <int W>
DEV bool take_static(const EnumParams &p, const Grp<W> &g, int w, const WorkerLDS &ws, JobHead &jh) {
    // strided partition (no job arrays, one queue per worker): worker w
    // holds roots n-1-w, n-1-w-nworkers, ...
    const bool strided = p.sq_jobs == nullptr;
    int64_t b = 0, e = 0;
    if (strided) e = w < p.n ? (int64_t)((p.n - 1 - w) / p.nworkers) + 1 : 0;
    else b = p.sq_off[w], e = p.sq_off[w + 1];
    int got = 0, r = 0, lo = 0, hi = 0, q = 0;
    // a shared queue's cursor sits at its first worker's slot, so cursors of
    // groups of 32 or more workers never share a cache line
    int32_t *cursor = &p.sq_cur[p.q_group > 1 ? w * p.q_group : w];
    if (g.gl == 0 && b < e) {
        int cur = __hip_atomic_load(cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (b + cur < e) {
            int j = atomicAdd(cursor, 1);
            if (b + j < e) {
                got = 1;
                if (strided) {
                    r = p.n - 1 - (w + j * p.nworkers);
                    lo = -1;
                    hi = r + 1;
                    q = 0;
                } else {
                    const int32_t *t = p.sq_jobs + 4 * (b + j);
                    r = t[0];
                    lo = t[1];
                    hi = t[2];
                    q = t[3];
                }
                // whoever claims a queue's last job retires the queue
                if (b + j == e - 1) {
                    if (p.qmask) atomicAnd(&p.qmask[w >> 6], ~(1ull << (w & 63)));
                    atomicSub(&p.sc->static_left, 1);
                }
            }
        }
    }
    got = g.bcast(got, 0);
    if (!got) return false;
    jh.lvl = 0;
    jh.donated = false;
    jh.shared = false;
    lo = g.bcast(lo, 0);
    hi = g.bcast(hi, 0);
    jh.q = g.bcast(q, 0);
    r = g.bcast(r, 0);
    if (lo == -2) {  // (root, -2, jobs on the root, q): children from the shared cursor
        jh.range_end = 0;
        jh.shared = true;
        jh.lo = 0;
        jh.hi = hi;
    } else if (lo < 0) {  // range job
        jh.range_end = hi;
        jh.lo = 0;
        jh.hi = 0x7fffffff;
    } else {
        jh.range_end = 0;
        jh.lo = lo;
        jh.hi = hi;
    }
    ws.jid()[0] = r;
    wsync();
    return true;
}

// Donated jobs are served in ticket order: an idle worker takes one ticket
// (atomicAdd on the high half of q_ht) and from then on polls only the slot of its ticket,
// so waiting workers never contend on a shared word.  A slot is complete when
// all its words carry this launch's epoch.  Returns true when the ticket's job
// has arrived (and clears the ticket).
template <int W>
DEV bool poll_ticket(const EnumParams &p, const Grp<W> &g, const WorkerLDS &ws, JobHead &jh, bool &has_ticket,
                     uint32_t &ticket, uint32_t ep) {
    if (!has_ticket) {
        uint32_t t = 0;
        if (g.gl == 0) t = (uint32_t)(atomicAdd(&p.sc->q_ht, 1ull << 32) >> 32);
        ticket = (uint32_t)g.bcast((int)t, 0);
        has_ticket = true;
    }
    if (ticket >= p.q_cap) return false;  // no slot: wait for the end
    const uint64_t *src = p.q[ticket].w;
    int v[(kJobWords + W - 1) / W];
    bool mine_ok = true;
#pragma unroll
    for (int r = 0; r < (kJobWords + W - 1) / W; r++) {
        const int j = g.gl + r * W;
        v[r] = 0;
        if (j < kJobWords) {
            uint64_t x = __hip_atomic_load(src + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((uint32_t)(x >> 32) != ep) mine_ok = false;
            v[r] = (int32_t)(uint32_t)x;
        }
    }
    if (g.ballot(!mine_ok) != 0) return false;
#pragma unroll
    for (int r = 0; r < (kJobWords + W - 1) / W; r++) {
        const int j = g.gl + r * W;
        if (j >= 4 && j < kJobWords) ws.jid()[j - 4] = v[r];
    }
    jh.lvl = g.bcast(v[0], 0);
    jh.lo = g.bcast(v[0], 1);
    jh.hi = g.bcast(v[0], 2);
    jh.q = g.bcast(v[0], 3);
    jh.range_end = 0;
    jh.donated = true;
    jh.shared = false;
    if (jh.lo < 0) {  // the rest of a range job: roots [jid[0], hi)
        jh.range_end = jh.hi;
        jh.lo = 0;
        jh.hi = 0x7fffffff;
    }
    has_ticket = false;
    wsync();
    return true;
}

template <int W>
DEV bool steal_static(const EnumParams &p, const Grp<W> &g, int w, const WorkerLDS &ws, JobHead &jh,
                      uint32_t &seed, bool &static_done) {
    int left = 0;
    if (g.gl == 0) left = __hip_atomic_load(&p.sc->static_left, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (g.bcast(left, 0) <= 0) {
        static_done = true;
        if (g.gl == 0) ws.st()[kStStatic] = wall_clock64();
        return false;
    }
    const int nq = num_queues(p);
    if (nq <= 1) return false;  // the one queue was tried by take_static
    if (p.qmask) {
        // a random queue among those still holding jobs: the lanes read the
        // mask words from a random word on, the first non-zero word (in that
        // order) gives a random set bit.  Plain reads of rarely written
        // words; thieves spread over the non-empty queues.
        const int nwords = (nq + 63) >> 6;
        for (int attempt = 0; attempt < 4; attempt++) {
            seed = seed * 1664525u + 1013904223u;
            const int start = (int)((seed >> 8) % (uint32_t)nwords);
            const int rot = (int)((seed >> 20) & 63);
            int v = -1;
            for (int base = 0; base < nwords && v < 0; base += W) {
                const int i = base + g.gl;
                int wi = start + i;
                if (wi >= nwords) wi -= nwords;
                unsigned long long m = 0;
                if (i < nwords) m = __hip_atomic_load(&p.qmask[wi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t bal = g.ballot(m != 0);
                if (bal) {
                    const int l = __builtin_ctzll(bal);
                    int q = 0;
                    if (g.gl == l) {
                        const unsigned long long r = rot ? (m >> rot) | (m << (64 - rot)) : m;
                        q = wi * 64 + ((__builtin_ctzll(r) + rot) & 63);
                    }
                    v = g.bcast(q, l);
                }
            }
            if (v < 0) return false;  // every queue retired or being retired
            if (v != w && v < nq && take_static<W>(p, g, v, ws, jh)) return true;
        }
        return false;
    }
    // random probes: idle workers spread over the queues (a scan for the
    // fullest queue sends them all to the same cursor, and device-scope
    // atomics on one address slow every worker's memory traffic)
    for (int probe = 0; probe < 8; probe++) {
        seed = seed * 1664525u + 1013904223u;
        int v = (int)((seed >> 8) % (uint32_t)nq);
        if (v == w) continue;
        if (take_static<W>(p, g, v, ws, jh)) return true;
    }
    return false;
}

// Adds a worker's counts for query `q` to the query's result: the clique count
// (one lane's total) and, in batch launches, the worker's class counters.
template <int W>
DEV void flush_query(const EnumParams &p, const Grp<W> &g, int q, unsigned long long count,
                     unsigned long long *whist, unsigned long long last_ticks) {
    if (q < 0) return;
    const QueryDesc &Q = p.qd[q];
    if (g.gl == 0 && count && !Q.motif) atomicAdd(&Q.out[0], count);
    if (g.gl == 0 && Q.done_ticks && last_ticks) atomicMax(Q.done_ticks, last_ticks);
    if (p.nq > 1 && Q.motif) {
        for (int i = g.gl; i < Q.ncls; i += W) {
            unsigned long long v = whist[i];
            if (v) {
                atomicAdd(&Q.out[i], v);
                whist[i] = 0;
            }
        }
        wsync();
    }
}

// MODE 0: cliques only, 1: motifs only, 2: mixed (batch launches); WAVES:
// the waves per SIMD the registers are budgeted for.
template <int W, int MODE, int WAVES = min_waves<W, MODE>()>
__global__ __launch_bounds__(kThreads, WAVES) void enum_kernel(EnumParams p) {
    extern __shared__ __align__(16) char smem[];
    Grp<W> g;
    const int nthr = (int)blockDim.x;  // 64, 128 or 256
    const int wpb = nthr / W;
    const int wib = uni<W>(threadIdx.x / W);
    const int wid = blockIdx.x * wpb + wib;
    // LDS: single-query class counters of the block | per-worker records |
    // per-worker class counters (batch) | per-worker stack slices
    const bool single = p.nq == 1;
    const int bslots = single && p.qd[0].motif ? (p.qd[0].ncls + 1) & ~1 : 0;
    const int wslots = single ? 0 : (p.hist_slots + 1) & ~1;
    const size_t wsb = worker_lds_bytes(p.kmax);
    unsigned long long *bhist = (unsigned long long *)smem;
    char *base = smem + (size_t)bslots * 8;
    const WorkerLDS ws = worker_lds(base + wib * wsb, p.kmax);
    base += wpb * wsb;
    unsigned long long *whist = (unsigned long long *)base + (size_t)wib * wslots;
    base += (size_t)wpb * wslots * 8;
    char *slice = base + (int64_t)wib * p.lds_per_worker;
    for (int i = threadIdx.x; i < bslots; i += nthr) bhist[i] = 0;
    for (int i = threadIdx.x; i < wpb * wslots; i += nthr) ((unsigned long long *)(smem + (size_t)bslots * 8 + wpb * wsb))[i] = 0;
    if (threadIdx.x == 0) {
        const unsigned long long t = wall_clock64();
        atomicMin(&p.sc->t_start, t);
        if (p.start_ticks) atomicMin(p.start_ticks, t);
    }
    __syncthreads();

    unsigned long long bytes = 0;
    if (wid < p.nworkers) {
        if (g.gl == 0)
            for (int i = 0; i < kStats; i++) ws.st()[i] = 0;
        wsync();
        Worker<W, MODE> wk(p, ws, single ? bhist : whist, slice, wid);
        uint32_t seed = 0x9E3779B9u * (uint32_t)(wid + 1);
        bool static_done = false, has_ticket = false;
        uint32_t ticket = 0;
        int backoff = 1;
        uint32_t spins = 0;
        const uint64_t t0 = wk.t_start;
        while (true) {
            JobHead jh;
            bool got = false;
            if (!static_done) {
                got = take_static<W>(p, g, own_queue(p, wid), ws, jh);
                if (!got) got = steal_static<W>(p, g, own_queue(p, wid), ws, jh, seed, static_done);
            }
            if (!got && static_done) got = poll_ticket<W>(p, g, ws, jh, has_ticket, ticket, wk.epoch);
            if (!got) {
                if (wk.pending) {
                    if (g.gl == 0) atomicSub(&p.sc->outstanding, (unsigned long long)wk.pending);
                    wk.pending = 0;
                }
                // the termination words are shared by every worker: read
                // them on every 4th empty poll only
                if ((++spins & 3) == 0) {
                    unsigned long long out = 0;
                    int err = 0;
                    if (g.gl == 0) {
                        out = __hip_atomic_load(&p.sc->outstanding, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        err = __hip_atomic_load(&p.sc->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    out = g.bcast64(out, 0);
                    err = g.bcast(err, 0);
                    if (out == 0 || err) break;
                }
                if (p.watchdog_ticks && now_ticks() - t0 > p.watchdog_ticks) {
                    if (g.gl == 0) atomicCAS(&p.sc->error, 0, 1);
                    break;
                }
                // sleep only when no lane of this wave has work: s_sleep
                // stalls the whole wave, including busy narrower workers
                if (__ballot(1) == ~0ull) {
                    for (int b = 0; b < backoff; b++) __builtin_amdgcn_s_sleep(8);
                    backoff = backoff < 32 ? backoff * 2 : 32;
                }
                continue;
            }
            backoff = 1;
            const int jq = job_query(jh.q);
            bool ok = jh.q >= 0 && jq < p.nq;
            // the job's own donation threshold travels with its query word
            // to every piece it gives away
            wk.qword = jh.q;
            wk.job_sm = job_steal_min(jh.q) >= 2 ? job_steal_min(jh.q) : p.steal_min;
            if (ok && jq != wk.qid) {
                if (!single) flush_query<W>(p, g, wk.qid, ws.st()[kStQCount], whist, ws.st()[kStLast]);
                if (g.gl == 0) ws.st()[kStQCount] = 0;
                const QueryDesc &Q = p.qd[jq];
                wk.qid = jq;
                wk.K = Q.K;
                wk.qmotif = Q.motif != 0;
                wk.lut = Q.lut;
                ok = (MODE == 2 || (MODE == 1) == wk.qmotif) && Q.K >= 2 && Q.K <= p.kmax;
            }
            const int lvl_max = wk.K == 2 ? 0 : wk.K - 3;
            const int r0 = ws.jid()[0];
            const int r1 = jh.range_end > 0 ? jh.range_end : r0 + 1;
            if (!ok || jh.lvl < 0 || jh.lvl > lvl_max || r0 < 0 || r1 > p.n || r0 >= r1) {
                wk.fail(3);
            } else {
                wk.range_hi = r1;
                for (int root = r0; root < wk.range_hi && !wk.abort; root++) {
                    if (jh.range_end > 0) {
                        // range jobs skip roots without a neighbour above them;
                        // a motif root still counts its own node's bytes
                        if (p.off[root + 1] - p.ostart[root] <= 0) {
                            if (wk.qmotif && g.gl == 0)
                                wk.bytes += 4ull * (unsigned long long)(p.off[root + 1] - p.off[root]);
                            continue;
                        }
                        ws.jid()[0] = root;
                        wsync();
                    }
                    // the job's start time goes through LDS: nothing of the
                    // bookkeeping stays live in registers across the DFS
                    if (g.gl == 0) {
                        if (!ws.st()[kStFirst]) ws.st()[kStFirst] = wall_clock64();
                        ws.st()[kStT0] = __builtin_amdgcn_s_memtime();
                    }
                    wk.run_job(jh.lvl, jh.lo, jh.hi,
                               !jh.shared && (jh.range_end > 0 || (!jh.donated && jh.lvl == 0 && jh.lo == 0)),
                               jh.shared);
                    unsigned long long jc = g.lsum(wk.cnt);
                    if (g.gl == 0) {
                        const unsigned long long dt = __builtin_amdgcn_s_memtime() - ws.st()[kStT0];
                        ws.st()[kStQCount] += jc;
                        ws.st()[kStBusy] += dt;
                        if (dt > ws.st()[kStJobMax]) ws.st()[kStJobMax] = dt;
                        ws.st()[kStLast] = wall_clock64();
                        if (p.root_cost) atomicAdd(&p.root_cost[root], dt);
                        if (p.root_count && jc) atomicAdd(&p.root_count[root], jc);
                    }
                }
            }
            if (g.gl == 0) ws.st()[kStJobs]++;
            wk.pending++;
            if (wk.abort) break;
        }
        wsync();
        if (!single) flush_query<W>(p, g, wk.qid, ws.st()[kStQCount], whist, ws.st()[kStLast]);
        bytes = g.lsum(wk.bytes);
    }
    __syncthreads();
    if (single && bslots) {
        for (int i = threadIdx.x; i < p.qd[0].ncls; i += nthr)
            if (bhist[i]) atomicAdd(&p.qd[0].out[i], bhist[i]);
    }
    // Per-worker statistics (held by lane 0 of each worker) reduced over the
    // block, so that each block, not each worker, adds to the shared words.
    // The reduction reuses the start of the block's dynamic LDS (at least the
    // worker records, > 88 B per wave): a static array would count towards
    // the LDS that decides how many blocks share a CU.  Fields 0-4 and 10 are
    // sums, the others maxima.
    constexpr int NF = 11;
    unsigned long long (*red)[NF] = (unsigned long long (*)[NF])smem;
    const bool holder = wid < p.nworkers && g.gl == 0;
    const unsigned long long *st = ws.st();
    unsigned long long f[NF] = {holder ? bytes : 0ull,          holder ? st[kStJobs] : 0ull,
                                holder ? st[kStBusy] : 0ull,    holder ? st[kStDonated] : 0ull,
                                holder && single ? st[kStQCount] : 0ull,
                                holder ? st[kStBusy] : 0ull,    holder ? st[kStFirst] : 0ull,
                                holder ? st[kStStatic] : 0ull,  holder ? st[kStLast] : 0ull,
                                holder ? st[kStJobMax] : 0ull,
                                holder && st[kStJobs] ? 1ull : 0ull};
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int i = 0; i < NF; i++) {
            const unsigned long long y = __shfl_xor(f[i], o, 64);
            f[i] = i < 5 || i == 10 ? f[i] + y : (y > f[i] ? y : f[i]);
        }
    }
    // every wave has read its records (and the class counters were flushed)
    // before the reduction overwrites them
    __syncthreads();
    if (__lane_id() == 0) {
#pragma unroll
        for (int i = 0; i < NF; i++) red[threadIdx.x / 64][i] = f[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nthr / 64; w++) {
            for (int i = 0; i < NF; i++)
                f[i] = i < 5 || i == 10 ? f[i] + red[w][i] : (red[w][i] > f[i] ? red[w][i] : f[i]);
        }
        if (f[0]) atomicAdd(&p.sc->alg_bytes, f[0]);
        if (f[1]) atomicAdd(&p.sc->jobs, f[1]);
        if (f[2]) atomicAdd(&p.sc->busy_sum, f[2]);
        if (f[3]) atomicAdd(&p.sc->donations, f[3]);
        if (f[4] && single && !p.qd[0].motif) atomicAdd(&p.qd[0].out[0], f[4]);
        atomicMax(&p.sc->busy_max, f[5]);
        atomicMax(&p.sc->t_first_job, f[6]);
        atomicMax(&p.sc->t_static_done, f[7]);
        atomicMax(&p.sc->t_last_job, f[8]);
        atomicMax(&p.sc->job_max, f[9]);
        if (f[10]) atomicAdd(&p.sc->active_workers, f[10]);
        atomicMax(&p.sc->t_exit, (unsigned long long)wall_clock64());
    }
}

// Local bitmaps of the launch's prebuilt roots, one block per root, in the
// layout clique_bitmap_root reads: rows (d x NW uint64, row i bit j set when
// L[j] is in N+(L[i]), L = N+(v)) then the out-degrees of L (d int32).  The
// host lists only roots with out-degree <= 256.
// threads per block of the bitmap pre-pass: the largest bitmaps (256 rows of
// up to a few hundred entries) finish sooner with more waves on their rows
// (R-MAT s17, 1,830 roots: 256 threads 94 us, 512 81 us, 1024 89 us)
constexpr int kPrebuildThreads = 512;
__global__ __launch_bounds__(kPrebuildThreads) void prebuild_kernel(EnumParams p) {
    __shared__ int32_t hkey[512];
    __shared__ uint8_t hval[512];
    __shared__ int64_t rst[256];
    __shared__ int32_t spre[257];
    __shared__ int32_t wsum[kPrebuildThreads / 64];
    __shared__ unsigned long long srows[256 * 4];
    // the roots are ascending by degree rank: the largest bitmaps (last) are
    // built by the first blocks, so they do not start last and set the tail
    const int b = p.npre - 1 - (int)blockIdx.x, tid = threadIdx.x;
    if (b < 0) return;
    const int32_t v = p.pre_roots[b];
    const int d = (int)(p.off[v + 1] - p.ostart[v]);
    if (d > 256) return;
    const int NW = d <= 64 ? 1 : d <= 128 ? 2 : 4;
    const int HB = NW == 1 ? 7 : NW == 2 ? 8 : 9;
    const int HS = 1 << HB;
    const int32_t *L = p.adj + p.ostart[v];
    uint64_t *rows = (uint64_t *)const_cast<char *>(p.pre + p.pre_offs[b]);
    int32_t *gdp = (int32_t *)(rows + (int64_t)d * NW);
    for (int t = tid; t < HS; t += blockDim.x) hkey[t] = -1;
    // one child per thread (d <= 256 <= blockDim.x): its list start, its
    // out-degree and the block-wide exclusive prefix of the out-degrees
    int od = 0;
    if (tid < d) {
        const int32_t u = L[tid];
        const int64_t os = p.ostart[u];
        rst[tid] = os;
        od = (int)(p.off[u + 1] - os);
        gdp[tid] = od;
    }
    int incl = od;  // inclusive scan within the wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if ((tid & 63) >= o) incl += y;
    }
    if ((tid & 63) == 63) wsum[tid >> 6] = incl;
    for (int t = tid; t < d * NW; t += blockDim.x) srows[t] = 0;
    __syncthreads();
    int wbase = 0;
    for (int w = 0; w < (tid >> 6); w++) wbase += wsum[w];
    if (tid < d) spre[tid + 1] = wbase + incl;
    if (tid == 0) spre[0] = 0;
    for (int j = tid; j < d; j += blockDim.x) {
        const int32_t key = L[j];
        uint32_t h = ((uint32_t)key * 2654435761u) >> (32 - HB);
        while (atomicCAS(&hkey[h], -1, key) != -1) h = (h + 1) & (HS - 1);
        hval[h] = (uint8_t)j;
    }
    __syncthreads();
    // each wave takes whole rows (child a's out-list, every 4th row) and its
    // lanes stride the list, U loads in flight per lane: coalesced reads and
    // no search for the row of an element; bits go into the LDS rows
    const int wv = tid >> 6, ln = tid & 63;
    constexpr int U = 4;
    for (int a = wv; a < d; a += kPrebuildThreads / 64) {
        const int64_t base = rst[a];
        const int len = spre[a + 1] - spre[a];
        for (int e0 = 0; e0 < len; e0 += U * 64) {
            int32_t xs[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = e0 + u * 64 + ln;
                xs[u] = e < len ? p.adj[base + e] : -1;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int32_t x = xs[u];
                if (x < 0) continue;
                uint32_t h = ((uint32_t)x * 2654435761u) >> (32 - HB);
                int k;
                while ((k = hkey[h]) != x && k != -1) h = (h + 1) & (HS - 1);
                if (k == x) {
                    const int j = hval[h];
                    atomicOr(&srows[a * NW + (j >> 6)], 1ull << (j & 63));
                }
            }
        }
    }
    __syncthreads();
    for (int t = tid; t < d * NW; t += blockDim.x) rows[t] = srows[t];
}

int prebuild_launch(const EnumParams &p, void *stream) {
    if (p.npre <= 0) return DM_OK;
    hipLaunchKernelGGL(prebuild_kernel, dim3(p.npre), dim3(kPrebuildThreads), 0, (hipStream_t)stream, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(DM_ERR_HIP, "prebuild launch: %s", hipGetErrorString(e));
    return DM_OK;
}

__global__ void fill_u32_kernel(uint32_t *ptr, uint32_t value, size_t count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) ptr[i] = value;
}

// Resets the scheduler state of a launch; graph-capturable (no memset or
// host copies).  Bumps the epoch so queue slots of earlier launches never
// match.
__global__ void enum_init_kernel(SchedCounters *sc, unsigned long long njobs, int32_t queues, int32_t *epoch, int32_t *sq_cur,
                                 int nworkers, unsigned long long *results, size_t nresults,
                                 unsigned long long *done_ticks, int nq, unsigned long long *start_ticks,
                                 unsigned long long *qmask, int nqueues, const int64_t *sq_off, int strided_n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    if (i == 0) {
        sc->outstanding = njobs;
        sc->static_left = queues;
        sc->q_ht = 0;
        sc->error = 0;
        sc->donations = 0;
        sc->jobs = 0;
        sc->alg_bytes = 0;
        sc->count = 0;
        sc->busy_sum = 0;
        sc->busy_max = 0;
        sc->t_start = ~0ull;
        sc->t_first_job = 0;
        sc->t_static_done = 0;
        sc->t_last_job = 0;
        sc->t_exit = 0;
        sc->job_max = 0;
        sc->active_workers = 0;
        *epoch = *epoch + 1;
        if (start_ticks) *start_ticks = ~0ull;
    }
    for (size_t t = i; t < (size_t)nworkers; t += stride) sq_cur[t] = 0;
    for (size_t t = i; t < nresults; t += stride) results[t] = 0;
    if (done_ticks)
        for (size_t t = i; t < (size_t)nq; t += stride) done_ticks[t] = 0;
    // non-empty static queues (strided partition: queue w holds roots iff w
    // < n), one queue per thread and one mask word per wave: a wave's lanes
    // hold 64 consecutive queues (blocks of 256 threads, a stride of whole
    // blocks), so the loop condition is uniform over the wave
    if (qmask) {
        const size_t nbits = ((size_t)nqueues + 63) / 64 * 64;
        for (size_t q = i; (q & ~(size_t)63) < nbits; q += stride) {
            bool ne = false;
            if (q < (size_t)nqueues) ne = sq_off ? sq_off[q + 1] > sq_off[q] : (int64_t)q < strided_n;
            const unsigned long long m = __ballot(ne);
            if ((q & 63) == 0) qmask[q >> 6] = m;
        }
    }
}

// ---------------------------------------------------------------------------
size_t enum_lds_bytes(int vwave, int threads, int kmax, int ncls_single, int hist_slots, int lds_per_worker) {
    int wpb = threads / vwave;
    size_t b = (size_t)((ncls_single + 1) & ~1) * 8;
    b += (size_t)wpb * worker_lds_bytes(kmax);
    b += (size_t)wpb * (size_t)((hist_slots + 1) & ~1) * 8;
    b += (size_t)wpb * lds_per_worker;
    return b;
}

// launch modes: 0 cliques, 1 motifs, 2 mixed, 3 motifs of size >= 5 (the
// full-wave motif kernel at kDeepMotifWaves; narrower widths keep mode 1's)
template <int W>
static const void *kernel_mode(int mode) {
    switch (mode) {
        case 0: return (const void *)enum_kernel<W, 0>;
        case 1: return (const void *)enum_kernel<W, 1>;
        case 2: return (const void *)enum_kernel<W, 2>;
        case 3:
            if constexpr (W == 64) return (const void *)enum_kernel<W, 1, kDeepMotifWaves>;
            else return (const void *)enum_kernel<W, 1>;
        default: return nullptr;
    }
}

static const void *kernel_for(int vwave, int mode) {
    switch (vwave) {
        case 8: return kernel_mode<8>(mode);
        case 16: return kernel_mode<16>(mode);
        case 32: return kernel_mode<32>(mode);
        case 64: return kernel_mode<64>(mode);
        default: return nullptr;
    }
}

static bool threads_ok(int threads) { return threads == 64 || threads == 128 || threads == 256; }

int enum_occupancy(int vwave, int mode, int threads, size_t lds, int *blocks_per_cu) {
    const void *k = kernel_for(vwave, mode);
    if (!k || !threads_ok(threads) || threads < vwave) return set_error(DM_ERR_ARG, "unsupported launch shape");
    if (lds > 160 * 1024) return set_error(DM_ERR_ARG, "LDS request %zu exceeds 160 KiB", lds);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return set_error(DM_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
    }
    int nb = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, threads, lds);
    if (e != hipSuccess) return set_error(DM_ERR_HIP, "occupancy query: %s", hipGetErrorString(e));
    // The query divides the CU's 160 KiB by the dynamic size alone.  gfx950
    // allocates a block's LDS, static and dynamic together, in 1280-byte
    // units (measured: 53,760 B per block keeps 3 blocks on a CU, 53,761 B
    // leaves 2), so a persistent launch sized by the query alone would hold
    // blocks that only become resident once others have exited.
    constexpr size_t kLdsGranule = 1280;
    static std::atomic<int> static_lds[4][4] = {};  // + 1: 0 = not read yet
    const int vi = vwave == 8 ? 0 : vwave == 16 ? 1 : vwave == 32 ? 2 : 3;
    int sl = static_lds[vi][mode].load(std::memory_order_relaxed) - 1;
    if (sl < 0) {
        hipFuncAttributes fa{};
        e = hipFuncGetAttributes(&fa, k);
        if (e != hipSuccess) return set_error(DM_ERR_HIP, "hipFuncGetAttributes: %s", hipGetErrorString(e));
        sl = (int)fa.sharedSizeBytes;
        static_lds[vi][mode].store(sl + 1, std::memory_order_relaxed);
    }
    const size_t per_block = (lds + (size_t)sl + kLdsGranule - 1) / kLdsGranule * kLdsGranule;
    const int fit = (int)((size_t)160 * 1024 / per_block);
    *blocks_per_cu = nb < fit ? nb : fit;
    return DM_OK;
}

int enum_launch(const EnumParams &p, int mode, int vwave, int blocks, int threads, size_t lds, void *stream) {
    const void *k = kernel_for(vwave, mode);
    if (!k || !threads_ok(threads) || threads < vwave) return set_error(DM_ERR_ARG, "unsupported launch shape");
    EnumParams copy = p;
    void *args[] = {&copy};
    hipError_t e = hipLaunchKernel(k, dim3(blocks), dim3(threads), args, lds, (hipStream_t)stream);
    if (e != hipSuccess) return set_error(DM_ERR_HIP, "enum launch: %s", hipGetErrorString(e));
    return DM_OK;
}

int enum_init_launch(const EnumParams &p, unsigned long long njobs, int queues, int32_t *epoch,
                     unsigned long long *results, size_t nresults, unsigned long long *done_ticks, int nq,
                     void *stream) {
    size_t work = std::max<size_t>({(size_t)p.nworkers, nresults, (size_t)nq, (size_t)1});
    int blocks = (int)std::min<size_t>((work + 255) / 256, 1024);
    const int g = p.q_group > 1 ? p.q_group : 1;
    const int nqueues = p.sq_jobs ? (p.nworkers + g - 1) / g : p.nworkers;
    hipLaunchKernelGGL(enum_init_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, p.sc, njobs, queues, epoch,
                       p.sq_cur, p.nworkers, results, nresults, done_ticks, nq, p.start_ticks, p.qmask, nqueues,
                       p.sq_jobs ? p.sq_off : (const int64_t *)nullptr, p.n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(DM_ERR_HIP, "init launch: %s", hipGetErrorString(e));
    return DM_OK;
}

int fill_u32_launch(void *ptr, uint32_t value, size_t count, void *stream) {
    if (count == 0) return DM_OK;
    int blocks = (int)((count + 255) / 256);
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       (uint32_t *)ptr, value, count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(DM_ERR_HIP, "fill launch: %s", hipGetErrorString(e));
    return DM_OK;
}

}  // namespace dm
