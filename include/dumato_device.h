/*
 * dumato_device.h -- the DFS-wide enumeration steps as a header-only HIP API
 * for gfx950 (wave64), for user kernels launched with dm_dfs_run.
 *
 * Replaces the reference's device class DuMatoGPU (src/main/DuMatoGPU.cu,
 * src/DM_WCV/DuMatoGPU.cu:12-491): one DFS per worker of W lanes (a virtual
 * wave: W = 8, 16, 32 or 64 lanes of a wave64), driven by a loop of steps.
 * The method names are the reference's, so a kernel in the shape of
 * src/main/clique_PA.cu:9-32 reads the same:
 *
 *   __global__ void clique(dm_dfs_data d) {
 *       dm_device::DuMatoGPU DM_GPU;          // Worker<32>, the reference's warp
 *       DM_GPU.start_clique(d);
 *       while (DM_GPU.active() && DM_GPU.balanced()) {
 *           if (DM_GPU.getCurrentNumberOfExtensions() == -1) DM_GPU.extend_clique();
 *           if (DM_GPU.getCurrentNumberOfExtensions() != 0) {
 *               if (DM_GPU.last_level()) { DM_GPU.aggregate_counter(); DM_GPU.backward_clique(); }
 *               else DM_GPU.forward();
 *           } else {
 *               DM_GPU.backward_clique();
 *           }
 *       }
 *       DM_GPU.end();
 *   }
 *
 * What differs inside (same results): vertex ids are degree ranks, so the
 * reference's "u > last" symmetry break is the degree orientation and an
 * extension list is a suffix of an ascending neighbour list; adjacency tests
 * are binary searches of the sorted lists instead of linear scans (findOne /
 * findMany); the union extend drops an element already contributed by an
 * earlier member's list by testing those lists instead of scanning what was
 * written; compaction is a 64-bit ballot masked to the worker's lanes and a
 * popcount of the lower lanes; class counts go to a per-block LDS histogram
 * flushed with 64-bit atomics at end(); a donated motif job carries its
 * induction instead of re-deriving it by a forward step.
 *
 * Rules for kernels: every thread of a block calls one start* and end()
 * (both synchronise the block); all lanes of a worker call the same steps;
 * the kernel uses no other dynamic LDS (dm_dfs_run sizes it for the
 * histogram); state lives in device memory, so a kernel stopped by the host
 * (balanced() false) resumes where it stopped when relaunched.
 */
#ifndef DUMATO_DEVICE_H
#define DUMATO_DEVICE_H

#include <hip/hip_runtime.h>

#include "dumato.h"

namespace dm_device {

extern __shared__ __align__(16) unsigned long long dm_dfs_lds[];

template <int W>
class Worker {
    static_assert(W == 8 || W == 16 || W == 32 || W == 64, "a worker is 8, 16, 32 or 64 lanes");

   public:
    // ---- start / end (DuMatoGPU::start, start_clique, start_induce, end) ----
    __device__ void start(const dm_dfs_data &dd) { init(dd, kWC); }
    __device__ void start_clique(const dm_dfs_data &dd) { init(dd, kPA); }
    __device__ void start_induce(const dm_dfs_data &dd) { init(dd, kMotif); }
    // pattern-aware near-clique query (start_q4 .. start_q13 with K = 4..13)
    __device__ void start_q(const dm_dfs_data &dd) { init(dd, kQ); }

    // Saves the level (the resume point) and adds the worker's counts to the
    // result.  Synchronises the block.
    __device__ void end() {
        if (lane_ == 0 && ok_) {
            d_.k[wid_] = k_;
            if (cnt_) atomicAdd(d_.result, cnt_);
        }
        __syncthreads();
        if (d_.motif) {
            for (int i = threadIdx.x; i < d_.ncls; i += blockDim.x)
                if (dm_dfs_lds[i]) atomicAdd(&d_.result[i], dm_dfs_lds[i]);
        }
    }

    // ---- state ----
    __device__ bool active() const { return k_ >= 0; }
    // false once the host has asked the kernel to stop (stopKernel)
    __device__ bool balanced() const {
        return __hip_atomic_load(d_.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    }
    __device__ int k() const { return k_; }
    __device__ bool last_level() const { return k_ == d_.K - 2; }
    // extensions left at the current level; -1 = not generated yet
    __device__ int getCurrentNumberOfExtensions() const { return row(d_.next)[k_]; }
    __device__ int getId(int level) const { return row(d_.id)[level]; }
    __device__ int getExtension(int pos) const { return list(k_)[pos]; }
    __device__ int getSource(int pos) const { return srcs(k_)[pos]; }
    __device__ int degree(int32_t v) const { return (int)(d_.off[v + 1] - d_.off[v]); }
    __device__ const int32_t *getAdjacency(int32_t v) const { return d_.adj + d_.off[v]; }
    __device__ int neighbour(int32_t v, int pos) const { return d_.adj[d_.off[v] + pos]; }
    __device__ int lane() const { return lane_; }

    // ---- compaction (DuMatoGPU::write, write_twice) ----
    // Lanes with valid write value to v[start + (valid lanes below them)];
    // returns the number of valid lanes (the same on every lane).
    __device__ int write(int32_t *v, int start, int32_t value, bool valid) const {
        const unsigned long long m = ballot(valid);
        if (valid) v[start + below(m)] = value;
        return __popcll(m);
    }
    __device__ int write_twice(int32_t *v1, int start, int32_t value1, int8_t *v2, int value2, bool valid) const {
        const unsigned long long m = ballot(valid);
        if (valid) {
            const int o = start + below(m);
            v1[o] = value1;
            v2[o] = (int8_t)value2;
        }
        return __popcll(m);
    }

    // ---- extend ----
    // clique (DM_WCV extend): the current level's list = N+(id[k])
    __device__ void extend() {
        const int32_t v = getId(k_);
        const int32_t *nb = d_.adj + d_.ostart[v];
        const int dn = (int)(d_.off[v + 1] - d_.ostart[v]);
        int32_t *out = list(k_);
        int w = 0;
        for (int base = 0; base < dn; base += W) {
            const int t = base + lane_;
            w += write(out, w, t < dn ? nb[t] : -1, t < dn);
        }
        set_count(k_, w, true);
    }

    // clique (clique_PA extend_clique): C_k = N+(id[k]) & C_{k-1}
    __device__ void extend_clique() { extend_clique_at(k_, true); }

    // motif (extend(begin, end)): union of N(id[begin..end]) above id[begin];
    // an element keeps the first list it appears in as its source
    __device__ void extend(int begin, int end_) {
        const int32_t v0 = getId(begin);
        int32_t *out = list(k_);
        int8_t *so = srcs(k_);
        int w = 0;
        for (int i = begin; i <= end_; i++) {
            const int32_t v = getId(i);
            const int32_t *nb = getAdjacency(v);
            const int dv = degree(v);
            for (int base = 0; base < dv; base += W) {
                const int t = base + lane_;
                const int32_t x = t < dv ? nb[t] : -1;
                bool ok = t < dv && x > v0;
                for (int j = begin; j < i && ok; j++) ok = !adjacent(getId(j), x);
                w += write_twice(out, w, x, so, i, ok);
            }
        }
        set_count(k_, w, true);
    }

    // pattern-aware near-clique query (extend_q4 .. extend_q13,
    // src/main/DuMatoGPU.cu:1222-1468): a K-vertex pattern of a
    // (K-2)-clique and two vertices adjacent to all of it.  Levels 0..K-4
    // build the clique like extend_clique; level K-3 is the common
    // neighbourhood of the clique (any id, members excluded); level K-2 the
    // entries of level K-3 above the vertex taken there, so the last level
    // counts the unordered pairs.
    __device__ void extend_q() { extend_q_at(k_, true); }

    // ---- filter ----
    // clique (DM_WCV filterClique): keep the extensions adjacent to id[0..k-1]
    __device__ void filterClique() {
        const int n = getCurrentNumberOfExtensions();
        int32_t *l = list(k_);
        int w = 0;
        for (int base = 0; base < n; base += W) {
            const int t = base + lane_;
            const int32_t x = t < n ? l[t] : -1;
            bool ok = t < n;
            for (int j = 0; j < k_ && ok; j++) ok = adjacent(getId(j), x);
            w += write(l, w, x, ok);
        }
        set_count(k_, w, true);
    }

    // motif (canonicalFilter2): drop (x, s) when some i in 1..k has
    // (i > s and x <= id[i]) or (i < s and x == id[i])
    __device__ void canonicalFilter2() {
        const int n = getCurrentNumberOfExtensions();
        int32_t *l = list(k_);
        int8_t *so = srcs(k_);
        int w = 0;
        for (int base = 0; base < n; base += W) {
            const int t = base + lane_;
            const int32_t x = t < n ? l[t] : -1;
            const int s = t < n ? (int)so[t] : 0;
            bool ok = t < n;
            for (int i = 1; i <= k_ && ok; i++) {
                const int32_t tg = getId(i);
                if ((i > s && x <= tg) || (i < s && x == tg)) ok = false;
            }
            w += write_twice(l, w, x, so, s, ok);
        }
        set_count(k_, w, true);
    }
    __device__ void canonicalFilter() { canonicalFilter2(); }

    // ---- aggregate ----
    __device__ void aggregate_counter() {
        if (lane_ == 0) cnt_ += (unsigned long long)getCurrentNumberOfExtensions();
    }
    // class of every leaf: induction of the prefix plus the leaf's adjacency
    // to id[0..k]
    __device__ void aggregate_pattern() {
        const int n = getCurrentNumberOfExtensions();
        const int32_t *l = list(k_);
        const int off = quick_off(k_ + 1);
        const unsigned long long base = row(d_.ind)[k_];
        for (int t = lane_; t < n; t += W) {
            const int32_t x = l[t];
            unsigned long long q = base;
            for (int j = 0; j <= k_; j++)
                if (adjacent(getId(j), x)) q |= 1ull << (off + j);
            atomicAdd(&dm_dfs_lds[d_.lut[q]], 1ull);
        }
    }

    // ---- move ----
    // take the last extension of the level as the next level's vertex
    __device__ void forward() { (void)step_down(); }
    __device__ void forward_induce() {
        const int32_t x = step_down();
        unsigned long long q = 0;
        if (k_ >= 2) {
            // quick bits of the new vertex against id[0..k-1], one lane each
            for (int b = 0; b < k_; b += W) {
                const int i = b + lane_;
                const unsigned long long m = ballot(i < k_ && adjacent(getId(i), x));
                q |= m << b;
            }
            q = row(d_.ind)[k_ - 1] + (q << quick_off(k_));
        }
        sync();
        if (lane_ == 0) row(d_.ind)[k_] = q;
        sync();
    }
    __device__ void backward() { up(kWC); }
    __device__ void backward_clique() { up(kPA); }
    __device__ void backward_induce() { up(kMotif); }
    __device__ void backward_q() { up(kQ); }

    // ---- lane primitives of the worker ----
    __device__ unsigned long long ballot(bool p) const {
        const unsigned long long b = __ballot(p);
        if (W == 64) return b;
        return (b >> (__lane_id() & ~(W - 1))) & ((1ull << W) - 1);
    }
    __device__ bool any(bool p) const { return ballot(p) != 0; }
    __device__ int below(unsigned long long m) const { return __popcll(m & ((1ull << lane_) - 1)); }

   private:
    enum { kWC = 0, kPA = 1, kMotif = 2, kQ = 3 };
    dm_dfs_data d_;
    int wid_, lane_, k_;
    bool ok_;  // a worker of the state (and the state's width)
    unsigned long long cnt_;

    __device__ static int quick_off(int s) { return ((s - 2) * (s + 1)) >> 1; }
    __device__ static void sync() { __threadfence_block(); }

    template <class T>
    __device__ T *row(T *a) const {
        return a + (size_t)wid_ * DM_DFS_MAXK;
    }
    __device__ size_t list_off(int l) const {
        return d_.motif ? (size_t)d_.cap_base * (size_t)(l * (l + 1) / 2) : (size_t)d_.cap_base * (size_t)l;
    }
    __device__ int32_t *list(int l) const { return d_.ext + (size_t)wid_ * d_.ext_per_worker + list_off(l); }
    __device__ int8_t *srcs(int l) const { return d_.src + (size_t)wid_ * d_.ext_per_worker + list_off(l); }

    // x in N(u): binary search of the ascending list
    __device__ bool adjacent(int32_t u, int32_t x) const {
        const int32_t *a = d_.adj + d_.off[u];
        int lo = 0, hi = (int)(d_.off[u + 1] - d_.off[u]);
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (a[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        return lo < (int)(d_.off[u + 1] - d_.off[u]) && a[lo] == x;
    }

    __device__ void set_count(int l, int n, bool also_next) {
        sync();
        if (lane_ == 0) {
            row(d_.fixed)[l] = n;
            if (also_next) row(d_.next)[l] = n;
        }
        sync();
    }

    __device__ void extend_clique_at(int l, bool set_next) {
        const int32_t last = getId(l);
        const int32_t *nb = d_.adj + d_.ostart[last];
        const int dn = (int)(d_.off[last + 1] - d_.ostart[last]);
        const int32_t *prev = l > 0 ? list(l - 1) : nullptr;
        const int np = l > 0 ? row(d_.fixed)[l - 1] : 0;
        int32_t *out = list(l);
        int w = 0;
        for (int base = 0; base < dn; base += W) {
            const int t = base + lane_;
            const int32_t x = t < dn ? nb[t] : -1;
            bool ok = t < dn;
            if (ok && l > 0) {
                int lo = 0, hi = np;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (prev[mid] < x) lo = mid + 1;
                    else hi = mid;
                }
                ok = lo < np && prev[lo] == x;
            }
            w += write(out, w, x, ok);
        }
        set_count(l, w, set_next);
    }

    __device__ void extend_q_at(int l, bool set_next) {
        const int K = d_.K;
        if (l <= K - 4) {  // the clique levels
            extend_clique_at(l, set_next);
            return;
        }
        int32_t *out = list(l);
        int w = 0;
        if (l == K - 3) {
            // N(id[0]) & ... & N(id[l]); a member is in no list of its own
            const int32_t v0 = getId(0);
            const int32_t *nb = getAdjacency(v0);
            const int dv = degree(v0);
            for (int base = 0; base < dv; base += W) {
                const int t = base + lane_;
                const int32_t x = t < dv ? nb[t] : -1;
                bool ok = t < dv;
                for (int i = 1; i <= l && ok; i++) ok = adjacent(getId(i), x);
                w += write(out, w, x, ok);
            }
        } else {
            // the previous level's full list above the vertex taken from it
            const int32_t last = getId(l);
            const int32_t *prev = list(l - 1);
            const int np = row(d_.fixed)[l - 1];
            for (int base = 0; base < np; base += W) {
                const int t = base + lane_;
                const int32_t x = t < np ? prev[t] : -1;
                w += write(out, w, x, t < np && x > last);
            }
        }
        set_count(l, w, set_next);
    }

    __device__ int32_t step_down() {
        const int nx = getCurrentNumberOfExtensions();
        const int32_t x = list(k_)[nx - 1];
        sync();
        if (lane_ == 0) {
            row(d_.next)[k_] = nx - 1;
            row(d_.id)[k_ + 1] = x;
            row(d_.next)[k_ + 1] = -1;
        }
        sync();
        k_++;
        return x;
    }

    __device__ void up(int mode) {
        k_--;
        if (k_ < 0) pop_job(mode);
    }

    // Next job of the worker's queue (reference popJob / popJob_clique /
    // popJob_induce): its prefix ids[0..L] at levels 0..L, level L to be
    // generated, lower levels spent; a clique_PA job rebuilds the full lists
    // of its lower levels (prepare_job_clique).
    __device__ void pop_job(int mode) {
        const int J = d_.jobs_per_worker;
        const int cur = d_.job_cur[wid_], valid = d_.job_valid[wid_];
        sync();
        if (cur >= valid) {
            if (lane_ == 0) d_.status[wid_] = 2;
            sync();
            k_ = -1;
            return;
        }
        const size_t j = (size_t)wid_ * J + cur;
        const int L = d_.job_level[j];
        const int32_t *jp = d_.jobs + j * DM_DFS_MAXK;
        for (int i = lane_; i < DM_DFS_MAXK; i += W) {
            row(d_.id)[i] = i <= L ? jp[i] : -1;
            row(d_.next)[i] = i == L ? -1 : 0;
        }
        if (lane_ == 0) {
            d_.job_cur[wid_] = cur + 1;
            d_.status[wid_] = 1;
            row(d_.ind)[L] = mode == kMotif ? d_.job_ind[j] : 0ull;
        }
        sync();
        k_ = L;
        if (mode == kPA)
            for (int l = 0; l < L; l++) extend_clique_at(l, false);
        if (mode == kQ)
            for (int l = 0; l < L; l++) extend_q_at(l, false);
    }

    __device__ void init(const dm_dfs_data &dd, int mode) {
        d_ = dd;
        wid_ = (int)((blockIdx.x * blockDim.x + threadIdx.x) / W);
        lane_ = (int)(threadIdx.x & (W - 1));
        cnt_ = 0;
        if (d_.motif)
            for (int i = threadIdx.x; i < d_.ncls; i += blockDim.x) dm_dfs_lds[i] = 0;
        __syncthreads();
        // lanes beyond the state's workers, or a kernel built for another
        // width, touch nothing
        ok_ = wid_ < d_.workers && d_.vwave == W;
        if (!ok_) {
            k_ = -1;
            return;
        }
        k_ = d_.k[wid_];
        if (lane_ == 0) d_.status[wid_] = 1;
        sync();
        if (k_ < 0) pop_job(mode);
    }
};

// The reference's 32-lane warp.
typedef Worker<32> DuMatoGPU;

}  // namespace dm_device

#endif
