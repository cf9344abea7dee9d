// Scheduling above the plain runtime: the virtual-wave width choice and the
// learned cost-model partition.  Nothing here changes what is counted: the
// policies only choose launch shapes and how roots are ordered, split into
// child ranges and dealt to workers (reference analogue: the static
// partition of DuMatoCPU, src/main/DuMatoCPU.cu:121-135, and the donation
// weights of EnumerationHelper::getWeight, src/main/EnumerationHelper.cpp:125-132).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>

#include "internal.h"
#include "mlp.h"
#include "runtime.h"

namespace dm {

#define HIPCHK(x)                                                                                    \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) return set_error(DM_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));     \
    } while (0)

// Width used when no tuned value is cached for (mode, k): full 64-lane
// waves, the fastest untuned choice on every benchmark graph measured.
int autotuned_vwave(dm_graph *g, int K, bool motif) {
    int t = g->tuned_vwave[motif ? 1 : 0][K];
    return t ? t : 64;
}

// Untuned donation threshold: 2 for motifs; 32 for cliques, whose bitmap
// roots rebuild part of their bitmap for every donated piece (R-MAT s17
// k = 5: 2.23 ms at 2, 1.43 ms at 32, 1.82 ms at 256).
int tuned_steal_min(const dm_graph *g, int K, bool motif) {
    int t = g->tuned_steal_min[motif ? 1 : 0][K];
    return t ? t : motif ? 2 : 32;
}

int tuned_lds(const dm_graph *g, int K, bool motif, int vwave, int fallback) {
    // a tuned LDS size belongs to the tuned width
    if (g->tuned_vwave[motif ? 1 : 0][K] == vwave && g->tuned_lds[motif ? 1 : 0][K])
        return g->tuned_lds[motif ? 1 : 0][K];
    return fallback;
}

// Autotuner: times candidate launch configurations on a degree-stratified
// sample of the roots (every S-th degree rank) and keeps the fastest per
// (mode, k).  Stage 1 picks the virtual-wave width, stage 2 the LDS stack
// bytes per worker and the steal granularity for that width.
int autotune(dm_graph *g, int K, bool motif, double budget_ms, dm_config *best_out) {
    int r = ensure_device(g);
    if (r) return r;
    const int mi = motif ? 1 : 0;
    std::vector<uint64_t> hist;
    if (motif) {
        const PatternTable *t = pattern_table(K);
        if (!t) return last_code();
        hist.assign(t->ncls, 0);
    }
    auto trial = [&](const dm_config &c, int S, double &ms) -> int {
        uint64_t tot = 0;
        int rr = run_enum(g, K, motif, c, motif ? hist.data() : nullptr, hist.size(), &tot, S);
        if (rr) return rr;
        ms = g->last.device_ms;
        return DM_OK;
    };
    dm_config base;
    dm_config_init(&base);
    base.steal_min = 2;
    base.vwave = 64;
    // roots with work, to size the sample: a trial must give every worker
    // several jobs, or it times the heaviest sampled root instead of the
    // throughput
    int64_t nroots = 0;
    for (int32_t v = 0; v < g->rcsr.n; v++) nroots += root_has_work(g, v, K, motif) ? 1 : 0;
    Shape sh64;
    if ((r = pick_shape(g, K, motif, base, sh64))) return r;
    int S = (int)std::max<int64_t>(1, nroots / (8 * (int64_t)sh64.workers));
    double ms = 0;
    if ((r = trial(base, S, ms))) return r;
    // too slow for the budget: fewer candidates rather than a thinner sample
    const bool full_search = ms * 18 <= budget_ms;
    const bool width_search = ms * 6 <= budget_ms;
    double best = ms;
    dm_config bc = base;
    if (width_search) {
        for (int w : {8, 16, 32}) {
            dm_config c = base;
            c.vwave = w;
            double t = 0;
            if ((r = trial(c, S, t))) return r;
            if (t < best) {
                best = t;
                bc = c;
            }
        }
    }
    const int ncls = motif ? pattern_table(K)->ncls : 0;
    const int dl = max_lds_per_worker(bc.vwave, 256, launch_mode(motif, K), K, ncls, 0);
    if (full_search) {
        const int d8 = 32768 / (256 / bc.vwave);  // 8 KiB for a full wave
        for (int lds : {(dl / 2) & ~255, std::min(d8, dl), dl}) {
            for (int sm : {2, 8, 32, 256}) {
                dm_config c = bc;
                c.lds_bytes = lds;
                c.steal_min = sm;
                double t = 0;
                if ((r = trial(c, S, t))) continue;
                if (t < best) {
                    best = t;
                    bc = c;
                }
            }
        }
    }
    if (bc.lds_bytes == 0) bc.lds_bytes = dl;
    g->tuned_vwave[mi][K] = bc.vwave;
    g->tuned_lds[mi][K] = bc.lds_bytes;
    g->tuned_steal_min[mi][K] = bc.steal_min;
    if (best_out) {
        *best_out = bc;
        best_out->vwave = bc.vwave;
    }
    return DM_OK;
}

void free_model(dm_graph *g) {
    if (g->model) {
        if (g->model->d_params) (void)hipFree(g->model->d_params);
        if (g->model->d_x) (void)hipFree(g->model->d_x);
        if (g->model->d_out) (void)hipFree(g->model->d_out);
        if (g->model->h_out) (void)hipHostFree(g->model->h_out);
    }
    delete g->model;
    g->model = nullptr;
}

// ---------------------------------------------------------------------------
// features (per degree-rank vertex)
static void core_numbers(const HostCSR &c, std::vector<int32_t> &core) {
    // Batagelj-Zaversnik bucket peeling
    const int32_t n = c.n;
    std::vector<int32_t> deg(n), bin(c.maxdeg + 2, 0), pos(n), vert(n);
    for (int32_t v = 0; v < n; v++) {
        deg[v] = (int32_t)(c.off[v + 1] - c.off[v]);
        bin[deg[v]]++;
    }
    int32_t start = 0;
    for (int32_t d = 0; d <= c.maxdeg; d++) {
        int32_t num = bin[d];
        bin[d] = start;
        start += num;
    }
    for (int32_t v = 0; v < n; v++) {
        pos[v] = bin[deg[v]];
        vert[pos[v]] = v;
        bin[deg[v]]++;
    }
    for (int32_t d = c.maxdeg; d >= 1; d--) bin[d] = bin[d - 1];
    bin[0] = 0;
    for (int32_t i = 0; i < n; i++) {
        int32_t v = vert[i];
        for (int64_t e = c.off[v]; e < c.off[v + 1]; e++) {
            int32_t u = c.adj[e];
            if (deg[u] > deg[v]) {
                int32_t du = deg[u], pu = pos[u], pw = bin[du], w = vert[pw];
                if (u != w) {
                    pos[u] = pw;
                    vert[pu] = w;
                    pos[w] = pu;
                    vert[pw] = u;
                }
                bin[du]++;
                deg[u]--;
            }
        }
    }
    core.swap(deg);
}

// 8 features per root (log1p): degree, out-degree, sum of neighbour degrees,
// sum of out-neighbour out-degrees, core number, triangles rooted at v
// (oriented, counted by the library's own k=3 kernel), largest neighbour
// degree, sum of out-neighbour degrees.
static int compute_features(dm_graph *g) {
    if (!g->features.empty()) return DM_OK;
    const HostCSR &c = g->rcsr;
    const int32_t n = c.n;
    // oriented triangles per root from a plain k=3 run
    dm_config cfg;
    dm_config_init(&cfg);
    cfg.per_root = 1;
    cfg.steal = 0;
    uint64_t tri = 0;
    int r = run_enum(g, 3, false, cfg, nullptr, 0, &tri);
    if (r) return r;
    std::vector<uint64_t> tri_orig = g->last_counts;
    std::vector<int32_t> core;
    core_numbers(c, core);
    g->features.assign((size_t)n * kMlpIn, 0.f);
    g->root_density.assign(n, 0.f);
    for (int32_t v = 0; v < n; v++) {
        double deg = (double)(c.off[v + 1] - c.off[v]);
        double up = (double)(c.off[v + 1] - g->ostart[v]);
        g->root_density[v] =
            (float)std::min(1.0, (double)tri_orig[g->inv[v]] / std::max(1.0, up * (up - 1) / 2));
        double s2 = 0, s2o = 0, mx = 0, s2u = 0;
        for (int64_t e = c.off[v]; e < c.off[v + 1]; e++) {
            int32_t u = c.adj[e];
            double du = (double)(c.off[u + 1] - c.off[u]);
            s2 += du;
            mx = std::max(mx, du);
            if (e >= g->ostart[v]) {
                s2o += (double)(c.off[u + 1] - g->ostart[u]);
                s2u += du;
            }
        }
        float *f = &g->features[(size_t)v * kMlpIn];
        f[0] = (float)std::log1p(deg);
        f[1] = (float)std::log1p(up);
        f[2] = (float)std::log1p(s2);
        f[3] = (float)std::log1p(s2o);
        f[4] = (float)std::log1p((double)core[v]);
        f[5] = (float)std::log1p((double)tri_orig[g->inv[v]]);
        f[6] = (float)std::log1p(mx);
        f[7] = (float)std::log1p(s2u);
    }
    return DM_OK;
}

// ---------------------------------------------------------------------------
// device buffers for training / inference
struct MlpDev {
    float *P = nullptr, *M = nullptr, *V = nullptr, *grad = nullptr, *loss = nullptr;
    float *X = nullptr, *Y = nullptr;
    MlpScratch s;
    int cap = 0;
};

static void mlp_dev_free(MlpDev &d) {
    float *ptrs[] = {d.P, d.M, d.V, d.grad, d.loss, d.X, d.Y, d.s.out, d.s.h1, d.s.h2, d.s.dz1, d.s.dz2, d.s.g};
    for (float *p : ptrs)
        if (p) (void)hipFree(p);
    d = MlpDev();
}

static int mlp_dev_alloc(MlpDev &d, int cap) {
    auto A = [](float **p, size_t n) { return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(float)); };
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = A(&d.P, kMlpParams);
    if (e == hipSuccess) e = A(&d.M, kMlpParams);
    if (e == hipSuccess) e = A(&d.V, kMlpParams);
    if (e == hipSuccess) e = A(&d.grad, kMlpParams);
    if (e == hipSuccess) e = A(&d.loss, 1);
    if (e == hipSuccess) e = A(&d.X, (size_t)cap * kMlpIn);
    if (e == hipSuccess) e = A(&d.Y, (size_t)cap);
    if (e == hipSuccess) e = A(&d.s.out, (size_t)cap);
    if (e == hipSuccess) e = A(&d.s.g, (size_t)cap);
    if (e == hipSuccess) e = A(&d.s.h1, (size_t)cap * kMlpHidden);
    if (e == hipSuccess) e = A(&d.s.h2, (size_t)cap * kMlpHidden);
    if (e == hipSuccess) e = A(&d.s.dz1, (size_t)cap * kMlpHidden);
    if (e == hipSuccess) e = A(&d.s.dz2, (size_t)cap * kMlpHidden);
    if (e != hipSuccess) {
        mlp_dev_free(d);
        return set_error(DM_ERR_NOMEM, "MLP buffers: %s", hipGetErrorString(e));
    }
    d.cap = cap;
    return DM_OK;
}

static void init_params(std::vector<float> &P, uint64_t seed) {
    Rng r(seed);
    P.assign(kMlpParams, 0.f);
    auto normal = [&]() {
        double u1 = std::max(r.uniform(), 1e-12), u2 = r.uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    const int oW1 = 0, ob1 = kMlpIn * kMlpHidden, oW2 = ob1 + kMlpHidden, ob2 = oW2 + kMlpHidden * kMlpHidden,
              oW3 = ob2 + kMlpHidden;
    for (int i = 0; i < kMlpIn * kMlpHidden; i++) P[oW1 + i] = (float)(normal() * std::sqrt(2.0 / kMlpIn));
    for (int i = 0; i < kMlpHidden * kMlpHidden; i++) P[oW2 + i] = (float)(normal() * std::sqrt(2.0 / kMlpHidden));
    for (int i = 0; i < kMlpHidden; i++) P[oW3 + i] = (float)(normal() * std::sqrt(1.0 / kMlpHidden));
}

// Train on (X, Y) (host, standardised) with minibatch Adam on the GPU.
static int train_model(const std::vector<float> &X, const std::vector<float> &Y, int epochs, uint64_t seed,
                       std::vector<float> &P, double &final_loss) {
    const int n = (int)Y.size();
    const int batch = std::min(n, 4096);
    MlpDev d;
    int r = mlp_dev_alloc(d, batch);
    if (r) return r;
    init_params(P, seed);
    HIPCHK(hipMemcpy(d.P, P.data(), kMlpParams * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(d.M, 0, kMlpParams * 4));
    HIPCHK(hipMemset(d.V, 0, kMlpParams * 4));
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    Rng rng(seed ^ 0x5bd1e995ull);
    // each epoch's shuffled samples go to the device in one copy; the
    // minibatches are consecutive slices of it
    float *dX = nullptr, *dY = nullptr;
    HIPCHK(hipMalloc((void **)&dX, (size_t)n * kMlpIn * 4));
    if (hipMalloc((void **)&dY, (size_t)n * 4) != hipSuccess) {
        (void)hipFree(dX);
        mlp_dev_free(d);
        return set_error(DM_ERR_NOMEM, "training samples");
    }
    std::vector<float> sx((size_t)n * kMlpIn), sy(n);
    const int nb_epoch = std::max(1, n / batch);
    // small graphs have few minibatches per epoch: train for at least
    // kMinSteps Adam steps
    constexpr int kMinSteps = 1500;
    const int total_steps = std::max(epochs * nb_epoch, kMinSteps);
    int step = 0;
    double last_loss = 0;
    int last_n = 0;
    while (step < total_steps && !r) {
        for (int i = n - 1; i > 0; i--) std::swap(idx[i], idx[rng.below((uint64_t)i + 1)]);
        for (int t = 0; t < n; t++) {
            memcpy(&sx[(size_t)t * kMlpIn], &X[(size_t)idx[t] * kMlpIn], kMlpIn * 4);
            sy[t] = Y[idx[t]];
        }
        if (hipMemcpy(dX, sx.data(), sx.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dY, sy.data(), sy.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
            r = set_error(DM_ERR_HIP, "training sample upload");
            break;
        }
        const bool last_epoch = step + nb_epoch >= total_steps;
        for (int b = 0; b < nb_epoch && step < total_steps; b++) {
            const int s0 = b * batch, m = std::min(batch, n - s0);
            if (last_epoch && hipMemset(d.loss, 0, 4) != hipSuccess) {
                r = set_error(DM_ERR_HIP, "loss reset");
                break;
            }
            if ((r = mlp_grad_launch(d.P, dX + (size_t)s0 * kMlpIn, dY + s0, m, d.s, d.grad, d.loss, nullptr))) break;
            if ((r = mlp_adam_launch(d.P, d.grad, d.M, d.V, 1e-3f, 0.9f, 0.999f, 1e-8f, ++step, nullptr))) break;
            if (last_epoch) {
                float l = 0;
                if (hipMemcpy(&l, d.loss, 4, hipMemcpyDeviceToHost) != hipSuccess) {
                    r = set_error(DM_ERR_HIP, "loss readback");
                    break;
                }
                last_loss += l;
                last_n++;
            }
        }
    }
    (void)hipFree(dX);
    (void)hipFree(dY);
    if (r) {
        mlp_dev_free(d);
        return r;
    }
    const double epoch_loss = last_loss / std::max(last_n, 1);
    HIPCHK(hipMemcpy(P.data(), d.P, kMlpParams * 4, hipMemcpyDeviceToHost));
    mlp_dev_free(d);
    final_loss = epoch_loss;
    return DM_OK;
}

// Predicted log-cycles per degree-rank vertex.  The standardised features
// and the parameters stay on the device after the first call; each call is
// one forward launch and one copy back.
// One forward launch over every vertex; *out (n floats, degree-rank ids) holds
// the standardised outputs until the next call: predicted log cycles are
// out[v] * y_std + y_mean.
static int predict_all(dm_graph *g, const float **out) {
    CostModel &M = *g->model;
    const int32_t n = g->rcsr.n;
    DeviceCtx *ctx = g->dev->ctx;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    if (!M.d_params) {
        std::vector<float> X((size_t)n * kMlpIn);
        for (int32_t v = 0; v < n; v++)
            for (int i = 0; i < kMlpIn; i++)
                X[(size_t)v * kMlpIn + i] = (g->features[(size_t)v * kMlpIn + i] - M.feat_mean[i]) / M.feat_std[i];
        if (hipMalloc(&M.d_params, kMlpParams * 4) != hipSuccess ||
            hipMalloc(&M.d_x, X.size() * 4 + 4) != hipSuccess || hipMalloc(&M.d_out, (size_t)n * 4 + 4) != hipSuccess ||
            hipHostMalloc((void **)&M.h_out, (size_t)n * 4 + 4, hipHostMallocDefault) != hipSuccess)
            return set_error(DM_ERR_NOMEM, "MLP inference buffers");
        HIPCHK(hipMemcpy(M.d_params, M.params.data(), kMlpParams * 4, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(M.d_x, X.data(), X.size() * 4, hipMemcpyHostToDevice));
    }
    int r = mlp_forward_launch(M.d_params, M.d_x, n, M.d_out, ctx->stream);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(M.h_out, M.d_out, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *out = M.h_out;
    return DM_OK;
}

// ---------------------------------------------------------------------------
// Learned partition.  The model's predicted cost ranks the roots (it is
// trained on the costs of a smaller k, so its ranking, not its magnitude, is
// what carries over to the queried k):
//  * the H = max(64, 5%) highest-ranked roots become P jobs each, 64 for
//    the top 32, 32 for the next 96, 16 for the next 384, 8 down to 2% of
//    the roots and 4 for the rest (the ranking of a larger k drifts from the
//    trained one's, so roots heavy at k can sit lower in it), and each job
//    donates from as few as 2 children left.  On the clique bitmap path
//    (k = 4..6, out-degree <= 256) a pre-pass builds the root's local bitmap
//    once and its P jobs claim children from a shared cursor (guided
//    self-scheduling: no weight model decides where the root's work lies;
//    on R-MAT the last children, hubs of the dense core, can hold half of
//    it).  Elsewhere the P jobs are first-level child ranges of equal
//    estimated work, child i of n weighing (n - i)^(k-3);
//  * the lower half of the roots keeps 16 children before it donates (a
//    light job is not worth rebuilding its prefix elsewhere); the rest keep
//    the launch's threshold.
// All jobs go into one list, heaviest predicted first, that the workers claim
// from in order (dynamic longest-first list scheduling).  The orderings are
// counting sorts (exact only over the heavy head), so the host cost stays
// linear in the number of roots.
// Ascending sort of non-negative vertex ids: two radix passes of 11 bits for
// ids below 2^22, std::sort otherwise.
static void sort_ids(std::vector<int32_t> &a) {
    int32_t mx = 0;
    for (int32_t x : a) mx = std::max(mx, x);
    if (mx >= (1 << 22) || a.size() < 64) {
        std::sort(a.begin(), a.end());
        return;
    }
    static thread_local std::vector<int32_t> tmp;
    tmp.resize(a.size());
    int32_t cnt[2048];
    for (int pass = 0; pass < 2; pass++) {
        const int sh = 11 * pass;
        std::vector<int32_t> &src = pass == 0 ? a : tmp, &dst = pass == 0 ? tmp : a;
        std::fill(cnt, cnt + 2048, 0);
        for (int32_t x : src) cnt[(x >> sh) & 2047]++;
        int32_t acc = 0;
        for (int b = 0; b < 2048; b++) {
            const int32_t c = cnt[b];
            cnt[b] = acc;
            acc += c;
        }
        for (int32_t x : src) dst[cnt[(x >> sh) & 2047]++] = x;
    }
}

int learned_partition(dm_graph *g, const std::vector<int32_t> &roots, int K, bool motif, const dm_config &cfg,
                      const Shape &sh, dm_graph::Partition &part) {
    if (!g->model) return set_error(DM_ERR_ARG, "learned scheduler: train a model first (dm_sched_train)");
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    int r = compute_features(g);
    if (r) return r;
    const float *z = nullptr;  // standardised predictions
    if ((r = predict_all(g, &z))) return r;
    const float ys = g->model->y_std, ym = g->model->y_mean;
    const auto t1 = clk::now();
    g->last_infer_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    struct BinTimer {
        dm_graph *g;
        clk::time_point t;
        ~BinTimer() { g->last_bin_ms = std::chrono::duration<double, std::milli>(clk::now() - t).count(); }
    } bin_timer{g, t1};

    struct Job {
        int32_t root, lo, hi, q;
        float lcost;  // predicted log cycles
    };
    const size_t nr = roots.size();
    const size_t W = (size_t)std::max(sh.workers, 1);
    const size_t H8 = std::min(nr, std::max<size_t>(64, nr / 50));
    const size_t H = std::min(nr, std::max<size_t>(64, nr / 20));
    // roots in descending predicted cost: a counting sort on 4096 buckets of
    // the predicted log cost (ascending root within a bucket), made exact
    // over the buckets that hold the H heaviest
    // per-root and per-job scratch keeps its capacity between calls (a fresh
    // allocation's first-touch page faults cost as much as the binning)
    static thread_local std::vector<float> kv;
    static thread_local std::vector<uint16_t> bk;
    static thread_local std::vector<int32_t> idx, hord;
    static thread_local std::vector<uint16_t> hb;
    kv.resize(nr);
    // A model trained on size k_train predicts that size's cost.  For a larger
    // clique size the prediction is carried up by each root's clique growth:
    // a j-clique of N+(v) extends to about (d - j + 1) / j * rho^(j-1)
    // (j+1)-cliques, d = |N+(v)| and rho = the density of N+(v) (oriented
    // triangles / C(d, 2)), both model features.  Without it a k = 4 model
    // ranks R-MAT's dense-core roots (the heaviest at k = 6) near 2,300th; with
    // it they are in the top 10 (median of the 20 heaviest).
    const int kt = g->model->k_train;
    const bool grow = !motif && K > kt;
    if (grow && (g->growth_from != kt || g->growth_to != K || g->growth.size() != (size_t)g->rcsr.n)) {
        // graph-derived like the features: computed once per (k_train, k)
        g->growth.assign(g->rcsr.n, 0.f);
        for (int32_t v = 0; v < g->rcsr.n; v++) {
            const double d = (double)(g->rcsr.off[v + 1] - g->ostart[v]);
            const double rho = (double)g->root_density[v];
            double gr = 1.0, rp = 1.0;  // rp = rho^(j-1)
            for (int j = 1; j < kt; j++) rp *= rho;
            for (int j = kt; j < K; j++, rp *= rho) gr *= std::max(1e-3, (d - j + 1) / j * rp);
            g->growth[v] = (float)std::log(gr);
        }
        g->growth_from = kt;
        g->growth_to = K;
    }
    // predicted log cycles at the training size (carried up by the growth),
    // then their range with 8 independent minima and maxima: one running
    // pair inside the gather loop made it a serial dependency chain
    const float *gr = grow ? g->growth.data() : nullptr;
    for (size_t i = 0; i < nr; i++) {
        const int32_t v = roots[i];
        const float lc = z[v] * ys + ym;
        kv[i] = gr ? lc + gr[v] : lc;
    }
    float pmin = 1e30f, pmax = -1e30f;
    {
        float mn[8], mx[8];
        for (int j = 0; j < 8; j++) mn[j] = 1e30f, mx[j] = -1e30f;
        size_t i = 0;
        for (; i + 8 <= nr; i += 8)
            for (int j = 0; j < 8; j++) {
                const float x = kv[i + j];
                mn[j] = x < mn[j] ? x : mn[j];
                mx[j] = x > mx[j] ? x : mx[j];
            }
        for (; i < nr; i++) pmin = std::min(pmin, kv[i]), pmax = std::max(pmax, kv[i]);
        for (int j = 0; j < 8; j++) pmin = std::min(pmin, mn[j]), pmax = std::max(pmax, mx[j]);
    }
    constexpr int NB = 4096;
    const float scale = pmax > pmin ? (float)(NB - 1) / (pmax - pmin) : 0.f;
    bk.resize(nr);
    std::vector<int32_t> bstart(NB + 1, 0);
    for (size_t i = 0; i < nr; i++) {
        bk[i] = (uint16_t)std::min(NB - 1, std::max(0, (int)((pmax - kv[i]) * scale)));
        bstart[bk[i] + 1]++;
    }
    for (int b = 0; b < NB; b++) bstart[b + 1] += bstart[b];
    idx.resize(nr);
    {
        std::vector<int32_t> fill(bstart.begin(), bstart.end() - 1);
        for (size_t i = 0; i < nr; i++) idx[fill[bk[i]]++] = (int32_t)i;
    }
    if (H > 0) {
        // buckets are monotone in the cost (equal costs share a bucket), so
        // sorting each bucket of the head on its own orders the head exactly
        const int hb_last = bk[idx[H - 1]];
        for (int b = 0; b <= hb_last; b++)
            if (bstart[b + 1] - bstart[b] > 1)
                std::sort(idx.begin() + bstart[b], idx.begin() + bstart[b + 1], [&](int32_t a, int32_t c) {
                    return kv[a] > kv[c] || (kv[a] == kv[c] && a < c);
                });
    }

    const bool can_split = !(K == 2 && !motif);
    const bool bitmap_path = !motif && cfg.clique_bitmap && K >= 4 && K <= 6;
    const int e = motif ? 1 : std::max(1, K - 3);
    // child i of a root with nch children carries about (nch - i)^e of its
    // work (later children see fewer candidates above them); cw[m] = sum of
    // x^e for x = 1..m, so children [lo, hi) weigh cw[nch - lo] - cw[nch - hi]
    int maxch = 0;
    for (size_t rk = 0; rk < H; rk++) {
        const int32_t v = roots[idx[rk]];
        maxch = std::max(maxch, (int)(g->rcsr.off[v + 1] - g->ostart[v]));
    }
    std::vector<double> cw((size_t)maxch + 1, 0.0);
    for (int m = 1; m <= maxch; m++) {
        double x = 1.0;
        for (int i = 0; i < e; i++) x *= (double)m;
        cw[m] = cw[m - 1] + x;
    }
    // heavy roots and their pieces; a piece's predicted log cost is its
    // root's less log P (the cuts give the P pieces equal estimated work)
    static thread_local std::vector<Job> jobs;
    jobs.clear();
    // log of a root's piece count (at most 64)
    static const auto logp = [] {
        std::array<float, 65> t{};
        for (int i = 1; i <= 64; i++) t[i] = (float)std::log((double)i);
        return t;
    }();
    part.pre_roots.clear();
    for (size_t rk = 0; rk < H; rk++) {
        const int32_t v = roots[idx[rk]];
        const int nch = (int)(g->rcsr.off[v + 1] - g->ostart[v]);  // children of the root
        int P = rk < 32 ? 64 : rk < 128 ? 32 : rk < 512 ? 16 : rk < H8 ? 8 : 4;
        P = can_split ? std::min(P, nch) : 1;
        const int32_t q = make_job_query(0, 2);
        if (P <= 1) {
            jobs.push_back({v, 0, INT_MAX, q, kv[idx[rk]]});
            continue;
        }
        const float lc = kv[idx[rk]] - logp[P];
        if (bitmap_path && nch <= 256) {
            // a prebuilt bitmap root: its jobs claim children from one shared
            // cursor, so no weight model fixes where its work is.  Below k = 6
            // a root holds about a tenth of the work, and P sharers' row
            // copies and claims cost more than they balance (R-MAT k = 5:
            // 1.83 ms with P, 1.46 ms with P / 4)
            const int PS = K >= 6 ? P : std::max(2, P / 4);
            part.pre_roots.push_back(v);
            const float lcs = kv[idx[rk]] - logp[PS];
            for (int pi = 0; pi < PS; pi++) jobs.push_back({v, -2, PS, q, lcs});
            continue;
        }
        const double ws = cw[nch];
        int lo = 0;
        for (int pi = 1; pi <= P && lo < nch; pi++) {
            // the first hi > lo whose prefix [0, hi) reaches pi/P of the work
            int hi = nch;
            if (pi < P) {
                const double goal = ws * pi / P;
                int a = lo + 1, b = nch;
                while (a < b) {
                    const int mid = (a + b) >> 1;
                    if (ws - cw[nch - mid] >= goal) b = mid;
                    else a = mid + 1;
                }
                hi = a;
            }
            jobs.push_back({v, lo, hi >= nch ? INT_MAX : hi, q, lc});
            lo = hi;
        }
    }
    const size_t nheavy = jobs.size();

    // one list, heaviest first: the pieces and heavy roots by predicted cost
    // (a counting sort of its log), then the other roots in rank order, the
    // lower half of them light.  It is dealt round-robin to queues shared by
    // kQueueGroup workers each, which claim in list order: longest-first list
    // scheduling inside a group, with a group's cursor shared by few enough
    // workers that its atomics do not contend (one list for all workers ran
    // 2.6x slower: device-scope atomics on one address slowed every worker's
    // memory traffic).
    constexpr size_t kQueueGroup = 8;
    const size_t J = nheavy + (nr - H);
    hord.resize(nheavy);
    {
        float cmin = 1e30f, cmax = -1e30f;
        for (size_t t = 0; t < nheavy; t++) {
            cmin = std::min(cmin, jobs[t].lcost);
            cmax = std::max(cmax, jobs[t].lcost);
        }
        const float cs = cmax > cmin ? (float)(NB - 1) / (cmax - cmin) : 0.f;
        std::vector<int32_t> cb(NB + 1, 0);
        hb.resize(nheavy);
        for (size_t t = 0; t < nheavy; t++) {
            hb[t] = (uint16_t)std::min(NB - 1, std::max(0, (int)((cmax - jobs[t].lcost) * cs)));
            cb[hb[t] + 1]++;
        }
        for (int b = 0; b < NB; b++) cb[b + 1] += cb[b];
        for (size_t t = 0; t < nheavy; t++) hord[cb[hb[t]]++] = (int32_t)t;
    }
    const size_t G = std::min(W, kQueueGroup);
    const size_t nq = (W + G - 1) / G;
    part.group = (int)G;
    // list position t goes to queue t % nq, in list order
    part.sq_off.assign(nq + 1, 0);
    for (size_t qi = 0; qi < nq; qi++) part.sq_off[qi + 1] = part.sq_off[qi] + (int64_t)((J + nq - 1 - qi) / nq);
    part.sq_jobs.resize(J * 4);
    // list positions in order: queue qi = t % nq, slot = t / nq, kept as
    // counters (no division per job)
    size_t qi = 0, slot = 0;
    auto put = [&](int32_t root, int32_t lo, int32_t hi, int32_t q) {
        int32_t *jp = &part.sq_jobs[4 * (size_t)(part.sq_off[qi] + (int64_t)slot)];
        jp[0] = root;
        jp[1] = lo;
        jp[2] = hi;
        jp[3] = q;
        if (++qi == nq) {
            qi = 0;
            slot++;
        }
    };
    for (size_t t = 0; t < nheavy; t++) {
        const Job &jb = jobs[(size_t)hord[t]];
        put(jb.root, jb.lo, jb.hi, jb.q);
    }
    const bool halves = nr > 2 * H;
    for (size_t t = H; t < nr; t++)
        put(roots[idx[t]], 0, INT_MAX, halves && t >= nr / 2 ? make_job_query(0, 16) : 0);

    // prebuilt bitmaps, ascending by root
    sort_ids(part.pre_roots);
    part.pre_offs.resize(part.pre_roots.size());
    int64_t off = 0;
    for (size_t i = 0; i < part.pre_roots.size(); i++) {
        const int32_t v = part.pre_roots[i];
        part.pre_offs[i] = off;
        off += prebuilt_bytes((int32_t)(g->rcsr.off[v + 1] - g->ostart[v]));
    }
    part.pre_bytes = std::max<int64_t>(off, 16);
    return DM_OK;
}

int sched_train(dm_graph *g, int mode_motif, int k_train, int epochs, uint64_t seed, double *final_loss) {
    if (!g) return set_error(DM_ERR_ARG, "graph is NULL");
    if (epochs < 1) return set_error(DM_ERR_ARG, "epochs must be >= 1");
    int r = compute_features(g);
    if (r) return r;
    dm_config cfg;
    dm_config_init(&cfg);
    cfg.record_costs = 1;
    cfg.steal = 0;
    const bool motif = mode_motif != 0;
    std::vector<uint64_t> hist;
    uint64_t tot = 0;
    // per-root cycles summed over a few runs: a single run's small roots
    // are dominated by timing noise
    constexpr int kRuns = 3;
    std::vector<double> cost(g->rcsr.n, 0.0);
    for (int run = 0; run < kRuns; run++) {
        if (motif) {
            const PatternTable *t = pattern_table(k_train);
            if (!t) return last_code();
            hist.assign(t->ncls, 0);
            r = run_enum(g, k_train, true, cfg, hist.data(), hist.size(), &tot);
        } else {
            r = run_enum(g, k_train, false, cfg, nullptr, 0, &tot);
        }
        if (r) return r;
        for (size_t v = 0; v < cost.size(); v++) cost[v] += (double)g->last_costs[v];
    }
    // samples: roots with work, target log(cycles)
    const int32_t n = g->rcsr.n;
    std::vector<int32_t> roots;
    for (int32_t v = 0; v < n; v++)
        if (cost[g->inv[v]] > 0) roots.push_back(v);
    if (roots.size() < 16) return set_error(DM_ERR_ARG, "too few roots with work to train on");
    CostModel *M = new CostModel();
    M->mode_motif = mode_motif;
    M->k_train = k_train;
    const size_t ns = roots.size();
    std::vector<float> X(ns * kMlpIn), Y(ns);
    for (int i = 0; i < kMlpIn; i++) {
        double s = 0, s2 = 0;
        for (int32_t v : roots) {
            double x = g->features[(size_t)v * kMlpIn + i];
            s += x;
            s2 += x * x;
        }
        double mu = s / ns, var = std::max(s2 / ns - mu * mu, 1e-12);
        M->feat_mean[i] = (float)mu;
        M->feat_std[i] = (float)std::sqrt(var);
    }
    double ys = 0, ys2 = 0;
    for (size_t t = 0; t < ns; t++) {
        double y = std::log(cost[g->inv[roots[t]]] / kRuns);
        Y[t] = (float)y;
        ys += y;
        ys2 += y * y;
        for (int i = 0; i < kMlpIn; i++)
            X[t * kMlpIn + i] = (g->features[(size_t)roots[t] * kMlpIn + i] - M->feat_mean[i]) / M->feat_std[i];
    }
    M->y_mean = (float)(ys / ns);
    M->y_std = (float)std::sqrt(std::max(ys2 / ns - (ys / ns) * (ys / ns), 1e-12));
    for (auto &y : Y) y = (y - M->y_mean) / M->y_std;
    double loss = 0;
    if ((r = train_model(X, Y, epochs, seed, M->params, loss))) {
        delete M;
        return r;
    }
    M->final_loss = loss;
    free_model(g);
    g->model = M;
    g->model_gen++;
    if (final_loss) *final_loss = loss;
    return DM_OK;
}

}  // namespace dm

using namespace dm;

extern "C" int dm_sched_train(dm_graph *g, int mode_motif, int k_train, int epochs, uint64_t seed,
                              double *final_loss) {
    return sched_train(g, mode_motif, k_train, epochs, seed, final_loss);
}

extern "C" int dm_autotune(dm_graph *g, int mode_motif, int k, double budget_ms, dm_config *best) {
    if (!g) return set_error(DM_ERR_ARG, "graph is NULL");
    if (budget_ms <= 0) return set_error(DM_ERR_ARG, "budget_ms must be > 0");
    return autotune(g, k, mode_motif != 0, budget_ms, best);
}

extern "C" int dm_sched_predict(dm_graph *g, float *pred, size_t n) {
    if (!g || !pred) return set_error(DM_ERR_ARG, "NULL argument");
    if (!g->model) return set_error(DM_ERR_ARG, "no trained model");
    if (n < (size_t)g->rcsr.n) return set_error(DM_ERR_ARG, "need %d entries", g->rcsr.n);
    int r = ensure_device(g);
    if (r) return r;
    if ((r = compute_features(g))) return r;
    const float *z = nullptr;
    if ((r = predict_all(g, &z))) return r;
    const CostModel &M = *g->model;
    for (int32_t v = 0; v < g->rcsr.n; v++) pred[g->inv[v]] = z[v] * M.y_std + M.y_mean;
    return DM_OK;
}

// Runs a HIP call unless an earlier one failed; records the first failure.
#define MLP_TRY(call, what)                                                                          \
    do {                                                                                             \
        if (!r) {                                                                                    \
            hipError_t e_ = (call);                                                                  \
            if (e_ != hipSuccess) r = set_error(DM_ERR_HIP, "%s: %s", what, hipGetErrorString(e_));    \
        }                                                                                            \
    } while (0)

// Test entry points for the MLP kernels (no graph involved).
extern "C" int dm_mlp_grad(const float *params, const float *X, const float *Y, int n, float *out, float *grad,
                           float *loss) {
    if (!params || !X || !Y || n < 1 || !grad || !loss) return set_error(DM_ERR_ARG, "NULL argument");
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return set_error(DM_ERR_HIP, "no HIP device available");
    MlpDev d;
    int r = mlp_dev_alloc(d, n);
    if (r) return r;
    MLP_TRY(hipMemcpy(d.P, params, kMlpParams * 4, hipMemcpyHostToDevice), "parameter upload");
    MLP_TRY(hipMemcpy(d.X, X, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice), "sample upload");
    MLP_TRY(hipMemcpy(d.Y, Y, (size_t)n * 4, hipMemcpyHostToDevice), "target upload");
    MLP_TRY(hipMemset(d.loss, 0, 4), "loss reset");
    if (!r) r = mlp_grad_launch(d.P, d.X, d.Y, n, d.s, d.grad, d.loss, nullptr);
    MLP_TRY(hipMemcpy(grad, d.grad, kMlpParams * 4, hipMemcpyDeviceToHost), "gradient readback");
    MLP_TRY(hipMemcpy(loss, d.loss, 4, hipMemcpyDeviceToHost), "loss readback");
    if (out) MLP_TRY(hipMemcpy(out, d.s.out, (size_t)n * 4, hipMemcpyDeviceToHost), "output readback");
    MLP_TRY(hipDeviceSynchronize(), "MLP gradient");
    mlp_dev_free(d);
    return r;
}

extern "C" int dm_mlp_adam(float *params, const float *X, const float *Y, int n, int steps, float lr,
                           float *losses) {
    if (!params || !X || !Y || n < 1 || steps < 1) return set_error(DM_ERR_ARG, "bad argument");
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return set_error(DM_ERR_HIP, "no HIP device available");
    MlpDev d;
    int r = mlp_dev_alloc(d, n);
    if (r) return r;
    MLP_TRY(hipMemcpy(d.P, params, kMlpParams * 4, hipMemcpyHostToDevice), "parameter upload");
    MLP_TRY(hipMemset(d.M, 0, kMlpParams * 4), "moment reset");
    MLP_TRY(hipMemset(d.V, 0, kMlpParams * 4), "moment reset");
    MLP_TRY(hipMemcpy(d.X, X, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice), "sample upload");
    MLP_TRY(hipMemcpy(d.Y, Y, (size_t)n * 4, hipMemcpyHostToDevice), "target upload");
    for (int s = 0; s < steps && !r; s++) {
        MLP_TRY(hipMemset(d.loss, 0, 4), "loss reset");
        if (!r) r = mlp_grad_launch(d.P, d.X, d.Y, n, d.s, d.grad, d.loss, nullptr);
        if (!r) r = mlp_adam_launch(d.P, d.grad, d.M, d.V, lr, 0.9f, 0.999f, 1e-8f, s + 1, nullptr);
        if (losses) MLP_TRY(hipMemcpy(&losses[s], d.loss, 4, hipMemcpyDeviceToHost), "loss readback");
    }
    MLP_TRY(hipMemcpy(params, d.P, kMlpParams * 4, hipMemcpyDeviceToHost), "parameter readback");
    MLP_TRY(hipDeviceSynchronize(), "MLP Adam");
    mlp_dev_free(d);
    return r;
}
