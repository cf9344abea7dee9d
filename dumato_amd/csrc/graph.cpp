// Graph input, CSR construction and seeded synthetic generators.
//
// Reader semantics follow Graph::Graph / addEdge (src/main/Graph.cpp:17-75):
// whitespace-separated "src dst" integer pairs until end of file; both
// directions inserted; duplicates and reversed pairs collapse; vertex ids
// 0..maxId with isolated ids kept.  Differences, all on malformed input: a
// non-numeric token, an odd number of integers or a negative id is an error
// here (the reference loops or misreads), and self-loops are dropped (the
// reference keeps u in N(u); no clique or motif count depends on it).
// The CSR is DuMatoCPU's (src/main/DuMatoCPU.cu:94-108) without the -1
// sentinel after each list, with int64 offsets.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_set>

#include "internal.h"

namespace dm {

int parse_edgelist(const char *path, std::vector<int32_t> &src, std::vector<int32_t> &dst) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return set_error(DM_ERR_IO, "cannot open %s", path);
    std::vector<char> buf;
    {
        fseek(fp, 0, SEEK_END);
        long sz = ftell(fp);
        fseek(fp, 0, SEEK_SET);
        if (sz < 0) {
            fclose(fp);
            return set_error(DM_ERR_IO, "cannot size %s", path);
        }
        buf.resize((size_t)sz + 1);
        size_t got = fread(buf.data(), 1, (size_t)sz, fp);
        buf[got] = 0;
        buf.resize(got + 1);
    }
    fclose(fp);
    src.clear();
    dst.clear();
    const char *s = buf.data();
    const char *e = s + buf.size() - 1;
    int64_t pending = -1;
    while (s < e) {
        while (s < e && (*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r')) s++;
        if (s >= e) break;
        if (*s == '-') return set_error(DM_ERR_IO, "%s: negative vertex id", path);
        if (*s == '+') s++;
        if (s >= e || *s < '0' || *s > '9') return set_error(DM_ERR_IO, "%s: non-numeric token", path);
        int64_t v = 0;
        while (s < e && *s >= '0' && *s <= '9') {
            v = v * 10 + (*s - '0');
            if (v > 2147483646) return set_error(DM_ERR_IO, "%s: vertex id too large", path);
            s++;
        }
        if (s < e && !(*s == ' ' || *s == '\t' || *s == '\n' || *s == '\r'))
            return set_error(DM_ERR_IO, "%s: non-numeric token", path);
        if (pending < 0) {
            pending = v;
        } else {
            src.push_back((int32_t)pending);
            dst.push_back((int32_t)v);
            pending = -1;
        }
    }
    if (pending >= 0) return set_error(DM_ERR_IO, "%s: odd number of integers", path);
    return DM_OK;
}

int build_csr(int32_t n_hint, const int32_t *src, const int32_t *dst, size_t m, HostCSR &out) {
    int64_t n = n_hint > 0 ? n_hint : 0;
    for (size_t i = 0; i < m; i++) {
        if (src[i] < 0 || dst[i] < 0) return set_error(DM_ERR_ARG, "negative vertex id");
        n = std::max<int64_t>(n, (int64_t)src[i] + 1);
        n = std::max<int64_t>(n, (int64_t)dst[i] + 1);
    }
    if (n > 2147483647) return set_error(DM_ERR_ARG, "too many vertices");
    std::vector<int64_t> deg((size_t)n + 1, 0);
    for (size_t i = 0; i < m; i++) {
        if (src[i] == dst[i]) continue;
        deg[src[i]]++;
        deg[dst[i]]++;
    }
    std::vector<int64_t> off((size_t)n + 1, 0);
    for (int64_t v = 0; v < n; v++) off[v + 1] = off[v] + deg[v];
    std::vector<int32_t> adj((size_t)off[n]);
    std::vector<int64_t> pos(off.begin(), off.end() - 1);
    for (size_t i = 0; i < m; i++) {
        if (src[i] == dst[i]) continue;
        adj[pos[src[i]]++] = dst[i];
        adj[pos[dst[i]]++] = src[i];
    }
    out.n = (int32_t)n;
    out.off.assign((size_t)n + 1, 0);
    int64_t w = 0;
    int32_t maxdeg = 0;
    for (int64_t v = 0; v < n; v++) {
        auto b = adj.begin() + off[v], e = adj.begin() + off[v + 1];
        std::sort(b, e);
        auto ue = std::unique(b, e);
        for (auto it = b; it != ue; ++it) adj[w++] = *it;
        out.off[v + 1] = w;
        maxdeg = std::max<int32_t>(maxdeg, (int32_t)(out.off[v + 1] - out.off[v]));
    }
    adj.resize((size_t)w);
    out.adj.swap(adj);
    out.maxdeg = maxdeg;
    return DM_OK;
}

// ---- generators ---------------------------------------------------------------
static inline uint64_t ekey(int32_t a, int32_t b) {
    if (a > b) std::swap(a, b);
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

// G(n, m): m distinct undirected edges, uniformly at random, no self-loops.
int gen_erdos_renyi(int32_t n, int64_t m, uint64_t seed, std::vector<int32_t> &s,
                    std::vector<int32_t> &d) {
    if (n < 2 || m < 0 || (double)m > 0.5 * (double)n * (double)(n - 1))
        return set_error(DM_ERR_ARG, "G(n,m): need n >= 2 and m <= n(n-1)/2");
    Rng r(seed);
    std::unordered_set<uint64_t> seen;
    seen.reserve((size_t)m * 2);
    s.clear();
    d.clear();
    while ((int64_t)s.size() < m) {
        int32_t a = (int32_t)r.below((uint64_t)n), b = (int32_t)r.below((uint64_t)n);
        if (a == b) continue;
        if (!seen.insert(ekey(a, b)).second) continue;
        s.push_back(std::min(a, b));
        d.push_back(std::max(a, b));
    }
    return DM_OK;
}

// Chung-Lu power law: expected weights w_i = wmax * (1 + i/i0)^(-1/(gamma-1)),
// i0 chosen so that sum(w) = 2 * m_target; edges are drawn with both endpoints
// proportional to w (rejecting self-loops and repeats) until m_target distinct
// edges exist.
int gen_chung_lu(int32_t n, int64_t m_target, double gamma, double wmax, uint64_t seed,
                 std::vector<int32_t> &s, std::vector<int32_t> &d) {
    if (n < 2 || gamma <= 1.0 || wmax <= 0 || m_target < 0 ||
        (double)m_target > 0.25 * (double)n * (double)(n - 1))
        return set_error(DM_ERR_ARG, "Chung-Lu: bad parameters");
    const double ex = -1.0 / (gamma - 1.0);
    auto total = [&](double i0) {
        double t = 0;
        for (int32_t i = 0; i < n; i++) t += wmax * std::pow(1.0 + i / i0, ex);
        return t;
    };
    double target = 2.0 * (double)m_target;
    double lo = 1e-6, hi = 1e12;
    if (total(hi) < target) return set_error(DM_ERR_ARG, "Chung-Lu: wmax too small for m");
    for (int it = 0; it < 200; it++) {
        double mid = std::sqrt(lo * hi);
        if (total(mid) < target) lo = mid;
        else hi = mid;
    }
    const double i0 = hi;
    std::vector<double> cdf((size_t)n);
    double acc = 0;
    for (int32_t i = 0; i < n; i++) {
        acc += wmax * std::pow(1.0 + i / i0, ex);
        cdf[i] = acc;
    }
    Rng r(seed);
    std::unordered_set<uint64_t> seen;
    seen.reserve((size_t)m_target * 2);
    s.clear();
    d.clear();
    auto draw = [&]() {
        double x = r.uniform() * acc;
        return (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), x) - cdf.begin());
    };
    int64_t attempts = 0;
    while ((int64_t)s.size() < m_target) {
        if (++attempts > 200 * (m_target + 10)) return set_error(DM_ERR_ARG, "Chung-Lu: graph too dense");
        int32_t a = std::min(draw(), n - 1), b = std::min(draw(), n - 1);
        if (a == b) continue;
        if (!seen.insert(ekey(a, b)).second) continue;
        s.push_back(a);
        d.push_back(b);
    }
    return DM_OK;
}

// R-MAT: edge_factor * 2^scale draws; each draw picks one quadrant per bit
// with probabilities (a, b, c, 1-a-b-c).  Self-loops are dropped; the CSR
// builder symmetrises and removes repeats.
int gen_rmat(int32_t scale, int32_t ef, double a, double b, double c, uint64_t seed,
             std::vector<int32_t> &s, std::vector<int32_t> &d) {
    if (scale < 1 || scale > 30 || ef < 1 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return set_error(DM_ERR_ARG, "R-MAT: bad parameters");
    Rng r(seed);
    int64_t draws = (int64_t)ef << scale;
    s.clear();
    d.clear();
    s.reserve((size_t)draws);
    d.reserve((size_t)draws);
    for (int64_t e = 0; e < draws; e++) {
        int32_t u = 0, v = 0;
        for (int bit = scale - 1; bit >= 0; bit--) {
            double x = r.uniform();
            if (x < a) {
            } else if (x < a + b) {
                v |= 1 << bit;
            } else if (x < a + b + c) {
                u |= 1 << bit;
            } else {
                u |= 1 << bit;
                v |= 1 << bit;
            }
        }
        if (u == v) continue;
        s.push_back(u);
        d.push_back(v);
    }
    return DM_OK;
}

// Barabasi-Albert preferential attachment (start: clique on m+1 vertices; each
// new vertex links to m distinct existing vertices drawn proportionally to
// degree), then `planted` cliques of size uniform in [kmin, kmax] on random
// vertex sets.
int gen_barabasi_albert(int32_t n, int32_t m, int32_t planted, int32_t kmin, int32_t kmax,
                        uint64_t seed, std::vector<int32_t> &s, std::vector<int32_t> &d) {
    if (m < 1 || n < m + 2 || planted < 0 || (planted > 0 && (kmin < 2 || kmax < kmin || kmax > n)))
        return set_error(DM_ERR_ARG, "Barabasi-Albert: bad parameters");
    Rng r(seed);
    s.clear();
    d.clear();
    std::vector<int32_t> ends;
    ends.reserve((size_t)n * m * 2);
    for (int32_t i = 0; i <= m; i++)
        for (int32_t j = i + 1; j <= m; j++) {
            s.push_back(i);
            d.push_back(j);
            ends.push_back(i);
            ends.push_back(j);
        }
    std::vector<int32_t> chosen;
    for (int32_t v = m + 1; v < n; v++) {
        chosen.clear();
        while ((int32_t)chosen.size() < m) {
            int32_t t = ends[r.below(ends.size())];
            if (std::find(chosen.begin(), chosen.end(), t) == chosen.end()) chosen.push_back(t);
        }
        for (int32_t t : chosen) {
            s.push_back(t);
            d.push_back(v);
            ends.push_back(t);
            ends.push_back(v);
        }
    }
    std::vector<int32_t> members;
    for (int32_t c = 0; c < planted; c++) {
        int32_t sz = kmin + (int32_t)r.below((uint64_t)(kmax - kmin + 1));
        members.clear();
        while ((int32_t)members.size() < sz) {
            int32_t v = (int32_t)r.below((uint64_t)n);
            if (std::find(members.begin(), members.end(), v) == members.end()) members.push_back(v);
        }
        for (int32_t i = 0; i < sz; i++)
            for (int32_t j = i + 1; j < sz; j++) {
                s.push_back(members[i]);
                d.push_back(members[j]);
            }
    }
    return DM_OK;
}

}  // namespace dm
