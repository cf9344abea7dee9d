/*
 * CPU oracle for DuMato's enumeration path.  TEST INFRASTRUCTURE ONLY: it is
 * the checker for the HIP product and the timed CPU port in bench.py, and is
 * never linked into or called by the product library.
 *
 * What each function restates, and what pins it independently of this file
 * (tests under tests/):
 *   or_read_edgelist, or_csr_build
 *       Graph::Graph / addEdge (src/main/Graph.cpp:17-75) and the CSR of
 *       DuMatoCPU::initializeCpuDataStructures (src/main/DuMatoCPU.cu:94-108).
 *       Pinned against the reference's own Graph.cpp compiled into
 *       oracle/_ref (same vertex count, edge count, max degree and neighbour
 *       sets on the shipped datasets) and by the invariants sum(deg) = 2m,
 *       sortedness, symmetry.
 *   or_relabel
 *       No reference counterpart (degree-ordered orientation).  Pinned by
 *       permutation validity and by clique counts being unchanged.
 *   or_clique_count[_roots]
 *       clique kernel of src/clique_DM_DFS/clique_DM_DFS.cu:43-75 with
 *       dm_popJob / dm_generateExtension / dm_findNeighbourhood2 /
 *       dm_forward / dm_backward (src/clique_DM_DFS/DuMato.h:66-178).
 *       Pinned by brute force over k-subsets, C(n,k) on complete graphs,
 *       and the reference datasets' known answers.
 *   or_motif_count[_roots]
 *       motifs kernel of src/motifs_DM_DFS/motifs_DM_DFS.cu:46-150 with
 *       dm_generateUniqueExtensions (DuMato.h:237-258), the lower-than and
 *       canonical filters (:312-341), dm_agregateValidSubgraphs (:383-423)
 *       and dm_forward (:425-446); the canonical rule is stated as in
 *       DuMatoGPU::canonicalFilter2 (src/main/DuMatoGPU.cu:989-1012) and the
 *       union as in DuMatoGPU::extend(begin,end) (:1702-1721).  The class
 *       table is the reference's own dictionary (dictionaries/k.csv), used
 *       as a full 16-bit table instead of the reference DFS build's 32-bit
 *       copy of a 64-bit table (src/motifs_DM_DFS/Manager.cpp:270,301).
 *       Pinned by the brute force below (which never uses the canonical
 *       rule), closed forms on stars, paths, cycles and complete graphs,
 *       sum(hist) = number of connected k-subsets, and known answers.
 *   or_near_clique_count
 *       the pattern-aware queries q4_PA .. q13_PA: extend_q4 .. extend_q13
 *       (src/main/DuMatoGPU.cu:1222-1468).  Pinned by or_brute_near_clique
 *       (a count per k-subset by its missing edges, no DFS) and the closed
 *       form C(n,k)*C(k,2) on complete graphs.
 *   or_brute_clique / or_brute_motif / or_brute_near_clique
 *       Enumerate all k-subsets; pinned against a pure-Python brute force
 *       and closed forms.
 *   or_clique_bytes / or_motif_bytes
 *       The algorithmic-byte formulas of the measurement plan; pinned by
 *       closed forms on complete graphs and stars.
 *
 * The counting loops hand out one root at a time to the OpenMP threads
 * (schedule(dynamic, 1)), as the reference DFS baselines hand one root to
 * each thread that asks (the atomicAdd root counter of dm_popJob,
 * src/clique_DM_DFS/DuMato.h:66-88): with larger chunks one thread can hold
 * several hubs while the others idle.
 */
#include "oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXK 16

void or_free(void *p) { free(p); }

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* ------------------------------------------------------------------------ */
/* Graph input: src/main/Graph.cpp:17-54 reads "%d %d" pairs until EOF and
 * inserts both directions into per-vertex std::set (addEdge :60-75).        */
int64_t or_read_edgelist(const char *path, int32_t **src_out,
                         int32_t **dst_out) {
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    int64_t cap = 1024, m = 0;
    int32_t *s = (int32_t *)malloc(cap * sizeof(int32_t));
    int32_t *d = (int32_t *)malloc(cap * sizeof(int32_t));
    long a, b;
    int r;
    while ((r = fscanf(fp, "%ld %ld", &a, &b)) == 2) {
        if (a < 0 || b < 0 || a > 2147483646L || b > 2147483646L) {
            r = -2;
            break;
        }
        if (m == cap) {
            cap *= 2;
            s = (int32_t *)realloc(s, cap * sizeof(int32_t));
            d = (int32_t *)realloc(d, cap * sizeof(int32_t));
        }
        s[m] = (int32_t)a;
        d[m] = (int32_t)b;
        m++;
    }
    fclose(fp);
    if (r != EOF) { /* odd token count, non-numeric token or bad id */
        free(s);
        free(d);
        return -1;
    }
    *src_out = s;
    *dst_out = d;
    return m;
}

static int cmp_i32(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

/* CSR as DuMatoCPU builds it (src/main/DuMatoCPU.cu:94-108): ascending
 * neighbour lists over ids 0..maxId (isolated ids kept, Graph.cpp:34-38),
 * without the -1 sentinel.  Duplicates and reversed pairs collapse as in the
 * std::set of Graph::addEdge. */
int or_csr_build(int32_t n_hint, const int32_t *src, const int32_t *dst,
                 int64_t m, int keep_loops, int32_t *n_out, int64_t **off_out,
                 int32_t **adj_out) {
    int32_t n = n_hint > 0 ? n_hint : 0;
    for (int64_t i = 0; i < m; i++) {
        if (src[i] < 0 || dst[i] < 0) return -1;
        if (src[i] + 1 > n) n = src[i] + 1;
        if (dst[i] + 1 > n) n = dst[i] + 1;
    }
    int64_t *deg = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t i = 0; i < m; i++) {
        if (src[i] == dst[i]) {
            if (keep_loops) deg[src[i]]++;
            continue;
        }
        deg[src[i]]++;
        deg[dst[i]]++;
    }
    int64_t *off = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    off[0] = 0;
    for (int32_t v = 0; v < n; v++) off[v + 1] = off[v] + deg[v];
    int32_t *tmp = (int32_t *)malloc((size_t)(off[n] > 0 ? off[n] : 1) *
                                     sizeof(int32_t));
    int64_t *pos = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    memcpy(pos, off, ((size_t)n + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < m; i++) {
        if (src[i] == dst[i]) {
            if (keep_loops) tmp[pos[src[i]]++] = src[i];
            continue;
        }
        tmp[pos[src[i]]++] = dst[i];
        tmp[pos[dst[i]]++] = src[i];
    }
    /* sort + unique per vertex */
    int64_t *off2 = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    int64_t w = 0;
    off2[0] = 0;
    for (int32_t v = 0; v < n; v++) {
        int64_t lo = off[v], hi = off[v + 1];
        qsort(tmp + lo, (size_t)(hi - lo), sizeof(int32_t), cmp_i32);
        for (int64_t i = lo; i < hi; i++) {
            if (i > lo && tmp[i] == tmp[i - 1]) continue;
            tmp[w++] = tmp[i];
        }
        off2[v + 1] = w;
    }
    free(deg);
    free(off);
    free(pos);
    *n_out = n;
    *off_out = off2;
    *adj_out = tmp;
    return 0;
}

typedef struct {
    int32_t deg, v;
} degv_t;
static int cmp_degv(const void *x, const void *y) {
    const degv_t *a = (const degv_t *)x, *b = (const degv_t *)y;
    if (a->deg != b->deg) return (a->deg > b->deg) - (a->deg < b->deg);
    return (a->v > b->v) - (a->v < b->v);
}

int or_relabel(int32_t n, const int64_t *off, const int32_t *adj,
               int oriented, int32_t *perm, int64_t **off2_out,
               int32_t **adj2_out) {
    degv_t *dv = (degv_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(degv_t));
    for (int32_t v = 0; v < n; v++) {
        dv[v].deg = (int32_t)(off[v + 1] - off[v]);
        dv[v].v = v;
    }
    qsort(dv, (size_t)n, sizeof(degv_t), cmp_degv);
    int32_t *inv = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    for (int32_t r = 0; r < n; r++) {
        perm[dv[r].v] = r;
        inv[r] = dv[r].v;
    }
    int64_t *off2 = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    int32_t *adj2 = (int32_t *)malloc((size_t)(off[n] > 0 ? off[n] : 1) *
                                      sizeof(int32_t));
    off2[0] = 0;
    for (int32_t r = 0; r < n; r++) {
        int32_t v = inv[r];
        int64_t w = off2[r];
        for (int64_t i = off[v]; i < off[v + 1]; i++) {
            int32_t u = perm[adj[i]];
            if (oriented && u <= r) continue;
            adj2[w++] = u;
        }
        qsort(adj2 + off2[r], (size_t)(w - off2[r]), sizeof(int32_t), cmp_i32);
        off2[r + 1] = w;
    }
    free(dv);
    free(inv);
    *off2_out = off2;
    *adj2_out = adj2;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Adjacency test.  faithful: the linear scan of dm_findNeighbourhood2
 * (src/clique_DM_DFS/DuMato.h:157-164, src/motifs_DM_DFS/DuMato.h:352-359);
 * otherwise binary search over the sorted list (same answer). */
static inline int has_nbr(const int64_t *off, const int32_t *adj, int32_t v,
                          int32_t x, int faithful) {
    int64_t lo = off[v], hi = off[v + 1];
    if (faithful) {
        for (int64_t i = lo; i < hi; i++)
            if (adj[i] == x) return 1;
        return 0;
    }
    while (lo < hi) {
        int64_t mid = lo + ((hi - lo) >> 1);
        int32_t a = adj[mid];
        if (a == x) return 1;
        if (a < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return 0;
}

/* ---- cliques: src/clique_DM_DFS/clique_DM_DFS.cu:43-75 ------------------ */
/* One root's DFS.  id[d] is the last vertex; an extension is a neighbour of
 * id[d] that is larger than id[d] and adjacent to every id[i], i < d
 * (clique_DM_DFS.cu:60-67); depth k-1 counts one subgraph (:49-52). */
static uint64_t clique_root(const int64_t *off, const int32_t *adj, int k,
                            int faithful, int32_t root) {
    int32_t id[OR_MAXK];
    int64_t cur[OR_MAXK];
    uint64_t count = 0;
    int d = 0;
    id[0] = root;
    cur[0] = off[root];
    if (k == 1) return 1;
    while (d >= 0) {
        if (d >= k - 1) { /* dm_validSubgraphs + dm_backward */
            count++;
            d--;
            continue;
        }
        if (cur[d] >= off[id[d] + 1]) { /* all extensions traversed */
            d--;
            continue;
        }
        int32_t ext = adj[cur[d]++]; /* dm_generateExtension */
        if (ext <= id[d]) continue;
        int ok = 1;
        for (int i = 0; i < d && ok; i++)
            ok = has_nbr(off, adj, id[i], ext, faithful);
        if (!ok) continue;
        d++; /* dm_forward */
        id[d] = ext;
        cur[d] = off[ext];
    }
    return count;
}

uint64_t or_clique_count_roots(int32_t n, const int64_t *off,
                               const int32_t *adj, int k, int nthreads,
                               int faithful, const int32_t *roots,
                               int64_t nroots, uint64_t *per_root_out) {
    if (k < 1 || k > OR_MAXK) return UINT64_MAX;
    (void)n;
    set_threads(nthreads);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int64_t i = 0; i < nroots; i++) {
        uint64_t c = clique_root(off, adj, k, faithful, roots[i]);
        if (per_root_out) per_root_out[i] = c;
        total += c;
    }
    return total;
}

uint64_t or_clique_count(int32_t n, const int64_t *off, const int32_t *adj,
                         int k, int nthreads, int faithful,
                         uint64_t *per_root) {
    if (k < 1 || k > OR_MAXK) return UINT64_MAX;
    set_threads(nthreads);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int32_t v = 0; v < n; v++) {
        uint64_t c = clique_root(off, adj, k, faithful, v);
        if (per_root) per_root[v] = c;
        total += c;
    }
    return total;
}

/* ---- near-clique query (q4_PA .. q13_PA) ----------------------------------
 * DuMatoGPU::extend_q4 .. extend_q13 (src/main/DuMatoGPU.cu:1222-1468) with
 * extend_single / copy_extensions (:1028-1057) and intersect /
 * intersect_unique (:353-383): levels 0..k-4 grow a clique in ascending ids
 * (neighbours of the last vertex above it, kept if adjacent to every earlier
 * one); level k-3 is the common neighbourhood of the (k-2)-clique (any id,
 * members excluded); level k-2 keeps the entries of level k-3 above the
 * vertex taken there, and the last level adds their number
 * (aggregate_counter).  So each clique adds C(c, 2) for its c common
 * neighbours. */
static uint64_t near_clique_root(const int64_t *off, const int32_t *adj,
                                 int k, int faithful, int32_t root) {
    int32_t id[OR_MAXK];
    int64_t cur[OR_MAXK];
    uint64_t count = 0;
    const int top = k - 3; /* depth at which id[0..k-3] is the clique */
    int d = 0;
    id[0] = root;
    cur[0] = off[root];
    while (d >= 0) {
        if (d == top) {
            /* common neighbours of id[0..top], then the pairs among them */
            uint64_t c = 0;
            for (int64_t p = off[id[0]]; p < off[id[0] + 1]; p++) {
                int32_t x = adj[p];
                int ok = 1;
                for (int i = 1; i <= top && ok; i++)
                    ok = x != id[i] && has_nbr(off, adj, id[i], x, faithful);
                c += ok;
            }
            count += c * (c - (c > 0)) / 2;
            d--;
            continue;
        }
        if (cur[d] >= off[id[d] + 1]) {
            d--;
            continue;
        }
        int32_t ext = adj[cur[d]++];
        if (ext <= id[d]) continue;
        int ok = 1;
        for (int i = 0; i < d && ok; i++)
            ok = has_nbr(off, adj, id[i], ext, faithful);
        if (!ok) continue;
        d++;
        id[d] = ext;
        cur[d] = off[ext];
    }
    return count;
}

uint64_t or_near_clique_count(int32_t n, const int64_t *off,
                              const int32_t *adj, int k, int nthreads,
                              int faithful) {
    if (k < 4 || k > OR_MAXK) return UINT64_MAX;
    set_threads(nthreads);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int32_t v = 0; v < n; v++)
        total += near_clique_root(off, adj, k, faithful, v);
    return total;
}

/* ---- motifs ------------------------------------------------------------ */
typedef struct {
    int k; /* pattern size */
    int faithful;
    const int64_t *off;
    const int32_t *adj;
    const uint16_t *lut;
    int32_t *ext[OR_MAXK];
    int8_t *src[OR_MAXK];
    int64_t cnt[OR_MAXK];
    int32_t id[OR_MAXK];
    uint64_t ind[OR_MAXK]; /* localSubgraphInduction */
    uint64_t *hist;
    uint64_t bytes; /* algorithmic bytes (or_motif_bytes) */
    /* non-faithful dedup: stamp[u] == stamp_val[level] <=> u already in E */
    int32_t *stamp;
    int32_t stamp_next;
    int32_t nstamp;
} mctx_t;

/* bit offset of the quick bits contributed by the vertex at position s:
 * (s-2)(s+1)/2 (src/main/DuMatoGPU.cu:856, :932) */
static inline int quick_off(int s) { return ((s - 2) * (s + 1)) / 2; }

/* E_k: union of N(id[0..k]) restricted to u > id[0], first occurrence keeps
 * its source index (dm_generateUniqueExtensions + dm_filterExtensionsLowerThan
 * + dm_compactExtensions, src/motifs_DM_DFS/DuMato.h:237-319), then the
 * canonical filter (DuMatoGPU::canonicalFilter2, src/main/DuMatoGPU.cu:989-1012):
 * drop (u, s) if exists i in 1..k with (i > s && u <= id[i]) or
 * (i < s && u == id[i]). */
static void motif_extend(mctx_t *c, int k) {
    int64_t w = 0;
    int32_t *E = c->ext[k];
    int8_t *S = c->src[k];
    int32_t v0 = c->id[0];
    int32_t sv = 0;
    if (!c->faithful) {
        if (c->stamp_next == 2147483646) { /* wrap: reset all stamps */
            memset(c->stamp, 0, (size_t)c->nstamp * sizeof(int32_t));
            c->stamp_next = 0;
        }
        sv = ++c->stamp_next;
    }
    for (int i = 0; i <= k; i++) {
        int32_t x = c->id[i];
        c->bytes += 4u * (uint64_t)(c->off[x + 1] - c->off[x]);
        for (int64_t p = c->off[x]; p < c->off[x + 1]; p++) {
            int32_t u = c->adj[p];
            if (u <= v0) continue;
            int dup = 0;
            if (c->faithful) {
                for (int64_t q = 0; q < w && !dup; q++) dup = (E[q] == u);
            } else {
                dup = (c->stamp[u] == sv);
                c->stamp[u] = sv;
            }
            if (dup) continue;
            E[w] = u;
            S[w] = (int8_t)i;
            w++;
        }
    }
    /* canonical filter + compaction */
    int64_t r = 0;
    for (int64_t q = 0; q < w; q++) {
        int32_t u = E[q];
        int s = S[q];
        int drop = 0;
        for (int i = 1; i <= k && !drop; i++)
            drop = (i > s && u <= c->id[i]) || (i < s && u == c->id[i]);
        if (drop) continue;
        E[r] = u;
        S[r] = (int8_t)s;
        r++;
    }
    c->cnt[k] = r;
}

/* quick bits of vertex u placed at position pos against id[0..pos-1] */
static inline uint64_t quick_bits(mctx_t *c, int pos, int32_t u) {
    uint64_t q = 0;
    int o = quick_off(pos);
    for (int j = 0; j < pos; j++)
        if (has_nbr(c->off, c->adj, c->id[j], u, c->faithful))
            q |= (uint64_t)1 << (o + j);
    return q;
}

/* DFS of one root; returns the number of k-subgraphs rooted at it. */
static uint64_t motif_root(mctx_t *c, int32_t root) {
    int K = c->k;
    int64_t cur[OR_MAXK];
    uint64_t total = 0;
    int d = 0;
    c->id[0] = root;
    c->ind[0] = 0;
    c->ind[1] = 0;
    motif_extend(c, 0);
    cur[0] = c->cnt[0];
    while (d >= 0) {
        if (d == K - 2) {
            /* last level: dm_accumulateValidSubgraphs +
             * dm_agregateValidSubgraphs (DuMato.h:379-423) */
            int64_t ne = c->cnt[d];
            total += (uint64_t)ne;
            c->bytes += 4u * (uint64_t)(K - 1) * (uint64_t)ne;
            if (c->lut) {
                for (int64_t q = 0; q < ne; q++) {
                    uint64_t quick = c->ind[d] + quick_bits(c, d + 1, c->ext[d][q]);
                    c->hist[c->lut[quick]]++;
                }
            }
            d--;
            continue;
        }
        if (cur[d] == 0) {
            d--;
            continue;
        }
        /* dm_forward (DuMato.h:425-446): take the last extension */
        int32_t u = c->ext[d][--cur[d]];
        c->id[d + 1] = u;
        if (d + 1 >= 2)
            c->ind[d + 1] = c->ind[d] + quick_bits(c, d + 1, u);
        else
            c->ind[d + 1] = 0;
        d++;
        motif_extend(c, d);
        cur[d] = c->cnt[d];
    }
    return total;
}

static int motif_ctx_init(mctx_t *c, int32_t n, const int64_t *off,
                          const int32_t *adj, int k, const uint16_t *lut,
                          int ncls, int faithful) {
    memset(c, 0, sizeof(*c));
    c->k = k;
    c->faithful = faithful;
    c->off = off;
    c->adj = adj;
    c->lut = lut;
    int64_t maxdeg = 0;
    for (int32_t v = 0; v < n; v++)
        if (off[v + 1] - off[v] > maxdeg) maxdeg = off[v + 1] - off[v];
    for (int l = 0; l <= k - 2; l++) {
        int64_t cap = (int64_t)(l + 1) * maxdeg + 1;
        c->ext[l] = (int32_t *)malloc((size_t)cap * sizeof(int32_t));
        c->src[l] = (int8_t *)malloc((size_t)cap);
    }
    c->hist = (uint64_t *)calloc((size_t)(ncls > 0 ? ncls : 1), sizeof(uint64_t));
    if (!faithful) {
        c->nstamp = n > 0 ? n : 1;
        c->stamp = (int32_t *)calloc((size_t)c->nstamp, sizeof(int32_t));
    }
    return 0;
}

static void motif_ctx_free(mctx_t *c) {
    for (int l = 0; l <= c->k - 2; l++) {
        free(c->ext[l]);
        free(c->src[l]);
    }
    free(c->hist);
    free(c->stamp);
}

static uint64_t motif_run(int32_t n, const int64_t *off, const int32_t *adj,
                          int k, const uint16_t *lut, int ncls, int nthreads,
                          int faithful, const int32_t *roots, int64_t nroots,
                          uint64_t *hist, uint64_t *per_root,
                          uint64_t *bytes_out) {
    if (k < 3 || k > OR_MAXK - 1) return UINT64_MAX;
    set_threads(nthreads);
    if (hist) memset(hist, 0, (size_t)ncls * sizeof(uint64_t));
    uint64_t total = 0, bytes = 0;
    int64_t cnt = roots ? nroots : (int64_t)n;
#pragma omp parallel reduction(+ : total, bytes)
    {
        mctx_t c;
        motif_ctx_init(&c, n, off, adj, k, lut, ncls, faithful);
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < cnt; i++) {
            int32_t v = roots ? roots[i] : (int32_t)i;
            uint64_t t = motif_root(&c, v);
            if (per_root) per_root[i] = t;
            total += t;
        }
        bytes += c.bytes;
        if (hist && lut) {
#pragma omp critical
            for (int j = 0; j < ncls; j++) hist[j] += c.hist[j];
        }
        motif_ctx_free(&c);
    }
    if (bytes_out) *bytes_out = bytes;
    return total;
}

uint64_t or_motif_count(int32_t n, const int64_t *off, const int32_t *adj,
                        int k, const uint16_t *lut, int ncls, int nthreads,
                        int faithful, uint64_t *hist, uint64_t *per_root) {
    if (!lut || !hist) return UINT64_MAX;
    return motif_run(n, off, adj, k, lut, ncls, nthreads, faithful, NULL, 0,
                     hist, per_root, NULL);
}

uint64_t or_motif_count_roots(int32_t n, const int64_t *off,
                              const int32_t *adj, int k, const uint16_t *lut,
                              int ncls, int nthreads, int faithful,
                              const int32_t *roots, int64_t nroots,
                              uint64_t *hist, uint64_t *per_root_out) {
    if (!lut || !hist) return UINT64_MAX;
    return motif_run(n, off, adj, k, lut, ncls, nthreads, faithful, roots,
                     nroots, hist, per_root_out, NULL);
}

uint64_t or_motif_bytes(int32_t n, const int64_t *off, const int32_t *adj,
                        int k, int nthreads) {
    uint64_t bytes = 0;
    uint64_t r = motif_run(n, off, adj, k, NULL, 1, nthreads, 0, NULL, 0,
                           NULL, NULL, &bytes);
    if (r == UINT64_MAX) return UINT64_MAX;
    return bytes;
}

/* ---- clique algorithmic bytes on the oriented CSR ----------------------- */
static uint64_t clique_bytes_rec(const int64_t *off, const int32_t *adj,
                                 int k, int l, const int32_t *C, int64_t nc,
                                 int32_t *buf) {
    /* node with l vertices and candidate set C (sorted) */
    uint64_t b = 0;
    if (l > k - 2) return 0;
    for (int64_t i = 0; i < nc; i++) {
        int32_t u = C[i];
        b += 4u * (uint64_t)(off[u + 1] - off[u]);
        if (l + 1 <= k - 2) {
            /* C' = N+(u) cap C */
            int64_t w = 0;
            int64_t p = off[u], pe = off[u + 1], q = i + 1;
            while (p < pe && q < nc) {
                if (adj[p] == C[q]) {
                    buf[w++] = adj[p];
                    p++;
                    q++;
                } else if (adj[p] < C[q])
                    p++;
                else
                    q++;
            }
            b += clique_bytes_rec(off, adj, k, l + 1, buf, w, buf + w);
        }
    }
    return b;
}

uint64_t or_clique_bytes(int32_t n, const int64_t *off, const int32_t *adj,
                         int k, int nthreads) {
    if (k < 2 || k > OR_MAXK) return UINT64_MAX;
    int64_t maxdeg = 0;
    for (int32_t v = 0; v < n; v++)
        if (off[v + 1] - off[v] > maxdeg) maxdeg = off[v + 1] - off[v];
    set_threads(nthreads);
    uint64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
        int32_t *buf = (int32_t *)malloc((size_t)(maxdeg + 1) * (size_t)k * sizeof(int32_t));
#pragma omp for schedule(dynamic, 16)
        for (int32_t v = 0; v < n; v++) {
            total += 4u * (uint64_t)(off[v + 1] - off[v]);
            total += clique_bytes_rec(off, adj, k, 1, adj + off[v],
                                      off[v + 1] - off[v], buf);
        }
        free(buf);
    }
    return total;
}

/* ---- brute force over k-subsets (n <= 64) ------------------------------- */
static int brute_setup(int32_t n, const int64_t *off, const int32_t *adj,
                       uint64_t *rows) {
    if (n > 64) return -1;
    for (int32_t v = 0; v < n; v++) {
        rows[v] = 0;
        for (int64_t i = off[v]; i < off[v + 1]; i++)
            if (adj[i] != v) rows[v] |= (uint64_t)1 << adj[i];
    }
    return 0;
}

/* next k-combination of 0..n-1 in lexicographic order */
static int next_comb(int *c, int k, int n) {
    int i = k - 1;
    while (i >= 0 && c[i] == n - k + i) i--;
    if (i < 0) return 0;
    c[i]++;
    for (int j = i + 1; j < k; j++) c[j] = c[j - 1] + 1;
    return 1;
}

uint64_t or_brute_clique(int32_t n, const int64_t *off, const int32_t *adj,
                         int k) {
    uint64_t rows[64];
    if (k < 1 || k > n || brute_setup(n, off, adj, rows)) return k > n ? 0 : UINT64_MAX;
    int c[OR_MAXK];
    for (int i = 0; i < k; i++) c[i] = i;
    uint64_t cnt = 0;
    do {
        int ok = 1;
        for (int i = 0; i < k && ok; i++)
            for (int j = i + 1; j < k && ok; j++)
                ok = (rows[c[i]] >> c[j]) & 1;
        cnt += ok;
    } while (next_comb(c, k, n));
    return cnt;
}

/* Each connected k-subset is classified by ANY vertex order whose first two
 * vertices are adjacent: the dictionary maps every quick code of a graph to
 * its isomorphism class.  No canonical DFS rule is involved. */
uint64_t or_brute_motif(int32_t n, const int64_t *off, const int32_t *adj,
                        int k, const uint16_t *lut, int ncls,
                        uint64_t *hist) {
    uint64_t rows[64];
    memset(hist, 0, (size_t)ncls * sizeof(uint64_t));
    if (k < 3 || k > n) return 0;
    if (brute_setup(n, off, adj, rows)) return UINT64_MAX;
    int c[OR_MAXK];
    for (int i = 0; i < k; i++) c[i] = i;
    uint64_t total = 0;
    do {
        uint64_t set = 0;
        for (int i = 0; i < k; i++) set |= (uint64_t)1 << c[i];
        /* connectivity by BFS inside the subset */
        uint64_t seen = (uint64_t)1 << c[0], frontier = seen;
        while (frontier) {
            uint64_t nxt = 0;
            for (int i = 0; i < k; i++)
                if ((frontier >> c[i]) & 1) nxt |= rows[c[i]] & set;
            nxt &= ~seen;
            seen |= nxt;
            frontier = nxt;
        }
        if (seen != set) continue;
        /* order: c[0], its first neighbour in the subset, then the rest */
        int ord[OR_MAXK], w = 0;
        ord[w++] = c[0];
        int second = -1;
        for (int i = 1; i < k; i++)
            if ((rows[c[0]] >> c[i]) & 1) {
                second = c[i];
                break;
            }
        ord[w++] = second;
        for (int i = 1; i < k; i++)
            if (c[i] != second) ord[w++] = c[i];
        uint64_t quick = 0;
        for (int s = 2; s < k; s++)
            for (int t = 0; t < s; t++)
                if ((rows[ord[s]] >> ord[t]) & 1)
                    quick |= (uint64_t)1 << (quick_off(s) + t);
        hist[lut[quick]]++;
        total++;
    } while (next_comb(c, k, n));
    return total;
}

/* Brute force: each k-subset adds C(k,2) when it is a clique and 1 when
 * exactly one of its vertex pairs is not an edge (the pair is then the
 * only choice of the two common neighbours). */
uint64_t or_brute_near_clique(int32_t n, const int64_t *off,
                              const int32_t *adj, int k) {
    uint64_t rows[64];
    if (k < 4 || k > n) return 0;
    if (brute_setup(n, off, adj, rows)) return UINT64_MAX;
    int c[OR_MAXK];
    for (int i = 0; i < k; i++) c[i] = i;
    uint64_t cnt = 0;
    do {
        int missing = 0;
        for (int i = 0; i < k && missing < 2; i++)
            for (int j = i + 1; j < k && missing < 2; j++)
                missing += !((rows[c[i]] >> c[j]) & 1);
        if (missing == 0) cnt += (uint64_t)k * (k - 1) / 2;
        else if (missing == 1) cnt += 1;
    } while (next_comb(c, k, n));
    return cnt;
}
