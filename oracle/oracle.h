/*
 * CPU oracle for DuMato's enumeration path (clique and motif counting).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only to check the
 * GPU results against (or, for bench.py, as the timed CPU port of the
 * reference's DFS logic).
 * The product library (dumato_amd/) never links or calls it.
 *
 * Every function restates a reference function; each definition in oracle.c
 * names the reference file:line it follows.  Function-level pinning (what
 * checks each function independently) is listed in the header comment of
 * oracle.c.
 */
#ifndef DM_ORACLE_H
#define DM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- graph input / CSR -------------------------------------------------- */
/* Reads "src dst" integer pairs until EOF.  Returns the number of pairs, or -1
 * on open/parse error.  The src and dst arrays are malloc'ed (free with or_free). */
int64_t or_read_edgelist(const char *path, int32_t **src, int32_t **dst);

/* Symmetric, sorted, de-duplicated CSR over ids 0..n-1 where
 * n = max(n_hint, max id + 1).  keep_loops=0 drops self-loops.
 * Returns 0, or -1 on a negative id.  *off (n+1 entries) and *adj malloc'ed. */
int or_csr_build(int32_t n_hint, const int32_t *src, const int32_t *dst,
                 int64_t m, int keep_loops, int32_t *n_out,
                 int64_t **off, int32_t **adj);

/* Degree rank relabelling: new id of v = rank of (deg(v), v) ascending.
 * perm[v] = new id.  Writes the relabelled CSR (same sizes) and, when
 * oriented != 0, keeps only neighbours with a larger new id. */
int or_relabel(int32_t n, const int64_t *off, const int32_t *adj,
               int oriented, int32_t *perm, int64_t **off2, int32_t **adj2);

/* ---- enumeration ------------------------------------------------------- */
/* k-clique count with the thread-per-root DFS of clique_DM_DFS.
 * faithful=1: adjacency tests are linear scans as in the reference;
 * faithful=0: binary search (same result).  per_root (nullable, n entries)
 * receives the count rooted at each vertex.  nthreads<=0: OpenMP default. */
uint64_t or_clique_count(int32_t n, const int64_t *off, const int32_t *adj,
                         int k, int nthreads, int faithful,
                         uint64_t *per_root);

/* Same, over a caller-given root list (roots[0..nroots)), single pass. */
uint64_t or_clique_count_roots(int32_t n, const int64_t *off,
                               const int32_t *adj, int k, int nthreads,
                               int faithful, const int32_t *roots,
                               int64_t nroots, uint64_t *per_root_out);

/* k-motif histogram with the DFS-wide extend / canonical filter / quick
 * induction of motifs_DM_DFS and DuMatoGPU.  lut[quick] = class id
 * (ncls classes).  hist (ncls entries) is overwritten.  per_root (nullable)
 * receives the number of subgraphs rooted at each vertex.  Returns the total,
 * or UINT64_MAX on an argument error. */
uint64_t or_motif_count(int32_t n, const int64_t *off, const int32_t *adj,
                        int k, const uint16_t *lut, int ncls, int nthreads,
                        int faithful, uint64_t *hist, uint64_t *per_root);

uint64_t or_motif_count_roots(int32_t n, const int64_t *off,
                              const int32_t *adj, int k, const uint16_t *lut,
                              int ncls, int nthreads, int faithful,
                              const int32_t *roots, int64_t nroots,
                              uint64_t *hist, uint64_t *per_root_out);

/* Pattern-aware near-clique query qk (q4_PA .. q13_PA): sum over ascending
 * (k-2)-cliques of C(c, 2), c = number of common neighbours of the clique.
 * 4 <= k <= 16; UINT64_MAX on a bad k. */
uint64_t or_near_clique_count(int32_t n, const int64_t *off,
                              const int32_t *adj, int k, int nthreads,
                              int faithful);

/* ---- brute force over vertex subsets (tiny graphs only) ---------------- */
/* qk by k-subsets: C(k,2) per k-clique, 1 per k-subset missing one edge. */
uint64_t or_brute_near_clique(int32_t n, const int64_t *off,
                              const int32_t *adj, int k);
uint64_t or_brute_clique(int32_t n, const int64_t *off, const int32_t *adj,
                         int k);
uint64_t or_brute_motif(int32_t n, const int64_t *off, const int32_t *adj,
                        int k, const uint16_t *lut, int ncls, uint64_t *hist);

/* ---- algorithmic bytes (roofline numerator) ---------------------------- */
/* Clique: oriented CSR; 4*sum_v d+(v) + 4*sum over DFS nodes P with
 * 1..k-2 vertices of sum_{u in C(P)} d+(u). */
uint64_t or_clique_bytes(int32_t n, const int64_t *off, const int32_t *adj,
                         int k, int nthreads);
/* Motif: symmetric CSR; 4*sum over DFS nodes with 1..k-1 vertices of the
 * members' degrees + 4*(k-1)*#leaves. */
uint64_t or_motif_bytes(int32_t n, const int64_t *off, const int32_t *adj,
                        int k, int nthreads);

int or_num_threads(void);
void or_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
