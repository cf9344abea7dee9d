// Drop-in for the reference's pattern-aware query executables q4_PA .. q13_PA
// (src/main/q4_PA.cu:34-130 .. q13_PA.cu; built from one source here): same
// argv, same stdout lines ("Result: N.", "Time (GPU|LB|IO|TOTAL)").  The
// query is the near-clique pattern of size k (argv): a (k-2)-clique and two
// vertices adjacent to all of it, counted per clique and unordered pair of
// common neighbours.  It runs in the reference's model: the q_PA kernel
// written against the step API, threads / 256 blocks, donationsPerWarp job
// slots per worker, and the host-driven stop / snapshot / donate loop with
// the argv threshold (>= 100: no balancing), polling every 100 ms
// (DM_LB_POLL_US overrides).  Environment overrides: DM_VWAVE (8/16/32/64).
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "../../include/dumato.h"

static double secs(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

int main(int argc, const char **argv) {
    printf("Usage: %s graphFile k threads blockSize donationsPerWarp threshold\n", argv[0]);
    printf("\t graphFile: \t\t url of graph dataset\n");
    printf("\t k: \t\t\t clique size\n");
    printf("\t threads: \t\t amount of GPU threads (recommended: 102400)\n");
    printf("\t blockSize: \t\t amount of threads per block (recommended: 256)\n");
    printf("\t donationsPerWarp: \t amount of donations during load-balancing (recommended: 16)\n");
    printf("\t threshold: \t\t load-balancing threshold (recommended: 30)\n");
    if (argc != 7) {
        printf("\nWrong amount of parameters!\n");
        printf("Exiting...\n");
        return 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    const char *path = argv[1];
    const int k = atoi(argv[2]);
    dm_graph *g = nullptr;
    if (dm_graph_load_edgelist(path, &g) != DM_OK) {
        fprintf(stderr, "%s\n", dm_last_error());
        return 1;
    }
    dm_dfs_config dc;
    dm_dfs_config_init(&dc);
    if (const char *v = getenv("DM_VWAVE")) dc.vwave = atoi(v);
    dc.blocks = atoi(argv[3]) / 256 > 0 ? atoi(argv[3]) / 256 : 1;
    dc.jobs_per_worker = atoi(argv[5]) > 0 ? atoi(argv[5]) : 16;
    dc.threshold = atoi(argv[6]);
    dc.poll_us = 100000;
    if (const char *v = getenv("DM_LB_POLL_US")) dc.poll_us = atoi(v);
    auto t1 = std::chrono::steady_clock::now();
    dm_dfs *s = nullptr;
    const void *kern = nullptr;
    dm_dfs_stats ds = {};
    uint64_t count = 0;
    int rc = dm_dfs_kernel(DM_DFS_NEAR_CLIQUE, dc.vwave, &kern);
    if (rc == DM_OK) rc = dm_dfs_create(g, k, DM_PATTERN_NEAR_CLIQUE, &dc, &s);
    if (rc == DM_OK) rc = dm_dfs_run(s, kern, &count, 1, &ds);
    dm_dfs_free(s);
    auto t2 = std::chrono::steady_clock::now();
    if (rc != DM_OK) {
        fprintf(stderr, "%s\n", dm_last_error());
        dm_graph_free(g);
        return 1;
    }
    printf("Result: %lu.\n", (unsigned long)count);
    const double gpu = ds.total_ms / 1e3, lb = ds.lb_ms / 1e3;
    printf("Time (GPU): %f secs.\n", gpu);
    printf("Time (LB): %f secs,%f%%\n", lb, gpu > 0 ? 100.0 * lb / gpu : 0.0);
    printf("Time (IO): %f secs\n", secs(t0, t1));
    printf("Time (TOTAL): %f secs\n", secs(t0, t2));
    dm_graph_free(g);
    return 0;
}
