// Drop-in for the reference's clique_PA executable (src/main/clique_PA.cu:34-130):
// same argv, same stdout lines ("Result: N.", "Time (GPU|LB|IO|TOTAL)").
// By default the library's fused kernel counts and threads / blockSize /
// donationsPerWarp are accepted and ignored (the launch shape is chosen for
// the MI355X); threshold >= 100 disables load balancing as in the reference
// (:78), any smaller value enables the on-device work stealing.
// DM_LB=host runs the reference's programming model instead: the clique_PA
// kernel written against the step API, threads / 256 blocks, donationsPerWarp
// job slots per worker, and the host-driven stop / snapshot / donate loop
// with the argv threshold, polling every 100 ms (DM_LB_POLL_US overrides).
// Environment overrides: DM_VWAVE (8/16/32/64).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

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
    int k = atoi(argv[2]);
    int threshold = atoi(argv[6]);
    dm_graph *g = nullptr;
    if (dm_graph_load_edgelist(path, &g) != DM_OK) {
        fprintf(stderr, "%s\n", dm_last_error());
        return 1;
    }
    dm_config cfg;
    dm_config_init(&cfg);
    cfg.steal = threshold < 100 ? 1 : 0;
    if (const char *v = getenv("DM_VWAVE")) cfg.vwave = atoi(v);
    const char *lbm = getenv("DM_LB");
    const bool host_lb = lbm && strcmp(lbm, "host") == 0;
    auto t1 = std::chrono::steady_clock::now();
    uint64_t count = 0;
    double gpu_ms = 0, lb_ms = 0;
    int rc;
    if (host_lb) {
        dm_dfs_config dc;
        dm_dfs_config_init(&dc);
        if (cfg.vwave) dc.vwave = cfg.vwave;
        dc.blocks = atoi(argv[3]) / 256 > 0 ? atoi(argv[3]) / 256 : 1;
        dc.jobs_per_worker = atoi(argv[5]) > 0 ? atoi(argv[5]) : 16;
        dc.threshold = threshold;
        dc.poll_us = 100000;
        if (const char *v = getenv("DM_LB_POLL_US")) dc.poll_us = atoi(v);
        dm_dfs *s = nullptr;
        const void *kern = nullptr;
        dm_dfs_stats ds = {};
        rc = dm_dfs_kernel(DM_DFS_CLIQUE_PA, dc.vwave, &kern);
        if (rc == DM_OK) rc = dm_dfs_create(g, k, 0, &dc, &s);
        if (rc == DM_OK) rc = dm_dfs_run(s, kern, &count, 1, &ds);
        dm_dfs_free(s);
        gpu_ms = ds.total_ms;
        lb_ms = ds.lb_ms;
    } else {
        rc = dm_clique_count(g, k, &cfg, &count);
    }
    auto t2 = std::chrono::steady_clock::now();
    if (rc != DM_OK) {
        fprintf(stderr, "%s\n", dm_last_error());
        dm_graph_free(g);
        return 1;
    }
    if (!host_lb) {
        dm_stats st;
        dm_last_stats(g, &st);
        gpu_ms = st.device_ms;
    }
    printf("Result: %lu.\n", (unsigned long)count);
    // DuMatoCPU::validateAggregateCounter (src/main/DuMatoCPU.cu:564-595),
    // which clique_PA never calls; DM_VALIDATE=1 calls it here.
    if (const char *v = getenv("DM_VALIDATE"); v && atoi(v)) {
        int match = 0;
        if (dm_validate_counter(path, k, count, &match) != DM_OK)
            fprintf(stderr, "Error opening the validation file of aggregate counter.\n");
        else
            printf("validateAggregateCounter: %s\n", match ? "MATCH" : "MISMATCH");
    }
    double gpu = gpu_ms / 1e3;
    printf("Time (GPU): %f secs.\n", gpu);
    printf("Time (LB): %f secs,%f%%\n", lb_ms / 1e3, gpu_ms > 0 ? 100.0 * lb_ms / gpu_ms : 0.0);
    printf("Time (IO): %f secs\n", secs(t0, t1));
    printf("Time (TOTAL): %f secs\n", secs(t0, t2));
    dm_graph_free(g);
    return 0;
}
