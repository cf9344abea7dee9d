// Scheduling policies above the plain runtime: the virtual-wave width choice
// and the learned cost-model partition.  Neither touches the counting
// arithmetic: they only choose launch shapes and how roots are ordered,
// split into child ranges and dealt to workers.
#include <algorithm>
#include <cmath>

#include "internal.h"
#include "runtime.h"

namespace dm {

// Width used when the caller asks for autotuning and no tuned value is
// cached for (mode, k): a worker's lanes stride adjacency lists, so the width
// follows the typical list length (rounded up to a power of two, 8..64).
int autotuned_vwave(dm_graph *g, int K, bool motif) {
    int t = g->tuned_vwave[motif ? 1 : 0][K];
    if (t) return t;
    const HostCSR &c = g->rcsr;
    double avg = c.n ? (double)c.nnz() / (double)c.n : 0.0;
    if (!motif) avg *= 0.5;
    int w = 8;
    while (w < 64 && w < avg) w *= 2;
    return w;
}

void free_model(dm_graph *g) { (void)g; }

int learned_partition(dm_graph *g, const std::vector<int32_t> &roots, int K, bool motif, const dm_config &cfg,
                      const Shape &sh, std::vector<int64_t> &sq_off, std::vector<int32_t> &sq_jobs) {
    (void)g; (void)roots; (void)K; (void)motif; (void)cfg; (void)sh; (void)sq_off; (void)sq_jobs;
    return set_error(DM_ERR_ARG, "learned scheduler: no trained model on this graph");
}

}  // namespace dm

extern "C" int dm_sched_train(dm_graph *g, int mode_motif, int k_train, int epochs, uint64_t seed,
                              double *final_loss) {
    (void)g; (void)mode_motif; (void)k_train; (void)epochs; (void)seed; (void)final_loss;
    return dm::set_error(DM_ERR_UNSUPPORTED, "not built yet");
}

extern "C" int dm_sched_predict(dm_graph *g, float *pred, size_t n) {
    (void)g; (void)pred; (void)n;
    return dm::set_error(DM_ERR_UNSUPPORTED, "not built yet");
}
