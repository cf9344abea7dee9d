// Kernels written against the step API of include/dumato_device.h, in the
// shape of the reference's user kernels:
//   dfs_clique_pa   src/main/clique_PA.cu:9-32
//   dfs_clique_wc   src/DM_WCV/clique_DM_WCV8.cu:8-34
//   dfs_motifs      src/main/motifs_DM_WCLB.cu:9-36
//   dfs_q_pa        src/main/q4_PA.cu:9-32 .. q13_PA.cu (near-clique query)
// They run under dm_dfs_run (host-driven load balancing).  The fused kernels
// of enum_kernel.hip compute the same counts faster; these exist so that the
// reference's programming model, and its stop/snapshot/donate balancing,
// have a drop-in here.
#include "../../include/dumato_device.h"
#include "internal.h"

namespace dm {

template <int W>
__global__ __launch_bounds__(256) void dfs_clique_pa(dm_dfs_data d) {
    dm_device::Worker<W> DM;
    DM.start_clique(d);
    while (DM.active() && DM.balanced()) {
        if (DM.getCurrentNumberOfExtensions() == -1) DM.extend_clique();
        if (DM.getCurrentNumberOfExtensions() != 0) {
            if (DM.last_level()) {
                DM.aggregate_counter();
                DM.backward_clique();
            } else {
                DM.forward();
            }
        } else {
            DM.backward_clique();
        }
    }
    DM.end();
}

template <int W>
__global__ __launch_bounds__(256) void dfs_clique_wc(dm_dfs_data d) {
    dm_device::Worker<W> DM;
    DM.start(d);
    while (DM.active() && DM.balanced()) {
        if (DM.getCurrentNumberOfExtensions() == -1) {
            DM.extend();
            if (DM.k() > 0) DM.filterClique();
        }
        if (DM.getCurrentNumberOfExtensions() != 0) {
            if (DM.last_level()) {
                DM.aggregate_counter();
                DM.backward();
            } else {
                DM.forward();
            }
        } else {
            DM.backward();
        }
    }
    DM.end();
}

template <int W>
__global__ __launch_bounds__(256) void dfs_motifs(dm_dfs_data d) {
    dm_device::Worker<W> DM;
    DM.start_induce(d);
    while (DM.active() && DM.balanced()) {
        if (DM.getCurrentNumberOfExtensions() == -1) {
            DM.extend(0, DM.k());
            if (DM.k() > 0) DM.canonicalFilter2();
        }
        if (DM.getCurrentNumberOfExtensions() != 0) {
            if (DM.last_level()) {
                DM.aggregate_pattern();
                DM.backward_induce();
            } else {
                DM.forward_induce();
            }
        } else {
            DM.backward_induce();
        }
    }
    DM.end();
}

// q4_PA .. q13_PA (src/main/q4_PA.cu:9-32): the near-clique query with the
// state's K
template <int W>
__global__ __launch_bounds__(256) void dfs_q_pa(dm_dfs_data d) {
    dm_device::Worker<W> DM;
    DM.start_q(d);
    while (DM.active() && DM.balanced()) {
        if (DM.getCurrentNumberOfExtensions() == -1) DM.extend_q();
        if (DM.getCurrentNumberOfExtensions() != 0) {
            if (DM.last_level()) {
                DM.aggregate_counter();
                DM.backward_q();
            } else {
                DM.forward();
            }
        } else {
            DM.backward_q();
        }
    }
    DM.end();
}

template <int W>
static const void *dfs_kernel_w(int kind) {
    switch (kind) {
        case DM_DFS_CLIQUE_PA: return (const void *)dfs_clique_pa<W>;
        case DM_DFS_CLIQUE_WC: return (const void *)dfs_clique_wc<W>;
        case DM_DFS_MOTIF: return (const void *)dfs_motifs<W>;
        case DM_DFS_NEAR_CLIQUE: return (const void *)dfs_q_pa<W>;
        default: return nullptr;
    }
}

}  // namespace dm

extern "C" int dm_dfs_kernel(int kind, int vwave, const void **kernel) {
    using namespace dm;
    if (!kernel) return set_error(DM_ERR_ARG, "kernel is NULL");
    const void *k = nullptr;
    switch (vwave) {
        case 8: k = dfs_kernel_w<8>(kind); break;
        case 16: k = dfs_kernel_w<16>(kind); break;
        case 32: k = dfs_kernel_w<32>(kind); break;
        case 64: k = dfs_kernel_w<64>(kind); break;
        default: return set_error(DM_ERR_ARG, "vwave must be 8, 16, 32 or 64 (got %d)", vwave);
    }
    if (!k) return set_error(DM_ERR_ARG, "unknown step-API kernel %d", kind);
    *kernel = k;
    return DM_OK;
}
