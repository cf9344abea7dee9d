// Cost-model MLP kernels (mlp.hip) and their host-side state.
#ifndef DM_MLP_H
#define DM_MLP_H

#include <cstdint>
#include <vector>

namespace dm {

constexpr int kMlpIn = 8;
constexpr int kMlpHidden = 64;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpHidden + 1;  // 4801

// Device scratch for one batch of up to cap samples.
struct MlpScratch {
    float *out = nullptr, *h1 = nullptr, *h2 = nullptr, *dz1 = nullptr, *dz2 = nullptr, *g = nullptr;
};

int mlp_forward_launch(const float *P, const float *X, int n, float *out, void *stream);
int mlp_grad_launch(const float *P, const float *X, const float *Y, int n, MlpScratch &s, float *grad,
                    float *loss, void *stream);
int mlp_adam_launch(float *P, const float *grad, float *M, float *V, float lr, float beta1, float beta2,
                    float eps, int step, void *stream);

// Trained model kept in a dm_graph: parameters plus the feature scaling.
struct CostModel {
    std::vector<float> params;      // kMlpParams
    float feat_mean[kMlpIn], feat_std[kMlpIn];
    float y_mean = 0, y_std = 1;    // target scaling (log cycles)
    int mode_motif = 0, k_train = 0;
    double final_loss = 0;
    float *d_params = nullptr, *d_x = nullptr, *d_out = nullptr;  // inference buffers
    float *h_out = nullptr;  // pinned host copy of the standardised outputs (n floats)
};

}  // namespace dm

#endif
