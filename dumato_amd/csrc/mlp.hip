// Cost-model MLP (8 -> 64 -> 64 -> 1, ReLU, FP32) with hand-written forward,
// backward and Adam kernels.  No reference counterpart: it predicts the
// enumeration cost of each root so the scheduler can order, split and bin
// roots.  It never touches the counting arithmetic.
//
// Parameter vector layout (kMlpParams floats):
//   W1[8][64], b1[64], W2[64][64], b2[64], W3[64], b3
// Row-major W1[i][j] = weight from input i to hidden unit j.
#include <hip/hip_runtime.h>

#include "internal.h"
#include "mlp.h"

namespace dm {

constexpr int kIn = kMlpIn, kH = kMlpHidden;
constexpr int oW1 = 0, ob1 = oW1 + kIn * kH, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH,
              ob3 = oW3 + kH;
static_assert(ob3 + 1 == kMlpParams, "parameter layout");

// Forward for samples [0, n): out[b]; when train != 0 also stores the hidden
// activations and the per-sample backward terms (dz1, dz2, g) for the loss
// sum_b (out_b - y_b)^2 / n.
__global__ __launch_bounds__(256) void mlp_fwd_bwd(const float *__restrict__ P, const float *__restrict__ X,
                                                   const float *__restrict__ Y, int n, int train,
                                                   float *__restrict__ out, float *__restrict__ H1,
                                                   float *__restrict__ H2, float *__restrict__ DZ1,
                                                   float *__restrict__ DZ2, float *__restrict__ G,
                                                   float *__restrict__ loss) {
    __shared__ float sP[kMlpParams];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sP[i] = P[i];
    __syncthreads();
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    float x[kIn];
#pragma unroll
    for (int i = 0; i < kIn; i++) x[i] = X[(size_t)b * kIn + i];
    float h1[kH];
#pragma unroll
    for (int j = 0; j < kH; j++) {
        float z = sP[ob1 + j];
#pragma unroll
        for (int i = 0; i < kIn; i++) z += x[i] * sP[oW1 + i * kH + j];
        h1[j] = z > 0.f ? z : 0.f;
    }
    float o = sP[ob3];
    float h2[kH];
#pragma unroll 4
    for (int j = 0; j < kH; j++) {
        float z = sP[ob2 + j];
#pragma unroll
        for (int i = 0; i < kH; i++) z += h1[i] * sP[oW2 + i * kH + j];
        h2[j] = z > 0.f ? z : 0.f;
        o += h2[j] * sP[oW3 + j];
    }
    out[b] = o;
    if (!train) return;
    const float diff = o - Y[b];
    const float g = 2.f * diff / (float)n;
    atomicAdd(loss, diff * diff / (float)n);
    G[b] = g;
    // dz2_j = g * W3_j * [h2_j > 0];  dz1_i = sum_j W2[i][j] dz2_j * [h1_i > 0]
    float dz2[kH];
#pragma unroll
    for (int j = 0; j < kH; j++) {
        dz2[j] = h2[j] > 0.f ? g * sP[oW3 + j] : 0.f;
        H2[(size_t)b * kH + j] = h2[j];
        DZ2[(size_t)b * kH + j] = dz2[j];
    }
#pragma unroll 4
    for (int i = 0; i < kH; i++) {
        float s = 0.f;
        if (h1[i] > 0.f) {
#pragma unroll
            for (int j = 0; j < kH; j++) s += sP[oW2 + i * kH + j] * dz2[j];
        }
        H1[(size_t)b * kH + i] = h1[i];
        DZ1[(size_t)b * kH + i] = s;
    }
}

// One block per parameter: its gradient is a sum over the batch, reduced
// across the block's lanes (wave shuffles, then LDS across waves).
__global__ __launch_bounds__(256) void mlp_grad(const float *__restrict__ X, const float *__restrict__ H1,
                                                const float *__restrict__ H2, const float *__restrict__ DZ1,
                                                const float *__restrict__ DZ2, const float *__restrict__ G,
                                                int n, float *__restrict__ grad) {
    const int p = blockIdx.x;
    float s = 0.f;
    if (p < ob1) {
        const int i = p / kH, j = p % kH;
        for (int b = threadIdx.x; b < n; b += blockDim.x) s += X[(size_t)b * kIn + i] * DZ1[(size_t)b * kH + j];
    } else if (p < oW2) {
        const int j = p - ob1;
        for (int b = threadIdx.x; b < n; b += blockDim.x) s += DZ1[(size_t)b * kH + j];
    } else if (p < ob2) {
        const int i = (p - oW2) / kH, j = (p - oW2) % kH;
        for (int b = threadIdx.x; b < n; b += blockDim.x) s += H1[(size_t)b * kH + i] * DZ2[(size_t)b * kH + j];
    } else if (p < oW3) {
        const int j = p - ob2;
        for (int b = threadIdx.x; b < n; b += blockDim.x) s += DZ2[(size_t)b * kH + j];
    } else if (p < ob3) {
        const int j = p - oW3;
        for (int b = threadIdx.x; b < n; b += blockDim.x) s += H2[(size_t)b * kH + j] * G[b];
    } else {
        for (int b = threadIdx.x; b < n; b += blockDim.x) s += G[b];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) grad[p] = part[0] + part[1] + part[2] + part[3];
}

// Adam (Kingma & Ba 2015), bias-corrected, one thread per parameter.
__global__ void mlp_adam(float *__restrict__ P, const float *__restrict__ grad, float *__restrict__ M,
                         float *__restrict__ V, float lr, float beta1, float beta2, float eps, float c1,
                         float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kMlpParams) return;
    const float g = grad[p];
    const float m = beta1 * M[p] + (1.f - beta1) * g;
    const float v = beta2 * V[p] + (1.f - beta2) * g * g;
    M[p] = m;
    V[p] = v;
    const float mh = m / c1, vh = v / c2;
    P[p] -= lr * mh / (sqrtf(vh) + eps);
}

int mlp_forward_launch(const float *P, const float *X, int n, float *out, void *stream) {
    if (n <= 0) return DM_OK;
    hipLaunchKernelGGL(mlp_fwd_bwd, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, P, X,
                       (const float *)nullptr, n, 0, out, nullptr, nullptr, nullptr, nullptr, nullptr,
                       nullptr);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? DM_OK : set_error(DM_ERR_HIP, "mlp forward: %s", hipGetErrorString(e));
}

int mlp_grad_launch(const float *P, const float *X, const float *Y, int n, MlpScratch &s, float *grad,
                    float *loss, void *stream) {
    if (n <= 0) return DM_OK;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(mlp_fwd_bwd, dim3((n + 255) / 256), dim3(256), 0, st, P, X, Y, n, 1, s.out, s.h1, s.h2,
                       s.dz1, s.dz2, s.g, loss);
    hipLaunchKernelGGL(mlp_grad, dim3(kMlpParams), dim3(256), 0, st, X, s.h1, s.h2, s.dz1, s.dz2, s.g, n, grad);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? DM_OK : set_error(DM_ERR_HIP, "mlp backward: %s", hipGetErrorString(e));
}

int mlp_adam_launch(float *P, const float *grad, float *M, float *V, float lr, float beta1, float beta2,
                    float eps, int step, void *stream) {
    float c1 = 1.f, c2 = 1.f;
    for (int i = 0; i < step; i++) {
        c1 *= beta1;
        c2 *= beta2;
    }
    c1 = 1.f - c1;
    c2 = 1.f - c2;
    hipLaunchKernelGGL(mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, (hipStream_t)stream, P, grad, M,
                       V, lr, beta1, beta2, eps, c1, c2);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? DM_OK : set_error(DM_ERR_HIP, "mlp adam: %s", hipGetErrorString(e));
}

}  // namespace dm
