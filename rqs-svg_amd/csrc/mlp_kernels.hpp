// Launch interface of the cost-model MLP kernels (mlp_kernels.hip).
#pragma once
#include "sv_kernels.hpp"

namespace qsv
{
    constexpr int kMlpHidden = 64;
    constexpr int kMlpMaxFeatures = 32;

    int mlp_num_params(int F);
    // Y[b] = MLP(X[b]); H1/H2 (B x 64 each) receive the activations when non-null.
    void mlp_forward(hipStream_t s, const float *params, int F, const float *X, int B, float *H1, float *H2, float *Y);
    // Gradient of mean squared error (1/B) sum (Y - T)^2 with respect to all parameters.
    void mlp_gradients(hipStream_t s, const float *params, int F, const float *X, const float *T, int B, float *H1, float *H2, float *Y, float *dH1, float *dH2, float *dY, float *G);
    // Adam step `step` (1-based): beta1 0.9, beta2 0.999, eps 1e-8.
    void mlp_adam(hipStream_t s, float *params, const float *G, float *m, float *v, int P, float lr, int step);
} // namespace qsv
