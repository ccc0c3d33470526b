// Cost-model network of the learned fusion planner: a 3-layer MLP, 64 units wide
// (F -> 64 -> 64 -> 1, ReLU), FP32, trained on the GPU.
//
// Parameter layout (floats): W1[64][F], b1[64], W2[64][64], b2[64], w3[64], b3.
// Training step = forward (one thread per sample, activations kept for the backward
// pass) -> activation gradients (one thread per sample) -> parameter gradients (one thread
// per parameter, summing over the batch in sample order, so results do not depend on
// scheduling) -> Adam update (one thread per parameter).
#include "mlp_kernels.hpp"

#include <stdexcept>

namespace qsv
{
    namespace
    {
        constexpr int H = kMlpHidden;

        struct Layout
        {
            int F;
            __host__ __device__ int w1() const { return 0; }
            __host__ __device__ int b1() const { return H * F; }
            __host__ __device__ int w2() const { return H * F + H; }
            __host__ __device__ int b2() const { return H * F + H + H * H; }
            __host__ __device__ int w3() const { return H * F + 2 * H + H * H; }
            __host__ __device__ int b3() const { return H * F + 3 * H + H * H; }
        };

        __global__ void k_forward(const float *__restrict__ p, int F, const float *__restrict__ X, int B, float *__restrict__ H1, float *__restrict__ H2, float *__restrict__ Y)
        {
            const int b = blockIdx.x * blockDim.x + threadIdx.x;
            if (b >= B)
                return;
            const Layout L{F};
            const float *x = X + size_t(b) * F;
            float h1[H], h2[H];
#pragma unroll
            for (int i = 0; i < H; ++i)
            {
                float acc = p[L.b1() + i];
                for (int k = 0; k < F; ++k)
                    acc = __builtin_fmaf(p[L.w1() + i * F + k], x[k], acc);
                h1[i] = acc > 0.f ? acc : 0.f;
            }
#pragma unroll
            for (int i = 0; i < H; ++i)
            {
                float acc = p[L.b2() + i];
#pragma unroll
                for (int j = 0; j < H; ++j)
                    acc = __builtin_fmaf(p[L.w2() + i * H + j], h1[j], acc);
                h2[i] = acc > 0.f ? acc : 0.f;
            }
            float y = p[L.b3()];
#pragma unroll
            for (int j = 0; j < H; ++j)
                y = __builtin_fmaf(p[L.w3() + j], h2[j], y);
            Y[b] = y;
            if (H1)
            {
#pragma unroll
                for (int i = 0; i < H; ++i)
                {
                    H1[size_t(b) * H + i] = h1[i];
                    H2[size_t(b) * H + i] = h2[i];
                }
            }
        }

        // dY = 2 (y - t) / B ; dH2 = dY w3 (h2 > 0) ; dH1 = W2^T dH2 (h1 > 0)
        __global__ void k_backward_act(const float *__restrict__ p, int F, const float *__restrict__ T, int B, const float *__restrict__ H1, const float *__restrict__ H2, const float *__restrict__ Y, float *__restrict__ dH1, float *__restrict__ dH2, float *__restrict__ dY)
        {
            const int b = blockIdx.x * blockDim.x + threadIdx.x;
            if (b >= B)
                return;
            const Layout L{F};
            const float dy = 2.f * (Y[b] - T[b]) / float(B);
            dY[b] = dy;
            float d2[H];
#pragma unroll
            for (int i = 0; i < H; ++i)
            {
                d2[i] = H2[size_t(b) * H + i] > 0.f ? dy * p[L.w3() + i] : 0.f;
                dH2[size_t(b) * H + i] = d2[i];
            }
#pragma unroll
            for (int j = 0; j < H; ++j)
            {
                float acc = 0.f;
#pragma unroll
                for (int i = 0; i < H; ++i)
                    acc = __builtin_fmaf(p[L.w2() + i * H + j], d2[i], acc);
                dH1[size_t(b) * H + j] = H1[size_t(b) * H + j] > 0.f ? acc : 0.f;
            }
        }

        __global__ void k_param_grad(int F, const float *__restrict__ X, int B, const float *__restrict__ H1, const float *__restrict__ H2, const float *__restrict__ dH1, const float *__restrict__ dH2, const float *__restrict__ dY, float *__restrict__ G)
        {
            const Layout L{F};
            const int P = L.b3() + 1;
            const int idx = blockIdx.x * blockDim.x + threadIdx.x;
            if (idx >= P)
                return;
            float acc = 0.f;
            if (idx < L.b1())
            {
                const int i = idx / F, k = idx % F;
                for (int b = 0; b < B; ++b)
                    acc = __builtin_fmaf(dH1[size_t(b) * H + i], X[size_t(b) * F + k], acc);
            }
            else if (idx < L.w2())
            {
                const int i = idx - L.b1();
                for (int b = 0; b < B; ++b)
                    acc += dH1[size_t(b) * H + i];
            }
            else if (idx < L.b2())
            {
                const int i = (idx - L.w2()) / H, j = (idx - L.w2()) % H;
                for (int b = 0; b < B; ++b)
                    acc = __builtin_fmaf(dH2[size_t(b) * H + i], H1[size_t(b) * H + j], acc);
            }
            else if (idx < L.w3())
            {
                const int i = idx - L.b2();
                for (int b = 0; b < B; ++b)
                    acc += dH2[size_t(b) * H + i];
            }
            else if (idx < L.b3())
            {
                const int j = idx - L.w3();
                for (int b = 0; b < B; ++b)
                    acc = __builtin_fmaf(dY[b], H2[size_t(b) * H + j], acc);
            }
            else
            {
                for (int b = 0; b < B; ++b)
                    acc += dY[b];
            }
            G[idx] = acc;
        }

        __global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v, int P, float lr, float b1, float b2, float eps, float c1, float c2)
        {
            const int i = blockIdx.x * blockDim.x + threadIdx.x;
            if (i >= P)
                return;
            const float gi = g[i];
            const float mi = b1 * m[i] + (1.f - b1) * gi;
            const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
            m[i] = mi;
            v[i] = vi;
            p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        }
    } // namespace

    int mlp_num_params(int F) { return Layout{F}.b3() + 1; }

    void mlp_forward(hipStream_t s, const float *params, int F, const float *X, int B, float *H1, float *H2, float *Y)
    {
        if (F < 1 || F > kMlpMaxFeatures || B < 0)
            throw std::runtime_error("mlp: bad shape");
        if (B == 0)
            return;
        hipLaunchKernelGGL(k_forward, dim3((B + 127) / 128), dim3(128), 0, s, params, F, X, B, H1, H2, Y);
        check(hipGetLastError(), "mlp forward");
    }

    void mlp_gradients(hipStream_t s, const float *params, int F, const float *X, const float *T, int B, float *H1, float *H2, float *Y, float *dH1, float *dH2, float *dY, float *G)
    {
        if (F < 1 || F > kMlpMaxFeatures || B < 1)
            throw std::runtime_error("mlp: bad shape");
        mlp_forward(s, params, F, X, B, H1, H2, Y);
        hipLaunchKernelGGL(k_backward_act, dim3((B + 127) / 128), dim3(128), 0, s, params, F, T, B, H1, H2, Y, dH1, dH2, dY);
        check(hipGetLastError(), "mlp backward");
        const int P = mlp_num_params(F);
        hipLaunchKernelGGL(k_param_grad, dim3((P + 255) / 256), dim3(256), 0, s, F, X, B, H1, H2, dH1, dH2, dY, G);
        check(hipGetLastError(), "mlp gradient");
    }

    void mlp_adam(hipStream_t s, float *params, const float *G, float *m, float *v, int P, float lr, int step)
    {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float c1 = 1.f - powf(b1, float(step)), c2 = 1.f - powf(b2, float(step));
        hipLaunchKernelGGL(k_adam, dim3((P + 255) / 256), dim3(256), 0, s, params, G, m, v, P, lr, b1, b2, eps, c1, c2);
        check(hipGetLastError(), "mlp adam");
    }
} // namespace qsv
