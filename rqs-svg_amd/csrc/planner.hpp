// Learned fusion planner: an MLP cost model of block-kernel time, trained on timings the
// simulator records, choosing among candidate fusion plans; the greedy plan is the
// fallback and the baseline every choice is judged against.
#pragma once
#include "fused.hpp"
#include "mlp_kernels.hpp"

#include <hip/hip_runtime.h>

#include <vector>

namespace qcore
{
    constexpr int kPlanFeatures = 20;

    // Features of one launch that sweeps the state (all in [0, ~1]; list in planner.cpp):
    // a block launch (lds: the LDS-staged block kernel) or a stage launch.
    void block_features(int n, float math, int kq, const BlockLayout &lay, int gates, bool lds, float *f);
    void stage_features(int n, float math, int kq, const std::vector<int> &tile, int nblocks, int gates, float *f);
    // Whether a block with this layout runs in the LDS-staged block kernel (enough free
    // qubits outside the block and its controls for the tile's column bits).
    bool block_uses_lds(const BlockLayout &lay, int n, bool f64);

    class CostModel
    {
    public:
        CostModel(int num_features, uint64_t seed);
        ~CostModel();
        CostModel(const CostModel &) = delete;
        CostModel &operator=(const CostModel &) = delete;

        int num_features() const { return F_; }
        int num_params() const { return P_; }
        void get_params(float *out) const;
        void set_params(const float *in);
        // predictions for B rows of features (host in/out), one GPU forward pass
        void predict(const float *X, int B, float *Y);
        // gradient of the mean squared error over the batch; returns the loss
        float gradients(const float *X, const float *T, int B, float *G);
        // Adam on shuffled mini-batches; returns the final mean squared error on all data
        float train(const float *X, const float *T, int N, int epochs, int batch, float lr, uint64_t seed);
        // The factor a candidate plan's predicted time must beat the default plan's by to be
        // chosen: 2^-rms, rms = the root mean squared error (log2 microseconds) of the last
        // training on its own data -- a plan is taken only when its predicted gain exceeds
        // the model's typical error -- kept within [0.75, 0.97]; 0.97 when the parameters
        // were set directly (no fit known).
        double choice_factor() const;

    private:
        void reserve(int B);
        int F_, P_;
        hipStream_t stream_ = nullptr;
        float *params_ = nullptr, *grad_ = nullptr, *m_ = nullptr, *v_ = nullptr;
        float *X_ = nullptr, *T_ = nullptr, *H1_ = nullptr, *H2_ = nullptr, *Y_ = nullptr;
        float *dH1_ = nullptr, *dH2_ = nullptr, *dY_ = nullptr;
        int cap_ = 0;
        int step_ = 0;
        float fit_mse_ = -1.f; // of the last train(); -1 after set_params
    };

    std::vector<PlannedBlock> plan_with(const std::vector<ResolvedGate> &gates, const PlanOption &opt);
    // Block plans: k from max(2, kq - 2) to kq x windows {4, 64} x rules {0, 1, 2}; stage
    // plans: each tile size in `tile_bits` x each block size in `stage_k` x tile choice
    // {grow, window}.
    std::vector<PlanOption> candidate_options(int kq, const std::vector<int> &tile_bits, const std::vector<int> &stage_k);
} // namespace qcore
