// Learned fusion planner: features, cost model (GPU training), candidate plans.
#include "planner.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace qcore
{
    /* Features of one launch (kPlanFeatures = 20):
     *  0 n/32      1 kq/5      2 arithmetic: complex128 FP64 1, complex64 BF16x3 0.5, FP32 0
     *  3 LDS-staged block kernel      4 stage kernel
     *  5..9  block launch: block qubits ascending / 32 (missing entries n/32);
     *        stage launch: the 5 highest tile qubits ascending / 32
     * 10 block qubits below 3 / 5      11 block qubits below 6 / 5      (stages: 0)
     * 12 controls outside the block / 4                                  (stages: 0)
     * 13 gates / 32 (stages: gates / 128)
     * 14 log2(columns) / 32 (stages: log2(tiles) / 32)
     * 15 lowest qubit outside the block / 8                              (stages: 0)
     * 16 blocks in the launch / 16     17 tile bits / 16 (blocks: 0)
     * 18 gates per block / 32          19 constant 1 */
    void block_features(int n, float math, int kq, const BlockLayout &lay, int gates, bool lds, float *f)
    {
        std::vector<int> q = lay.qubits;
        std::sort(q.begin(), q.end());
        const int outside = (int)(lay.pos_out.size() + lay.neg_out.size());
        int lowest_free = 0;
        while (std::find(q.begin(), q.end(), lowest_free) != q.end() ||
               std::find(lay.pos_out.begin(), lay.pos_out.end(), lowest_free) != lay.pos_out.end() ||
               std::find(lay.neg_out.begin(), lay.neg_out.end(), lowest_free) != lay.neg_out.end())
            ++lowest_free;
        std::fill(f, f + kPlanFeatures, 0.f);
        f[0] = n / 32.f;
        f[1] = kq / 5.f;
        f[2] = math;
        f[3] = lds ? 1.f : 0.f;
        f[4] = 0.f;
        for (int i = 0; i < 5; ++i)
            f[5 + i] = (i < (int)q.size() ? q[i] : n) / 32.f;
        int below3 = 0, below6 = 0;
        for (int x : q)
        {
            below3 += x < 3;
            below6 += x < 6;
        }
        f[10] = below3 / 5.f;
        f[11] = below6 / 5.f;
        f[12] = outside / 4.f;
        f[13] = gates / 32.f;
        f[14] = (n - (int)q.size() - outside) / 32.f;
        f[15] = lowest_free / 8.f;
        f[16] = 1.f / 16.f;
        f[17] = 0.f;
        f[18] = gates / 32.f;
        f[19] = 1.f;
    }

    void stage_features(int n, float math, int kq, const std::vector<int> &tile, int nblocks, int gates, float *f)
    {
        std::fill(f, f + kPlanFeatures, 0.f);
        const int m = (int)tile.size();
        f[0] = n / 32.f;
        f[1] = kq / 5.f;
        f[2] = math;
        f[3] = 0.f;
        f[4] = 1.f;
        for (int i = 0; i < 5; ++i)
            f[5 + i] = (m - 5 + i >= 0 ? tile[m - 5 + i] : n) / 32.f;
        f[13] = gates / 128.f;
        f[14] = (n - m) / 32.f;
        f[15] = float(nblocks << kq) / 256.f; // arithmetic per amplitude, relative (blocks x 2^k)
        f[16] = nblocks / 16.f;
        f[17] = m / 16.f;
        f[18] = nblocks ? float(gates) / nblocks / 32.f : 0.f;
        f[19] = 1.f;
    }

    bool block_uses_lds(const BlockLayout &lay, int n, bool f64)
    {
        const int free_q = n - (int)lay.qubits.size() - (int)(lay.pos_out.size() + lay.neg_out.size());
        return free_q >= (f64 ? 6 : 5);
    }

    std::vector<PlanOption> candidate_options(int kq, const std::vector<int> &tile_bits, const std::vector<int> &stage_k)
    {
        std::vector<PlanOption> out;
        for (int k = std::max(2, kq - 2); k <= kq; ++k)
            for (int w : {4, 64})
                for (int rule : {0, 1, 2})
                    out.push_back(PlanOption{0, k, w, rule});
        for (int m : tile_bits)
            for (int k : stage_k)
                for (int tiles : {0, 1})
                    out.push_back(PlanOption{m, k, 64, 0, tiles});
        return out;
    }

    std::vector<PlannedBlock> plan_with(const std::vector<ResolvedGate> &gates, const PlanOption &o)
    {
        return plan_rule(gates, o.k, o.window, o.rule);
    }

    // ---- cost model ----
    namespace
    {
        uint64_t splitmix(uint64_t &s)
        {
            uint64_t z = (s += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        }
        double unif(uint64_t &s) { return double(splitmix(s) >> 11) * 0x1.0p-53; }
    } // namespace

    CostModel::CostModel(int F, uint64_t seed) : F_(F)
    {
        if (F < 1 || F > qsv::kMlpMaxFeatures)
            throw std::runtime_error("cost model: feature count out of range");
        P_ = qsv::mlp_num_params(F);
        qsv::check(hipStreamCreate(&stream_), "hipStreamCreate");
        for (float **p : {&params_, &grad_, &m_, &v_})
            qsv::check(hipMalloc((void **)p, P_ * sizeof(float)), "hipMalloc mlp");
        // He-uniform weights from SplitMix64, zero biases
        std::vector<float> init(P_, 0.f);
        const int H = qsv::kMlpHidden;
        auto fill = [&](int off, int count, int fan_in) {
            const double a = std::sqrt(6.0 / fan_in);
            for (int i = 0; i < count; ++i)
                init[off + i] = float((2 * unif(seed) - 1) * a);
        };
        fill(0, H * F, F);
        fill(H * F + H, H * H, H);
        fill(H * F + 2 * H + H * H, H, H);
        set_params(init.data());
        std::vector<float> zero(P_, 0.f);
        qsv::check(hipMemcpy(m_, zero.data(), P_ * sizeof(float), hipMemcpyHostToDevice), "mlp init");
        qsv::check(hipMemcpy(v_, zero.data(), P_ * sizeof(float), hipMemcpyHostToDevice), "mlp init");
    }

    CostModel::~CostModel()
    {
        (void)hipStreamSynchronize(stream_);
        for (float *p : {params_, grad_, m_, v_, X_, T_, H1_, H2_, Y_, dH1_, dH2_, dY_})
            if (p)
                (void)hipFree(p);
        (void)hipStreamDestroy(stream_);
    }

    void CostModel::reserve(int B)
    {
        if (B <= cap_)
            return;
        qsv::check(hipStreamSynchronize(stream_), "sync");
        for (float *p : {X_, T_, H1_, H2_, Y_, dH1_, dH2_, dY_})
            if (p)
                (void)hipFree(p);
        const int H = qsv::kMlpHidden;
        qsv::check(hipMalloc((void **)&X_, size_t(B) * F_ * sizeof(float)), "hipMalloc mlp");
        qsv::check(hipMalloc((void **)&T_, size_t(B) * sizeof(float)), "hipMalloc mlp");
        qsv::check(hipMalloc((void **)&Y_, size_t(B) * sizeof(float)), "hipMalloc mlp");
        qsv::check(hipMalloc((void **)&dY_, size_t(B) * sizeof(float)), "hipMalloc mlp");
        for (float **p : {&H1_, &H2_, &dH1_, &dH2_})
            qsv::check(hipMalloc((void **)p, size_t(B) * H * sizeof(float)), "hipMalloc mlp");
        cap_ = B;
    }

    void CostModel::get_params(float *out) const
    {
        qsv::check(hipMemcpy(out, params_, P_ * sizeof(float), hipMemcpyDeviceToHost), "mlp params");
    }

    void CostModel::set_params(const float *in)
    {
        qsv::check(hipMemcpy(params_, in, P_ * sizeof(float), hipMemcpyHostToDevice), "mlp params");
        step_ = 0;
        fit_mse_ = -1.f;
    }

    double CostModel::choice_factor() const
    {
        if (!(fit_mse_ > 0.f))
            return 0.97;
        return std::min(0.97, std::max(0.75, std::exp2(-std::sqrt(double(fit_mse_)))));
    }

    void CostModel::predict(const float *X, int B, float *Y)
    {
        if (B <= 0)
            return;
        reserve(B);
        qsv::check(hipMemcpyAsync(X_, X, size_t(B) * F_ * sizeof(float), hipMemcpyHostToDevice, stream_), "mlp upload");
        qsv::mlp_forward(stream_, params_, F_, X_, B, nullptr, nullptr, Y_);
        qsv::check(hipMemcpyAsync(Y, Y_, size_t(B) * sizeof(float), hipMemcpyDeviceToHost, stream_), "mlp download");
        qsv::check(hipStreamSynchronize(stream_), "sync");
    }

    float CostModel::gradients(const float *X, const float *T, int B, float *G)
    {
        if (B <= 0)
            throw std::runtime_error("cost model: empty batch");
        reserve(B);
        qsv::check(hipMemcpyAsync(X_, X, size_t(B) * F_ * sizeof(float), hipMemcpyHostToDevice, stream_), "mlp upload");
        qsv::check(hipMemcpyAsync(T_, T, size_t(B) * sizeof(float), hipMemcpyHostToDevice, stream_), "mlp upload");
        qsv::mlp_gradients(stream_, params_, F_, X_, T_, B, H1_, H2_, Y_, dH1_, dH2_, dY_, grad_);
        std::vector<float> y(B);
        qsv::check(hipMemcpyAsync(y.data(), Y_, size_t(B) * sizeof(float), hipMemcpyDeviceToHost, stream_), "mlp download");
        if (G)
            qsv::check(hipMemcpyAsync(G, grad_, P_ * sizeof(float), hipMemcpyDeviceToHost, stream_), "mlp download");
        qsv::check(hipStreamSynchronize(stream_), "sync");
        double loss = 0;
        for (int b = 0; b < B; ++b)
            loss += double(y[b] - T[b]) * double(y[b] - T[b]);
        return float(loss / B);
    }

    float CostModel::train(const float *X, const float *T, int N, int epochs, int batch, float lr, uint64_t seed)
    {
        if (N <= 0 || epochs < 0 || batch < 1)
            throw std::runtime_error("cost model: bad training arguments");
        batch = std::min(batch, N);
        reserve(N);
        std::vector<int> idx(N);
        std::iota(idx.begin(), idx.end(), 0);
        std::vector<float> xs(size_t(N) * F_), ts(N);
        float *dX = nullptr, *dT = nullptr;
        qsv::check(hipMalloc((void **)&dX, size_t(N) * F_ * sizeof(float)), "hipMalloc mlp");
        qsv::check(hipMalloc((void **)&dT, size_t(N) * sizeof(float)), "hipMalloc mlp");
        try
        {
            for (int e = 0; e < epochs; ++e)
            {
                for (int i = N - 1; i > 0; --i)
                    std::swap(idx[i], idx[splitmix(seed) % uint64_t(i + 1)]);
                for (int i = 0; i < N; ++i)
                {
                    std::memcpy(&xs[size_t(i) * F_], X + size_t(idx[i]) * F_, F_ * sizeof(float));
                    ts[i] = T[idx[i]];
                }
                qsv::check(hipMemcpyAsync(dX, xs.data(), xs.size() * sizeof(float), hipMemcpyHostToDevice, stream_), "mlp upload");
                qsv::check(hipMemcpyAsync(dT, ts.data(), ts.size() * sizeof(float), hipMemcpyHostToDevice, stream_), "mlp upload");
                for (int b0 = 0; b0 + batch <= N; b0 += batch)
                {
                    qsv::mlp_gradients(stream_, params_, F_, dX + size_t(b0) * F_, dT + b0, batch, H1_, H2_, Y_, dH1_, dH2_, dY_, grad_);
                    qsv::mlp_adam(stream_, params_, grad_, m_, v_, P_, lr, ++step_);
                }
                // the host buffers are rewritten next epoch: wait for the uploads
                qsv::check(hipStreamSynchronize(stream_), "sync");
            }
            // final loss on all data
            qsv::mlp_forward(stream_, params_, F_, dX, N, nullptr, nullptr, Y_);
            std::vector<float> y(N);
            qsv::check(hipMemcpyAsync(y.data(), Y_, size_t(N) * sizeof(float), hipMemcpyDeviceToHost, stream_), "mlp download");
            qsv::check(hipStreamSynchronize(stream_), "sync");
            double loss = 0;
            for (int i = 0; i < N; ++i)
                loss += double(y[i] - ts[i]) * double(y[i] - ts[i]);
            (void)hipFree(dX);
            (void)hipFree(dT);
            fit_mse_ = float(loss / N);
            return fit_mse_;
        }
        catch (...)
        {
            (void)hipFree(dX);
            (void)hipFree(dT);
            throw;
        }
    }
} // namespace qcore
