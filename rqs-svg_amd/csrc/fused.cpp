// Precision form: gate queue, greedy fusion planner, FP64 block composition, block launches
// on the complex64 state.  FP64 is used for the block matrices (host) and, through the
// shared support kernels, for every reduction, prefix sum and renormalisation; the state
// itself and the block products are FP32.
#include "fused.hpp"
#include "planner.hpp"

#include <atomic>
#include <exception>
#include <mutex>
#include <thread>

#include <qcs_ext.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>

namespace qcore
{
    namespace
    {
        std::vector<int> gate_qubits(const ResolvedGate &g)
        {
            std::vector<int> q = g.tphys;
            q.insert(q.end(), g.pphys.begin(), g.pphys.end());
            q.insert(q.end(), g.nphys.begin(), g.nphys.end());
            return q;
        }

        int union_size(const std::vector<int> &a, const std::vector<int> &b)
        {
            int extra = 0;
            for (int x : b)
                if (std::find(a.begin(), a.end(), x) == a.end())
                    ++extra;
            return (int)a.size() + extra;
        }

        int local_pos(const std::vector<int> &qubits, int q)
        {
            auto it = std::find(qubits.begin(), qubits.end(), q);
            return it == qubits.end() ? -1 : int(it - qubits.begin());
        }
    } // namespace

    // Dependency-aware greedy fusion.  A gate may join any block at or after the last
    // block touching one of its qubits (blocks in between touch none of them, so the gate
    // commutes past them); among the blocks within `window` of that frontier that stay
    // within k qubits, rule 0 takes the one that grows least, rule 1 the fullest one,
    // rule 2 the earliest one; with none, the gate opens a new block.
    std::vector<PlannedBlock> plan_rule(const std::vector<ResolvedGate> &gates, int k, int window, int rule)
    {
        std::vector<PlannedBlock> blocks;
        std::vector<int> last(qsv::kMaxQubits, -1);
        for (int gi = 0; gi < (int)gates.size(); ++gi)
        {
            auto Q = gate_qubits(gates[gi]);
            int lo = -1;
            for (int q : Q)
                lo = std::max(lo, last[q]);
            int chosen = -1;
            if ((int)Q.size() <= k)
            {
                int best_cost = 1 << 30;
                int start = std::max(lo, 0);
                int stop = std::min((int)blocks.size(), start + window);
                for (int j = start; j < stop; ++j)
                {
                    if (blocks[j].wide)
                        continue;
                    int u = union_size(blocks[j].qubits, Q);
                    if (u > k)
                        continue;
                    int cost = rule == 0 ? u - (int)blocks[j].qubits.size() : rule == 1 ? -u : j;
                    if (cost < best_cost)
                    {
                        best_cost = cost;
                        chosen = j;
                    }
                }
            }
            if (chosen < 0)
            {
                PlannedBlock b;
                b.wide = (int)Q.size() > k;
                blocks.push_back(b);
                chosen = (int)blocks.size() - 1;
            }
            auto &b = blocks[chosen];
            for (int q : Q)
                if (std::find(b.qubits.begin(), b.qubits.end(), q) == b.qubits.end())
                    b.qubits.push_back(q);
            b.gates.push_back(gi);
            for (int q : Q)
                last[q] = chosen;
        }
        return blocks;
    }

    std::vector<PlannedBlock> plan_greedy(const std::vector<ResolvedGate> &gates, int k)
    {
        return plan_rule(gates, k, 64, 0);
    }

    BlockLayout layout_block(const std::vector<ResolvedGate> &gates, const PlannedBlock &b, int kq, int n)
    {
        BlockLayout lay;
        if (!b.wide)
        {
            lay.qubits = b.qubits;
        }
        else
        {
            // one gate wider than the block: targets inside, controls outside as far as the
            // free qubits allow padding; otherwise some controls move inside
            const ResolvedGate &g = gates[b.gates[0]];
            lay.qubits = g.tphys;
            std::vector<int> ctrls = g.pphys;
            ctrls.insert(ctrls.end(), g.nphys.begin(), g.nphys.end());
            int free_q = n - (int)g.tphys.size() - (int)ctrls.size();
            int need = kq - (int)lay.qubits.size();
            int inside = std::max(0, need - free_q);
            for (int i = 0; i < inside; ++i)
                lay.qubits.push_back(ctrls[i]);
            for (size_t i = inside; i < ctrls.size(); ++i)
            {
                int q = ctrls[i];
                if (std::find(g.pphys.begin(), g.pphys.end(), q) != g.pphys.end())
                    lay.pos_out.push_back(q);
                else
                    lay.neg_out.push_back(q);
            }
        }
        if ((int)lay.qubits.size() > kq)
            throw std::runtime_error("fusion: block larger than the kernel block size");
        // pad with the highest free qubits (low qubits stay in the column index, which keeps
        // consecutive lanes on consecutive amplitudes)
        for (int q = n - 1; q >= 0 && (int)lay.qubits.size() < kq; --q)
        {
            bool used = local_pos(lay.qubits, q) >= 0 || local_pos(lay.pos_out, q) >= 0 || local_pos(lay.neg_out, q) >= 0;
            if (!used)
                lay.qubits.push_back(q);
        }
        if ((int)lay.qubits.size() != kq)
            throw std::runtime_error("fusion: not enough qubits to pad the block");
        return lay;
    }

    std::vector<cd> block_matrix(const std::vector<ResolvedGate> &gates, const PlannedBlock &b, const BlockLayout &lay)
    {
        const int kq = (int)lay.qubits.size();
        const int K = 1 << kq;
        std::vector<cd> M(size_t(K) * K, 0.0);
        for (int i = 0; i < K; ++i)
            M[size_t(i) * K + i] = 1.0;
        std::vector<cd> col(K), in(16), out(16);
        for (int gi : b.gates)
        {
            const ResolvedGate &g = gates[gi];
            const int t = g.spec.ntargets;
            const int T = 1 << t;
            auto U = dense_matrix(g.spec);
            int tpos[4] = {0, 0, 0, 0};
            for (int j = 0; j < t; ++j)
                tpos[j] = local_pos(lay.qubits, g.tphys[j]);
            unsigned need1 = 0, need0 = 0; // control bits inside the block
            for (int q : g.pphys)
            {
                int p = local_pos(lay.qubits, q);
                if (p >= 0)
                    need1 |= 1u << p;
            }
            for (int q : g.nphys)
            {
                int p = local_pos(lay.qubits, q);
                if (p >= 0)
                    need0 |= 1u << p;
            }
            unsigned tmask = 0;
            for (int j = 0; j < t; ++j)
                tmask |= 1u << tpos[j];
            // M <- G M, applying the controlled gate to every column of M
            for (int c = 0; c < K; ++c)
            {
                for (int r = 0; r < K; ++r)
                    col[r] = M[size_t(r) * K + c];
                for (int x = 0; x < K; ++x)
                {
                    if (x & tmask)
                        continue;
                    if ((x & need1) != need1 || (x & need0) != 0)
                        continue;
                    for (int r = 0; r < T; ++r)
                    {
                        int y = x;
                        for (int j = 0; j < t; ++j)
                            if ((r >> j) & 1)
                                y |= 1 << tpos[j];
                        in[r] = col[y];
                    }
                    for (int r = 0; r < T; ++r)
                    {
                        cd acc = 0.0;
                        for (int cc = 0; cc < T; ++cc)
                            acc += U[size_t(r) * T + cc] * in[cc];
                        out[r] = acc;
                    }
                    for (int r = 0; r < T; ++r)
                    {
                        int y = x;
                        for (int j = 0; j < t; ++j)
                            if ((r >> j) & 1)
                                y |= 1 << tpos[j];
                        col[y] = out[r];
                    }
                }
                for (int r = 0; r < K; ++r)
                    M[size_t(r) * K + c] = col[r];
            }
        }
        return M;
    }

    FusedEngine::FusedEngine(Simulator &sim, int k, int mode) : sim_(sim), k_(k), f64_(mode == 2), x3_(mode == 3)
    {
        const int n = sim_.n_;
        if (mode < 1 || mode > 3)
            throw std::runtime_error("fused engine: mode must be 1, 2 or 3");
        if (n < 4)
            throw std::runtime_error("precision mode needs at least 4 qubits");
        kq_ = (f64_ || k <= 4 || n < 5) ? 4 : 5;
        if (k_ > kq_)
            k_ = kq_;
        if (f64_)
            qsv::check(hipMalloc((void **)&state128_, (uint64_t(1) << n) * sizeof(qsv::c128)), "hipMalloc c128 state");
        else
            qsv::check(hipMalloc((void **)&state_, (uint64_t(1) << n) * sizeof(qsv::c64)), "hipMalloc c64 state");
        qsv::check(hipEventCreateWithFlags(&copy_done_, hipEventDisableTiming), "hipEventCreate");
    }

    FusedEngine::~FusedEngine()
    {
        (void)hipStreamSynchronize(sim_.stream_);
        if (state_)
            (void)hipFree(state_);
        if (state128_)
            (void)hipFree(state128_);
        if (rdev_)
            (void)hipFree(rdev_);
        if (rhost_)
            (void)hipHostFree(rhost_);
        if (copy_done_)
            (void)hipEventDestroy(copy_done_);
        for (hipEvent_t e : ev_pool_)
            (void)hipEventDestroy(e);
        clear_trials();
    }

    void FusedEngine::clear_trials()
    {
        for (auto &kv : trials_)
            for (hipEvent_t e : kv.second.ev)
                if (e)
                    (void)hipEventDestroy(e);
        trials_.clear();
    }

    void FusedEngine::set_model(CostModel *m)
    {
        // measured checks belong to the model that made the choices
        if (m != model_)
            clear_trials();
        model_ = m;
    }

    void FusedEngine::set_check(bool on)
    {
        check_ = on;
        clear_trials();
    }

    uint64_t FusedEngine::structure_key(const std::vector<ResolvedGate> &gates)
    {
        // FNV-1a over what decides a plan: gate names and kinds, targets and controls
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](uint64_t v) {
            for (int b = 0; b < 8; ++b)
            {
                h ^= (v >> (8 * b)) & 0xff;
                h *= 1099511628211ull;
            }
        };
        mix(gates.size());
        for (const auto &g : gates)
        {
            for (char ch : g.spec.name)
                mix(uint64_t(uint8_t(ch)));
            mix(uint64_t(g.spec.kind));
            for (const auto *v : {&g.tphys, &g.pphys, &g.nphys})
            {
                mix(v->size());
                for (int q : *v)
                    mix(uint64_t(q));
            }
        }
        return h;
    }

    void FusedEngine::reserve_matrices(size_t bytes)
    {
        if (copy_pending_)
        {
            qsv::check(hipEventSynchronize(copy_done_), "hipEventSynchronize");
            copy_pending_ = false;
        }
        if (bytes <= rcap_)
            return;
        // the device buffer may still be read by queued kernels: wait before freeing it
        qsv::check(hipStreamSynchronize(sim_.stream_), "hipStreamSynchronize");
        if (rdev_)
            (void)hipFree(rdev_);
        if (rhost_)
            (void)hipHostFree(rhost_);
        rdev_ = nullptr;
        rhost_ = nullptr;
        size_t cap = std::max<size_t>(bytes, size_t(1) << 20);
        qsv::check(hipMalloc(&rdev_, cap), "hipMalloc block matrices");
        qsv::check(hipHostMalloc(&rhost_, cap, hipHostMallocDefault), "hipHostMalloc block matrices");
        rcap_ = cap;
    }

    void FusedEngine::enqueue(const ResolvedGate &g)
    {
        if (g.mmask)
        {
            // Inputs on the unobserved branch of a measured target read as zero
            // (src/qcs.cu:1153-1171): zero that branch in memory first, then apply normally.
            flush();
            for (int j = 0; j < g.spec.ntargets; ++j)
                if ((g.mmask >> j) & 1)
                {
                    if (f64_)
                        qsv::launch_collapse(sim_.stream_, state128_, sim_.n_, g.tphys[j], (g.mval >> j) & 1, 1.0);
                    else
                        qsv::launch_collapse(sim_.stream_, state_, sim_.n_, g.tphys[j], (g.mval >> j) & 1, 1.0);
                }
            ResolvedGate h = g;
            h.mmask = h.mval = 0;
            queue_.push_back(h);
        }
        else
        {
            queue_.push_back(g);
        }
        if (queue_.size() >= 4096)
            flush();
    }

    void FusedEngine::flush()
    {
        if (queue_.empty())
            return;
        std::vector<ResolvedGate> gates;
        gates.swap(queue_);
        apply_plan(gates);
        renormalize();
    }

    void FusedEngine::renormalize()
    {
        // complex64 block products are unitary only to their arithmetic (BF16x3 holds each
        // matrix as two bfloat16 parts, about 2^-17 per entry), which adds up to a norm drift
        // of up to ~1e-5 over a few hundred blocks: each flush ends with an FP64 norm and
        // rescale, queued on the stream (no host round trip)
        if (!f64_)
            qsv::renormalize_async(sim_.stream_, state_, uint64_t(1) << sim_.n_, sim_.work_);
    }

    std::vector<PlanOption> FusedEngine::plan_options() const
    {
        // candidate 0: the plan the engine runs without a model (stages when enabled, else
        // greedy blocks)
        const int n = sim_.n_;
        const int m0 = tile_bits();
        // complex128 stages on 2^13 tiles (one 512-thread workgroup per CU) stay out of the
        // candidates: at 33 qubits they ran the depth-40 brick in 5.2-6.2 s against 4.7 s for
        // 2^12 tiles, and a model trained at up to 32 qubits took them for faster; they remain
        // available through a fixed plan (qcs_ext_set_plan)
        std::vector<int> tiles;
        for (int m : f64_ ? std::vector<int>{11, 12} : std::vector<int>{11, 12, 13, 14})
            if (n >= m + 4)
                tiles.push_back(m);
        std::vector<PlanOption> opts;
        opts.push_back(PlanOption{m0, m0 > 0 ? stage_k() : k_, 64, 0, kTilesAuto});
        for (const auto &o : candidate_options(kq_, tiles, stage_block_sizes()))
            if (o.k <= k_ || o.stage_bits > 0)
                opts.push_back(o);
        return opts;
    }

    void FusedEngine::option_features(const std::vector<ResolvedGate> &gates, const PlanOption &o, std::vector<float> &X, PlannedOption *keep) const
    {
        // one row per launch, in the order apply_option issues them
        const int n = sim_.n_;
        const float math = math_feature();
        auto add_block = [&](const std::vector<ResolvedGate> &gs, const PlannedBlock &b) {
            const BlockLayout lay = layout_block(gs, b, kq_, n);
            float f[kPlanFeatures];
            block_features(n, math, kq_, lay, (int)b.gates.size(), block_uses_lds(lay, n, f64_), f);
            X.insert(X.end(), f, f + kPlanFeatures);
        };
        if (o.stage_bits == 0)
        {
            auto blocks = plan_with(gates, o);
            for (const auto &b : blocks)
                add_block(gates, b);
            if (keep)
                keep->blocks = std::move(blocks);
            return;
        }
        auto stages = plan_stages(gates, n, o.stage_bits, o.k, o.tiles);
        for (const auto &st : stages)
        {
            std::vector<ResolvedGate> sub;
            for (int gi : st.gates)
                sub.push_back(gates[gi]);
            if (st.tile.empty())
            {
                for (const auto &b : plan_greedy(sub, k_))
                    add_block(sub, b);
                continue;
            }
            float f[kPlanFeatures];
            const int nb = f64_ ? stage128_lds_blocks(sub, o.k) : stage64_lds_blocks(sub, o.k);
            stage_features(n, math, o.k, st.tile, nb, (int)st.gates.size(), f);
            X.insert(X.end(), f, f + kPlanFeatures);
        }
        if (keep)
            keep->stages = std::move(stages);
    }

    void FusedEngine::apply_option(const std::vector<ResolvedGate> &gates, const PlanOption &o)
    {
        if (o.stage_bits > 0)
            apply_stages(gates, o.stage_bits, o.k, o.tiles);
        else
            apply_blocks(gates, plan_with(gates, o));
    }

    void FusedEngine::apply_plan(const std::vector<ResolvedGate> &gates)
    {
        const int m0 = tile_bits();
        if (!model_ && forced_)
        {
            if (forced_bits_ > 0)
                apply_stages(gates, forced_bits_, forced_k_, forced_tiles_);
            else
                apply_blocks(gates, plan_rule(gates, forced_k_, forced_window_, forced_rule_));
            return;
        }
        if (!model_)
        {
            if (m0 > 0)
                apply_stages(gates, m0, stage_k());
            else
                apply_blocks(gates, plan_greedy(gates, k_));
            return;
        }
        // Learned planner: predicted cost of each candidate plan = sum over its launches of
        // 2^prediction (the model predicts log2 microseconds).  Candidate 0 is the plan the
        // engine runs without a model, and it judges the others: it is kept unless a
        // candidate's predicted gain exceeds the model's own error (choice_factor).  The
        // candidates are planned on up to 16 host threads: a circuit that runs as one flush
        // waits for its planning with the GPU idle.
        const auto opts = plan_options();
        std::vector<std::vector<float>> rows(opts.size());
        std::vector<PlannedOption> planned(opts.size());
        {
            const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::vector<std::thread> pool;
            std::exception_ptr err;
            std::mutex err_mu;
            std::atomic<size_t> next{0}; // candidates to threads as they free up (window-tile plans cost most)
            for (unsigned t = 0; t < hw; ++t)
                pool.emplace_back([&] {
                    try
                    {
                        for (size_t oi = next++; oi < opts.size(); oi = next++)
                            option_features(gates, opts[oi], rows[oi], &planned[oi]);
                    }
                    catch (...)
                    {
                        std::lock_guard<std::mutex> lock(err_mu);
                        err = std::current_exception();
                    }
                });
            for (auto &th : pool)
                th.join();
            if (err)
                std::rethrow_exception(err);
        }
        std::vector<float> X;
        std::vector<size_t> first(opts.size() + 1, 0);
        for (size_t oi = 0; oi < opts.size(); ++oi)
        {
            first[oi] = X.size() / kPlanFeatures;
            X.insert(X.end(), rows[oi].begin(), rows[oi].end());
        }
        first[opts.size()] = X.size() / kPlanFeatures;
        std::vector<float> Y(first.back());
        model_->predict(X.data(), (int)Y.size(), Y.data());
        std::vector<double> cost(opts.size(), 0.0);
        for (size_t oi = 0; oi < opts.size(); ++oi)
            for (size_t r = first[oi]; r < first[oi + 1]; ++r)
                cost[oi] += std::exp2(double(Y[r]));
        size_t best = 0;
        for (size_t oi = 1; oi < opts.size(); ++oi)
            if (std::isfinite(cost[oi]) && cost[oi] < cost[best])
                best = oi;
        if (!(best != 0 && std::isfinite(cost[0]) && cost[best] < model_->choice_factor() * cost[0]))
            best = 0;
        // The default plan checks an override by measurement: the first flush of this
        // structure runs the model's choice timed, the second the default timed, and the
        // later ones the faster of the two.
        size_t run = best;
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (best != 0 && check_)
        {
            const uint64_t key = structure_key(gates);
            auto it = trials_.find(key);
            if (it == trials_.end() && trials_.size() < kMaxTrials)
            {
                it = trials_.emplace(key, PlanTrial{}).first;
                for (hipEvent_t &e : it->second.ev)
                    qsv::check(hipEventCreate(&e), "hipEventCreate");
            }
            if (it != trials_.end())
            {
                PlanTrial &v = it->second;
                if (v.stage == 0 || v.option != best)
                {
                    v.stage = 1;
                    v.option = best;
                    t0 = v.ev[0];
                    t1 = v.ev[1];
                }
                else if (v.stage == 1)
                {
                    v.stage = 2;
                    run = 0;
                    t0 = v.ev[2];
                    t1 = v.ev[3];
                }
                else
                {
                    if (v.stage == 2)
                    {
                        float ms_learned = 0.f, ms_default = 0.f;
                        qsv::check(hipEventSynchronize(v.ev[3]), "hipEventSynchronize");
                        qsv::check(hipEventElapsedTime(&ms_learned, v.ev[0], v.ev[1]), "hipEventElapsedTime");
                        qsv::check(hipEventElapsedTime(&ms_default, v.ev[2], v.ev[3]), "hipEventElapsedTime");
                        v.learned = ms_learned < ms_default;
                        v.stage = 3;
                    }
                    run = v.learned ? best : 0;
                }
            }
        }
        stat_flushes_ += 1;
        stat_pred_greedy_ += cost[0];
        stat_pred_chosen_ += cost[run];
        if (run != 0)
            stat_learned_ += 1;
        last_choice_ = opts[run];
        // the chosen plan as planned above
        if (t0)
            qsv::check(hipEventRecord(t0, sim_.stream_), "hipEventRecord");
        if (opts[run].stage_bits > 0)
            apply_planned_stages(gates, planned[run].stages, opts[run].stage_bits, opts[run].k);
        else
            apply_blocks(gates, planned[run].blocks);
        if (t1)
            qsv::check(hipEventRecord(t1, sim_.stream_), "hipEventRecord");
    }

    void FusedEngine::launch_mark_begin()
    {
        if (!launch_log_)
            return;
        if (ev_used_ + 2 > ev_pool_.size())
            for (int i = 0; i < 2; ++i)
            {
                hipEvent_t e;
                qsv::check(hipEventCreate(&e), "hipEventCreate");
                ev_pool_.push_back(e);
            }
        hipEvent_t a = ev_pool_[ev_used_++];
        qsv::check(hipEventRecord(a, sim_.stream_), "hipEventRecord");
        launch_log_->push_back({a, nullptr});
    }

    void FusedEngine::launch_mark_end()
    {
        if (!launch_log_)
            return;
        hipEvent_t b = ev_pool_[ev_used_++];
        qsv::check(hipEventRecord(b, sim_.stream_), "hipEventRecord");
        launch_log_->back().second = b;
    }

    int FusedEngine::record_plans(int capacity, float *features, float *log2_us)
    {
        if (capacity < 0 || (capacity > 0 && (!features || !log2_us)))
            throw std::runtime_error("planner record: bad output buffers");
        if (queue_.empty())
            return 0;
        std::vector<ResolvedGate> gates;
        gates.swap(queue_);
        int count = 0;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> log;
        for (const auto &o : plan_options())
        {
            // stage plans only (block plans are sampled by record()); the default's "fewer
            // stages" choice repeats the grown or the window plan
            if (o.stage_bits == 0 || o.tiles == kTilesAuto || count >= capacity)
                continue;
            std::vector<float> X;
            option_features(gates, o, X);
            const size_t nrows = X.size() / kPlanFeatures;
            log.clear();
            ev_used_ = 0;
            launch_log_ = &log;
            try
            {
                apply_option(gates, o);
            }
            catch (...)
            {
                launch_log_ = nullptr;
                throw;
            }
            launch_log_ = nullptr;
            qsv::check(hipStreamSynchronize(sim_.stream_), "sync");
            if (log.size() != nrows)
                throw std::runtime_error("planner record: launches and feature rows differ");
            for (size_t r = 0; r < nrows && count < capacity; ++r)
            {
                float ms = 0.f;
                qsv::check(hipEventElapsedTime(&ms, log[r].first, log[r].second), "hipEventElapsedTime");
                std::memcpy(features + size_t(count) * kPlanFeatures, &X[r * kPlanFeatures], kPlanFeatures * sizeof(float));
                log2_us[count] = float(std::log2(std::max(1e3 * double(ms), 0.01)));
                ++count;
            }
        }
        renormalize();
        return count;
    }

    void FusedEngine::set_plan(bool enable, int bits, int k, int tiles, int window, int rule)
    {
        flush();
        if (!enable)
        {
            forced_ = false;
            return;
        }
        if (bits != 0)
        {
            const bool ok_bits = bits >= 11 && bits <= (f64_ ? 13 : 14);
            if (!ok_bits || sim_.n_ < bits + 4)
                throw std::runtime_error("plan: stage tile bits must be 11..13 (complex128) or 11..14 (complex64), with at least tile_bits + 4 qubits");
            const auto ks = stage_block_sizes();
            if (std::find(ks.begin(), ks.end(), k) == ks.end())
                throw std::runtime_error("plan: this arithmetic's stage kernels take other block sizes");
            if (tiles != kTilesGrow && tiles != kTilesWindow && tiles != kTilesAuto)
                throw std::runtime_error("plan: tile choice must be 0 (grow), 1 (window) or 2 (fewer stages)");
        }
        else if (k < 1 || k > k_ || window < 1 || rule < 0 || rule > 2)
            throw std::runtime_error("plan: block plans take 1 <= k <= the block limit, window >= 1, rule 0..2");
        forced_ = true;
        forced_bits_ = bits;
        forced_k_ = k;
        forced_tiles_ = tiles;
        forced_window_ = window;
        forced_rule_ = rule;
    }

    void FusedEngine::planner_stats(double out[5]) const
    {
        out[0] = stat_flushes_;
        out[1] = stat_learned_;
        out[2] = stat_pred_greedy_;
        out[3] = stat_pred_chosen_;
        out[4] = double(blocks_applied_);
    }

    void FusedEngine::record(int count, uint64_t seed, float *features, float *log2_us)
    {
        flush();
        const int n = sim_.n_;
        const float math = math_feature();
        auto rnd = [&seed]() {
            uint64_t z = (seed += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        auto unif = [&]() { return double(rnd() >> 11) * 0x1.0p-53; };
        std::vector<int> tiles;
        for (int m : f64_ ? std::vector<int>{11, 12} : std::vector<int>{11, 12, 13, 14})
            if (n >= m + 4)
                tiles.push_back(m);
        // 1..24 random gates (U3, or CX/CZ with probability 0.3) on the qubit set qs
        // diagonal: controlled phases and rz only (stages fold such blocks into phase
        // tables, much cheaper than dense blocks)
        auto random_gates = [&](const std::vector<int> &qs, std::vector<ResolvedGate> &out, bool diagonal = false) {
            const int ng = 1 + int(rnd() % 24);
            for (int g = 0; g < ng; ++g)
            {
                ResolvedGate r;
                if (diagonal)
                {
                    double prm[1] = {6.283 * unif()};
                    const int a = qs[rnd() % qs.size()];
                    r.tphys = {a};
                    if (qs.size() >= 2 && unif() < 0.7)
                    {
                        int b = a;
                        while (b == a)
                            b = qs[rnd() % qs.size()];
                        r.spec = make_gate("p", prm);
                        r.pphys = {b};
                    }
                    else
                        r.spec = make_gate("rz", prm);
                }
                else if (qs.size() >= 2 && unif() < 0.3)
                {
                    double none[1] = {0.0};
                    r.spec = make_gate(unif() < 0.5 ? "z" : "x", none);
                    int a = qs[rnd() % qs.size()], b = a;
                    while (b == a)
                        b = qs[rnd() % qs.size()];
                    r.tphys = {a};
                    r.pphys = {b};
                }
                else
                {
                    double prm[3] = {6.283 * unif(), 6.283 * unif(), 6.283 * unif()};
                    r.spec = make_gate("u3", prm);
                    r.tphys = {qs[rnd() % qs.size()]};
                }
                out.push_back(r);
            }
        };
        auto random_subset = [&](const std::vector<int> &from, int size) {
            std::vector<int> qs;
            while ((int)qs.size() < size)
            {
                int q = from[rnd() % from.size()];
                if (std::find(qs.begin(), qs.end(), q) == qs.end())
                    qs.push_back(q);
            }
            return qs;
        };
        std::vector<int> all(n);
        std::iota(all.begin(), all.end(), 0);
        const int chunk = 64;
        std::vector<hipEvent_t> ev(2 * chunk);
        for (auto &e : ev)
            qsv::check(hipEventCreate(&e), "hipEventCreate");
        try
        {
            // start = -3: three untimed warm-up launches, then rows [0, count) in chunks
            for (int start = -3; start < count; start = start < 0 ? 0 : start + chunk)
            {
                const int mcount = start < 0 ? 3 : std::min(chunk, count - start);
                std::vector<float> mats;
                std::vector<double> mats64;
                struct Launch
                {
                    int tile_bits = 0;
                    PreparedBlock b;
                    PreparedBlock128 b128;
                    PreparedStage st;
                    PreparedStage128 st128;
                    float f[kPlanFeatures];
                };
                std::vector<Launch> ls(mcount);
                for (int i = 0; i < mcount; ++i)
                {
                    Launch &L = ls[i];
                    // half of the launches are stages when the state is large enough
                    if (!tiles.empty() && (rnd() & 1))
                    {
                        const int m = tiles[rnd() % tiles.size()];
                        const auto ks = stage_block_sizes();
                        const int kb = ks[rnd() % ks.size()];
                        std::vector<int> upper(all.begin() + 4, all.end());
                        auto extra = random_subset(upper, m - 4);
                        PlannedStage st;
                        st.tile = {0, 1, 2, 3};
                        st.tile.insert(st.tile.end(), extra.begin(), extra.end());
                        std::sort(st.tile.begin(), st.tile.end());
                        std::vector<ResolvedGate> gs;
                        const int nb = 1 + int(rnd() % 12);
                        // a quarter of the stages all diagonal (phase tables), a quarter
                        // ending in a run of diagonal blocks (a table at the store)
                        const int shape = int(rnd() % 4);
                        for (int b = 0; b < nb; ++b)
                            random_gates(random_subset(st.tile, 1 + int(rnd() % kb)), gs, shape == 0 || (shape == 1 && 2 * b >= nb));
                        st.gates.resize(gs.size());
                        std::iota(st.gates.begin(), st.gates.end(), 0);
                        L.tile_bits = m;
                        int nblocks;
                        if (f64_)
                        {
                            L.st128 = prepare_stage128(gs, st, n, mats64, kb);
                            nblocks = stage128_lds_blocks(gs, kb);
                        }
                        else
                        {
                            L.st = prepare_stage(gs, st, kb, m, n, mats, x3_, true);
                            attach_stage(L.st, state_, mats);
                            nblocks = stage64_lds_blocks(gs, kb);
                        }
                        stage_features(n, math, kb, st.tile, nblocks, (int)gs.size(), L.f);
                        continue;
                    }
                    std::vector<ResolvedGate> gs;
                    random_gates(random_subset(all, 1 + int(rnd() % uint64_t(k_))), gs);
                    PlannedBlock pb;
                    for (const auto &g : gs)
                        for (int q : gate_qubits(g))
                            if (local_pos(pb.qubits, q) < 0)
                                pb.qubits.push_back(q);
                    for (size_t g = 0; g < gs.size(); ++g)
                        pb.gates.push_back((int)g);
                    const BlockLayout lay = layout_block(gs, pb, kq_, n);
                    if (f64_)
                        L.b128 = prepare_blocks128(gs, {pb}, n, mats64)[0];
                    else
                        L.b = prepare_blocks(gs, {pb}, kq_, n, mats, x3_)[0];
                    block_features(n, math, kq_, lay, (int)gs.size(), block_uses_lds(lay, n, f64_), L.f);
                }
                const size_t bytes = f64_ ? mats64.size() * sizeof(double) : mats.size() * sizeof(float);
                reserve_matrices(bytes);
                qsv::check(hipStreamSynchronize(sim_.stream_), "sync");
                qsv::check(hipMemcpy(rdev_, f64_ ? (const void *)mats64.data() : (const void *)mats.data(), bytes, hipMemcpyHostToDevice), "matrix upload");
                const float *devf = static_cast<const float *>(rdev_);
                const double *devd = static_cast<const double *>(rdev_);
                for (int i = 0; i < mcount; ++i)
                {
                    Launch &L = ls[i];
                    qsv::check(hipEventRecord(ev[2 * i], sim_.stream_), "hipEventRecord");
                    if (L.tile_bits && f64_)
                        launch_prepared_stage128(sim_.stream_, state128_, n, L.st128, devd);
                    else if (L.tile_bits)
                        launch_prepared_stage(sim_.stream_, L.st, devf, stage_stream_);
                    else if (f64_)
                        launch_prepared128(sim_.stream_, state128_, n, L.b128, devd);
                    else
                        launch_prepared(sim_.stream_, state_, n, L.b, devf);
                    qsv::check(hipEventRecord(ev[2 * i + 1], sim_.stream_), "hipEventRecord");
                }
                qsv::check(hipStreamSynchronize(sim_.stream_), "sync");
                if (start < 0)
                    continue;
                for (int i = 0; i < mcount; ++i)
                {
                    float ms = 0.f;
                    qsv::check(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]), "hipEventElapsedTime");
                    const int row = start + i;
                    std::memcpy(features + size_t(row) * kPlanFeatures, ls[i].f, sizeof(ls[i].f));
                    log2_us[row] = float(std::log2(std::max(1e-3, double(ms) * 1e3)));
                }
            }
        }
        catch (...)
        {
            for (auto e : ev)
                (void)hipEventDestroy(e);
            throw;
        }
        for (auto e : ev)
            (void)hipEventDestroy(e);
    }

    namespace
    {
        // float -> bfloat16 bits, round to nearest even (NaN stays NaN)
        uint16_t bf16_bits(float f)
        {
            uint32_t u;
            std::memcpy(&u, &f, 4);
            if ((u & 0x7fffffffu) > 0x7f800000u)
                return uint16_t((u >> 16) | 0x40u);
            u += 0x7fffu + ((u >> 16) & 1u);
            return uint16_t(u >> 16);
        }
        float bf16_value(uint16_t b)
        {
            const uint32_t u = uint32_t(b) << 16;
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        }
    } // namespace

    // An FP32 block matrix in the layout the block kernels read: 5-qubit blocks as Re M,
    // Im M and Re M + Im M (each 32 x 32 row-major, the sum formed in FP64; three real
    // products per block on MFMA), smaller blocks as the real form R (2K x 2K on
    // interleaved (re, im) vectors).  dst holds (2K)^2 floats.
    static void write_fp32_matrix(const std::vector<cd> &M, const std::vector<double> &R, int kq, float *dst)
    {
        const int K = 1 << kq;
        if (kq == 5)
        {
            for (int r = 0; r < K; ++r)
                for (int c = 0; c < K; ++c)
                {
                    const cd z = M[size_t(r) * K + c];
                    dst[r * K + c] = float(z.real());
                    dst[K * K + r * K + c] = float(z.imag());
                    dst[2 * K * K + r * K + c] = float(z.real() + z.imag());
                }
            return;
        }
        for (size_t i = 0; i < R.size(); ++i)
            dst[i] = float(R[i]);
    }

    void split_bf16(const double *R, size_t count, uint16_t *hi, uint16_t *lo)
    {
        for (size_t i = 0; i < count; ++i)
        {
            hi[i] = bf16_bits(float(R[i]));
            lo[i] = bf16_bits(float(R[i] - double(bf16_value(hi[i]))));
        }
    }

    std::vector<PreparedBlock> prepare_blocks(const std::vector<ResolvedGate> &gates, const std::vector<PlannedBlock> &plan, int kq, int n, std::vector<float> &mats, bool x3)
    {
        const int D = 2 << kq;
        const size_t per = size_t(D) * D;
        const int K = 1 << kq;
        std::vector<PreparedBlock> out;
        out.reserve(plan.size());
        for (const auto &pb : plan)
        {
            const BlockLayout lay = layout_block(gates, pb, kq, n);
            const auto M = block_matrix(gates, pb, lay);
            PreparedBlock b;
            b.kq = kq;
            b.gates = (int)pb.gates.size();
            b.matrix_offset = mats.size();
            mats.resize(mats.size() + per);
            std::vector<double> R(per);
            for (int r = 0; r < K; ++r)
                for (int c = 0; c < K; ++c)
                {
                    const cd z = M[size_t(r) * K + c];
                    R[(2 * r) * D + 2 * c] = z.real();
                    R[(2 * r) * D + 2 * c + 1] = -z.imag();
                    R[(2 * r + 1) * D + 2 * c] = z.imag();
                    R[(2 * r + 1) * D + 2 * c + 1] = z.real();
                }
            if (x3)
            { // bfloat16 high parts then low parts, 2 * per halves = per floats of storage
                std::vector<uint16_t> parts(2 * per);
                split_bf16(R.data(), per, parts.data(), parts.data() + per);
                std::memcpy(mats.data() + b.matrix_offset, parts.data(), 2 * per * sizeof(uint16_t));
            }
            else
            {
                write_fp32_matrix(M, R, kq, mats.data() + b.matrix_offset);
            }
            std::vector<int> outside = lay.pos_out;
            outside.insert(outside.end(), lay.neg_out.begin(), lay.neg_out.end());
            uint64_t setb = 0;
            for (int q : lay.pos_out)
                setb |= uint64_t(1) << q;
            // the 5 lowest qubits outside the block and its controls become the tile's column
            // bits when there are that many (coalesced LDS-staged kernel)
            std::vector<int> cols;
            for (int q = 0; q < n && cols.size() < 5; ++q)
                if (local_pos(lay.qubits, q) < 0 && local_pos(outside, q) < 0)
                    cols.push_back(q);
            b.bytes = 2.0 * 8.0 * double(uint64_t(1) << (n - (int)outside.size()));
            b.sub = qsv::Subspace{};
            b.sub.set_bits = setb;
            std::vector<int> ops;
            if (cols.size() == 5)
            {
                b.lds = true;
                std::vector<int> tb = lay.qubits;
                tb.insert(tb.end(), cols.begin(), cols.end());
                std::sort(tb.begin(), tb.end());
                b.lop = qsv::BlockLdsOp{};
                b.lop.kq = kq;
                b.lop.x3 = x3 ? 1 : 0;
                const int row = K + 2;
                for (size_t i = 0; i < tb.size(); ++i)
                {
                    b.lop.gw[i] = uint64_t(1) << tb[i];
                    int j = local_pos(lay.qubits, tb[i]);
                    b.lop.lw[i] = j >= 0 ? (1 << j) : (row << local_pos(cols, tb[i]));
                }
                ops = tb;
            }
            else
            {
                b.lds = false;
                b.op = qsv::BlockOp{};
                b.op.kq = kq;
                b.op.x3 = x3 ? 1 : 0;
                for (int e = 0; e < K; ++e)
                {
                    uint64_t o = 0;
                    for (int j = 0; j < kq; ++j)
                        if ((e >> j) & 1)
                            o |= uint64_t(1) << lay.qubits[j];
                    b.op.off[e] = o;
                }
                ops = lay.qubits;
            }
            ops.insert(ops.end(), outside.begin(), outside.end());
            std::sort(ops.begin(), ops.end());
            b.sub.nops = (int)ops.size();
            for (int i = 0; i < b.sub.nops; ++i)
                b.sub.ops[i] = ops[i];
            out.push_back(b);
        }
        return out;
    }

    void launch_prepared(hipStream_t s, qsv::c64 *state, int n, PreparedBlock &b, const float *mats_dev)
    {
        if (b.lds)
        {
            b.lop.R = mats_dev + b.matrix_offset;
            qsv::launch_block_lds(s, state, n, b.sub, b.lop);
        }
        else
        {
            b.op.R = mats_dev + b.matrix_offset;
            qsv::launch_block(s, state, n, b.sub, b.op);
        }
    }

    std::vector<PreparedBlock128> prepare_blocks128(const std::vector<ResolvedGate> &gates, const std::vector<PlannedBlock> &plan, int n, std::vector<double> &mats)
    {
        std::vector<PreparedBlock128> out;
        out.reserve(plan.size());
        for (const auto &pb : plan)
        {
            const BlockLayout lay = layout_block(gates, pb, 4, n);
            const auto M = block_matrix(gates, pb, lay);
            PreparedBlock128 b;
            b.gates = (int)pb.gates.size();
            b.matrix_offset = mats.size();
            for (int i = 0; i < 256; ++i)
            {
                mats.push_back(M[i].real());
                mats.push_back(M[i].imag());
            }
            std::vector<int> outside = lay.pos_out;
            outside.insert(outside.end(), lay.neg_out.begin(), lay.neg_out.end());
            b.sub = qsv::Subspace{};
            for (int q : lay.pos_out)
                b.sub.set_bits |= uint64_t(1) << q;
            b.bytes = 2.0 * 16.0 * double(uint64_t(1) << (n - (int)outside.size()));
            std::vector<int> cols;
            for (int q = 0; q < n && cols.size() < 6; ++q)
                if (local_pos(lay.qubits, q) < 0 && local_pos(outside, q) < 0)
                    cols.push_back(q);
            for (int e = 0; e < 16; ++e)
            {
                uint64_t o = 0;
                for (int j = 0; j < 4; ++j)
                    if ((e >> j) & 1)
                        o |= uint64_t(1) << lay.qubits[j];
                b.op.off[e] = o;
            }
            std::vector<int> ops;
            b.staged = cols.size() == 6;
            if (b.staged)
            {
                std::vector<int> tb = lay.qubits;
                tb.insert(tb.end(), cols.begin(), cols.end());
                std::sort(tb.begin(), tb.end());
                for (size_t i = 0; i < tb.size(); ++i)
                {
                    b.op.gw[i] = uint64_t(1) << tb[i];
                    int j = local_pos(lay.qubits, tb[i]);
                    b.op.lw[i] = j >= 0 ? (1 << j) : (17 << local_pos(cols, tb[i]));
                }
                ops = tb;
            }
            else
            {
                ops = lay.qubits;
            }
            ops.insert(ops.end(), outside.begin(), outside.end());
            std::sort(ops.begin(), ops.end());
            b.sub.nops = (int)ops.size();
            for (int i = 0; i < b.sub.nops; ++i)
                b.sub.ops[i] = ops[i];
            out.push_back(b);
        }
        return out;
    }

    void launch_prepared128(hipStream_t s, qsv::c128 *state, int n, PreparedBlock128 &b, const double *mats_dev)
    {
        b.op.M = mats_dev + b.matrix_offset;
        qsv::launch_block128(s, state, n, b.sub, b.op, b.staged);
    }

    namespace
    {
        // One stage taken from `rem` with the tile starting as `start` (grown up to m qubits
        // when `grow`): the gates taken, the gates deferred, the tile set.
        struct StageTake
        {
            std::vector<int> taken, deferred;
            std::vector<char> in_tile;
        };
        // Gates are taken in order; a gate may pass a deferred gate only if they commute:
        // disjoint qubits, or both diagonal.  blocked[q]: 0 no deferred gate on q, 1 only
        // diagonal ones, 2 some other gate.
        // Once every qubit of a tile that cannot grow is blocked by a non-diagonal deferred
        // gate, no later gate can be taken (a gate needs all its qubits in the tile, none
        // blocked): the rest is deferred without looking at it.
        StageTake take_stage(const std::vector<std::vector<int>> &qubits, const std::vector<char> &diag, const std::vector<int> &rem, int n, int m, int k, const std::vector<int> &start, bool grow)
        {
            StageTake r;
            r.in_tile.assign(n, 0);
            std::vector<char> blocked(n, 0);
            int tile_size = 0;
            for (int q : start)
                if (!r.in_tile[q])
                    r.in_tile[q] = 1, ++tile_size;
            int closed = 0; // tile qubits blocked at level 2 (fixed tiles only)
            for (size_t ri = 0; ri < rem.size(); ++ri)
            {
                const int gi = rem[ri];
                if (!grow && closed == tile_size)
                {
                    r.deferred.insert(r.deferred.end(), rem.begin() + ri, rem.end());
                    break;
                }
                const auto &qs = qubits[gi];
                const char level = diag[gi] ? 1 : 2;
                bool take = (int)qs.size() <= k;
                int need = 0;
                for (int q : qs)
                {
                    if (blocked[q] > (diag[gi] ? 1 : 0))
                        take = false;
                    need += r.in_tile[q] ? 0 : 1;
                }
                if (take && need > 0 && (!grow || tile_size + need > m))
                    take = false;
                if (take)
                {
                    for (int q : qs)
                        if (!r.in_tile[q])
                            r.in_tile[q] = 1, ++tile_size;
                    r.taken.push_back(gi);
                }
                else
                {
                    r.deferred.push_back(gi);
                    for (int q : qs)
                    {
                        if (level == 2 && blocked[q] < 2 && r.in_tile[q])
                            ++closed;
                        blocked[q] = std::max(blocked[q], level);
                    }
                }
            }
            return r;
        }

        // The number of gates take_stage(..., start, false) would take, by the same rule,
        // without building its lists (the window search ranks every window by this count).
        int count_stage(const std::vector<std::vector<int>> &qubits, const std::vector<char> &diag, const std::vector<int> &rem, int n, int k,
                        const std::vector<int> &start)
        {
            char in_tile[qsv::kMaxQubits] = {}, blocked[qsv::kMaxQubits] = {};
            int size = 0;
            for (int q : start)
                if (!in_tile[q])
                    in_tile[q] = 1, ++size;
            int taken = 0, closed = 0;
            for (int gi : rem)
            {
                if (closed == size) // every tile qubit blocked: nothing more can be taken
                    break;
                const auto &qs = qubits[gi];
                const char level = diag[gi] ? 1 : 2;
                bool take = (int)qs.size() <= k;
                for (int q : qs)
                    if (blocked[q] > (diag[gi] ? 1 : 0) || !in_tile[q])
                        take = false;
                if (take)
                    ++taken;
                else
                    for (int q : qs)
                    {
                        if (level == 2 && blocked[q] < 2 && in_tile[q])
                            ++closed;
                        blocked[q] = std::max(blocked[q], level);
                    }
            }
            (void)n;
            return taken;
        }

        std::vector<PlannedStage> plan_stages_with(const std::vector<std::vector<int>> &qubits, const std::vector<char> &diag, int n, int m, int k, bool windows)
        {
            std::vector<PlannedStage> out;
            std::vector<int> rem(qubits.size());
            std::iota(rem.begin(), rem.end(), 0);
            const std::vector<int> low = {0, 1, 2, 3};
            while (!rem.empty())
            {
                StageTake best = take_stage(qubits, diag, rem, n, m, k, low, true);
                bool grown = true;
                if (windows)
                {
                    // rank the windows by count, then take the first best one
                    int best_count = (int)best.taken.size(), best_a = -1;
                    std::vector<int> w;
                    for (int a = 4; a + (m - 4) <= n; ++a)
                    {
                        w = low;
                        for (int q = a; q < a + m - 4; ++q)
                            w.push_back(q);
                        const int c = count_stage(qubits, diag, rem, n, k, w);
                        if (c > best_count)
                            best_count = c, best_a = a;
                    }
                    if (best_a >= 0)
                    {
                        w = low;
                        for (int q = best_a; q < best_a + m - 4; ++q)
                            w.push_back(q);
                        best = take_stage(qubits, diag, rem, n, m, k, w, false);
                        grown = false;
                    }
                }
                PlannedStage st;
                if (best.taken.empty())
                { // the first remaining gate is wider than k: it runs on its own
                    st.gates.push_back(rem.front());
                    rem.erase(rem.begin());
                    out.push_back(std::move(st));
                    continue;
                }
                st.gates = best.taken;
                if (grown)
                {
                    // pad the tile with the lowest free qubits (longer contiguous runs)
                    int tile_size = 0;
                    for (int q = 0; q < n; ++q)
                        tile_size += best.in_tile[q];
                    for (int q = 4; q < n && tile_size < m; ++q)
                        if (!best.in_tile[q])
                            best.in_tile[q] = 1, ++tile_size;
                }
                for (int q = 0; q < n; ++q)
                    if (best.in_tile[q])
                        st.tile.push_back(q);
                out.push_back(std::move(st));
                rem.swap(best.deferred);
            }
            return out;
        }
    } // namespace

    std::vector<PlannedStage> plan_stages(const std::vector<ResolvedGate> &gates, int n, int m, int k, int tiles)
    {
        if (m < 8 || m > n)
            throw std::runtime_error("stages: tile bits out of range");
        if (tiles != kTilesGrow && tiles != kTilesWindow && tiles != kTilesAuto)
            throw std::runtime_error("stages: unknown tile choice");
        std::vector<std::vector<int>> qubits;
        std::vector<char> diag;
        qubits.reserve(gates.size());
        for (const auto &g : gates)
        {
            qubits.push_back(gate_qubits(g));
            // diagonal in the computational basis (controls keep it so); a gate on a measured
            // target is kept in place
            const Kind kd = g.spec.kind;
            diag.push_back(g.mmask == 0 && (kd == Kind::GlobalPhase || kd == Kind::PhaseOne || kd == Kind::Diag1 || kd == Kind::Diag2 || kd == Kind::Identity));
        }
        if (tiles != kTilesAuto)
            return plan_stages_with(qubits, diag, n, m, k, tiles == kTilesWindow);
        auto grow = plan_stages_with(qubits, diag, n, m, k, false);
        auto window = plan_stages_with(qubits, diag, n, m, k, true);
        return window.size() < grow.size() ? window : grow;
    }

    PreparedStage prepare_stage(const std::vector<ResolvedGate> &gates, const PlannedStage &st, int kq, int m, int n, std::vector<float> &mats, bool x3, bool table)
    {
        if ((int)st.tile.size() != m)
            throw std::runtime_error("stage: tile size mismatch");
        std::vector<ResolvedGate> sub;
        sub.reserve(st.gates.size());
        for (int gi : st.gates)
            sub.push_back(gates[gi]);
        const auto plan = plan_greedy(sub, kq);
        const int D = 2 << kq;
        const size_t per = size_t(D) * D;
        const int K = 1 << kq;
        std::vector<qsv::StageBlock> desc;
        std::vector<std::vector<cd>> diagonals; // per block: its diagonal, empty if not diagonal
        for (const auto &pb : plan)
        {
            if (pb.wide)
                throw std::runtime_error("stage: wide gate inside a stage");
            BlockLayout lay;
            lay.qubits = pb.qubits;
            // pad with the highest tile qubits: low tile bits stay column bits
            for (int i = m - 1; i >= 0 && (int)lay.qubits.size() < kq; --i)
                if (local_pos(lay.qubits, st.tile[i]) < 0)
                    lay.qubits.push_back(st.tile[i]);
            const auto M = block_matrix(sub, pb, lay);
            {
                // products of diagonal gates have exactly zero off-diagonal entries
                bool diagonal = table;
                for (int r = 0; r < K && diagonal; ++r)
                    for (int c = 0; c < K && diagonal; ++c)
                        diagonal = r == c || M[size_t(r) * K + c] == cd(0.0, 0.0);
                std::vector<cd> d;
                for (int e = 0; diagonal && e < K; ++e)
                    d.push_back(M[size_t(e) * K + e]);
                diagonals.push_back(std::move(d));
            }
            std::vector<double> R(per);
            for (int r = 0; r < K; ++r)
                for (int c = 0; c < K; ++c)
                {
                    const cd z = M[size_t(r) * K + c];
                    R[(2 * r) * D + 2 * c] = z.real();
                    R[(2 * r) * D + 2 * c + 1] = -z.imag();
                    R[(2 * r + 1) * D + 2 * c] = z.imag();
                    R[(2 * r + 1) * D + 2 * c + 1] = z.real();
                }
            qsv::StageBlock sb{};
            sb.matrix = uint32_t(mats.size());
            mats.resize(mats.size() + per);
            if (x3)
            {
                std::vector<uint16_t> parts(2 * per);
                split_bf16(R.data(), per, parts.data(), parts.data() + per);
                std::memcpy(mats.data() + sb.matrix, parts.data(), 2 * per * sizeof(uint16_t));
            }
            else
            {
                write_fp32_matrix(M, R, kq, mats.data() + sb.matrix);
            }
            std::vector<char> is_block(m, 0);
            for (int j = 0; j < kq; ++j)
            {
                const int b = local_pos(st.tile, lay.qubits[j]);
                if (b < 0)
                    throw std::runtime_error("stage: block qubit outside the tile");
                sb.w[j] = 1 << b;
                is_block[b] = 1;
            }
            int j = kq;
            for (int b = 0; b < m; ++b)
                if (!is_block[b])
                    sb.w[j++] = 1 << b;
            for (int b = 0; b < m; ++b)
                sb.lds[b] = qsv::stage_lds_offset(uint32_t(sb.w[b]));
            desc.push_back(sb);
        }
        static_assert(sizeof(qsv::StageBlock) % 16 == 0, "StageBlock keeps 16-B alignment");
        PreparedStage p;
        p.gates = (int)st.gates.size();
        p.n = n;
        p.kq = kq;
        p.m = m;
        p.x3 = x3;
        p.bytes = 2.0 * 8.0 * double(uint64_t(1) << n);
        p.tsub.set_bits = 0;
        p.tsub.nops = m;
        for (int b = 0; b < m; ++b)
            p.tsub.ops[b] = st.tile[b];
        // With table: the diagonal blocks after the last other block become one table over
        // the tile indices (tile index bit b = tile qubit b = the blocks' weight 1 << b),
        // their diagonals multiplied in block order (FP64, rounded to FP32).  A stage of
        // diagonal blocks only runs as that table alone; otherwise the stage kernel applies
        // it as the tile is stored, and the blocks before it run as usual.
        size_t first_trail = desc.size();
        while (first_trail > 0 && !diagonals[first_trail - 1].empty())
            --first_trail;
        if (table && first_trail < desc.size())
        {
            while (mats.size() % 4)
                mats.push_back(0.f);
            (first_trail == 0 ? p.table_offset : p.trail_offset) = mats.size();
            for (int idx = 0; idx < (1 << m); ++idx)
            {
                cd v(1.0, 0.0);
                for (size_t bi = first_trail; bi < desc.size(); ++bi)
                {
                    int e = 0;
                    for (int jj = 0; jj < kq; ++jj)
                        if (idx & desc[bi].w[jj])
                            e |= 1 << jj;
                    v *= diagonals[bi][e];
                }
                mats.push_back(float(v.real()));
                mats.push_back(float(v.imag()));
            }
            if (first_trail > 0)
                desc.resize(first_trail);
        }
        p.nblocks = (int)desc.size();
        while (mats.size() % 4)
            mats.push_back(0.f);
        p.desc_offset = mats.size();
        mats.resize(mats.size() + desc.size() * sizeof(qsv::StageBlock) / sizeof(float));
        std::memcpy(mats.data() + p.desc_offset, desc.data(), desc.size() * sizeof(qsv::StageBlock));
        return p;
    }

    int stage64_lds_blocks(const std::vector<ResolvedGate> &gates, int kq) { return stage128_lds_blocks(gates, kq); }

    qsv::StageTask stage_task(const PreparedStage &p, qsv::c64 *state, uint64_t tile_begin)
    {
        if (p.desc_offset > 0xffffffffu)
            throw std::runtime_error("stage: matrix buffer beyond 2^32 floats");
        qsv::StageTask t{};
        t.state = state;
        t.tile_begin = tile_begin;
        for (int b = 0; b < p.m; ++b)
            t.gw[b] = uint64_t(1) << p.tsub.ops[b];
        t.tsub = p.tsub;
        t.blocks = uint32_t(p.desc_offset);
        t.nblocks = p.nblocks;
        if (p.trail_offset != SIZE_MAX)
        {
            if (p.trail_offset > 0xffffffffu)
                throw std::runtime_error("stage: matrix buffer beyond 2^32 floats");
            t.flags |= qsv::kStageTrailTable;
            t.trail = uint32_t(p.trail_offset);
        }
        qsv::check_stage_task(t, p.n, p.kq, p.m);
        return t;
    }

    size_t append_stage_tasks(std::vector<float> &mats, const std::vector<qsv::StageTask> &tasks)
    {
        static_assert(sizeof(qsv::StageTask) % 16 == 0, "StageTask keeps 16-B alignment");
        if (mats.empty())
            mats.resize(4, 0.f);
        while (mats.size() % 4)
            mats.push_back(0.f);
        const size_t off = mats.size();
        mats.resize(off + tasks.size() * sizeof(qsv::StageTask) / sizeof(float));
        std::memcpy(mats.data() + off, tasks.data(), tasks.size() * sizeof(qsv::StageTask));
        return off;
    }

    void attach_stage(PreparedStage &p, qsv::c64 *state, std::vector<float> &mats)
    {
        p.task_offset = append_stage_tasks(mats, {stage_task(p, state, 0)});
        p.state = state;
    }

    void launch_prepared_stage(hipStream_t s, const PreparedStage &p, const float *mats_dev, bool stream)
    {
        if (p.task_offset == 0)
            throw std::runtime_error("stage: launched without a task");
        if (p.table_offset != SIZE_MAX)
            qsv::launch_stage_table64(s, p.state, p.n, p.tsub, mats_dev + p.table_offset);
        else if (stream && qsv::stage_stream_supported(p.kq, p.m, p.n))
            qsv::launch_stage_stream(s, p.kq, p.m, p.x3, mats_dev, reinterpret_cast<const qsv::StageTask *>(mats_dev + p.task_offset), uint64_t(1) << (p.n - p.m));
        else
            launch_stage_tasks(s, p.kq, p.m, p.x3, mats_dev, p.task_offset, 1, uint64_t(1) << (p.n - p.m));
    }

    void launch_stage_tasks(hipStream_t s, int kq, int m, bool x3, const float *mats_dev, size_t task_offset, int ntasks, uint64_t total_tiles)
    {
        if (task_offset % 4)
            throw std::runtime_error("stage: misaligned task array");
        qsv::launch_stage(s, kq, m, x3, mats_dev, reinterpret_cast<const qsv::StageTask *>(mats_dev + task_offset), ntasks, total_tiles);
    }

    PreparedStage128 prepare_stage128(const std::vector<ResolvedGate> &gates, const PlannedStage &st, int n, std::vector<double> &mats, int kq)
    {
        const int m = (int)st.tile.size();
        if (kq != 3 && kq != 4)
            throw std::runtime_error("stage128: block qubits must be 3 or 4");
        const int K = 1 << kq;
        if (m < 11 || m > 13)
            throw std::runtime_error("stage128: tile bits must be 11, 12 or 13");
        std::vector<ResolvedGate> sub;
        sub.reserve(st.gates.size());
        for (int gi : st.gates)
            sub.push_back(gates[gi]);
        const auto plan = plan_greedy(sub, kq);
        std::vector<qsv::StageBlock> desc;
        std::vector<std::vector<std::complex<double>>> diagonals; // of the diagonal blocks
        for (const auto &pb : plan)
        {
            if (pb.wide)
                throw std::runtime_error("stage128: wide gate inside a stage");
            BlockLayout lay;
            lay.qubits = pb.qubits;
            for (int i = m - 1; i >= 0 && (int)lay.qubits.size() < kq; --i)
                if (local_pos(lay.qubits, st.tile[i]) < 0)
                    lay.qubits.push_back(st.tile[i]);
            const auto M = block_matrix(sub, pb, lay);
            qsv::StageBlock sb{};
            sb.matrix = uint32_t(mats.size());
            for (int i = 0; i < K * K; ++i)
            {
                mats.push_back(M[i].real());
                mats.push_back(M[i].imag());
            }
            std::vector<char> is_block(m, 0);
            for (int j = 0; j < kq; ++j)
            {
                const int b = local_pos(st.tile, lay.qubits[j]);
                if (b < 0)
                    throw std::runtime_error("stage128: block qubit outside the tile");
                sb.w[j] = 1 << b;
                is_block[b] = 1;
            }
            int j = kq;
            for (int b = 0; b < m; ++b)
                if (!is_block[b])
                    sb.w[j++] = 1 << b;
            // products of diagonal gates have exactly zero off-diagonal entries
            bool diagonal = true;
            for (int r = 0; r < K && diagonal; ++r)
                for (int c = 0; c < K && diagonal; ++c)
                    diagonal = r == c || M[r * K + c] == std::complex<double>(0.0, 0.0);
            if (diagonal)
            {
                sb.pad[0] = qsv::kStageBlockDiagonal;
                std::vector<std::complex<double>> d(K);
                for (int e = 0; e < K; ++e)
                    d[e] = M[e * K + e];
                diagonals.push_back(std::move(d));
            }
            desc.push_back(sb);
        }
        // Diagonal blocks after the last other block (all of them in an all-diagonal stage)
        // become one table over the tile indices: their diagonals multiplied in block order
        // (tile index bit b = tile qubit b = LDS weight 1 << b).  An all-diagonal stage runs
        // as that table alone; otherwise the table is applied as the tile is stored.
        size_t first_trail = desc.size();
        while (first_trail > 0 && (desc[first_trail - 1].pad[0] & qsv::kStageBlockDiagonal))
            --first_trail;
        const size_t ntrail = desc.size() - first_trail;
        PreparedStage128 p;
        p.gates = (int)st.gates.size();
        p.bytes = 2.0 * 16.0 * double(uint64_t(1) << n);
        size_t table = SIZE_MAX;
        if (ntrail > 0)
        {
            while (mats.size() % 2)
                mats.push_back(0.0);
            table = mats.size();
            const size_t d0 = diagonals.size() - ntrail; // the trailing blocks' diagonals
            for (int idx = 0; idx < (1 << m); ++idx)
            {
                std::complex<double> v(1.0, 0.0);
                for (size_t bi = first_trail; bi < desc.size(); ++bi)
                {
                    int e = 0;
                    for (int jj = 0; jj < kq; ++jj)
                        if (idx & desc[bi].w[jj])
                            e |= 1 << jj;
                    v *= diagonals[d0 + (bi - first_trail)][e];
                }
                mats.push_back(v.real());
                mats.push_back(v.imag());
            }
            desc.resize(first_trail);
        }
        if (desc.empty())
            p.table_offset = table;
        else
            p.trail_offset = table;
        p.nblocks = (int)desc.size();
        while (mats.size() % 2)
            mats.push_back(0.0);
        p.desc_offset = mats.size();
        mats.resize(mats.size() + desc.size() * sizeof(qsv::StageBlock) / sizeof(double));
        if (!desc.empty())
            std::memcpy(mats.data() + p.desc_offset, desc.data(), desc.size() * sizeof(qsv::StageBlock));
        p.tsub.set_bits = 0;
        p.tsub.nops = m;
        p.op.m = m;
        p.op.kq = kq;
        p.op.nblocks = p.nblocks;
        for (int b = 0; b < m; ++b)
        {
            p.tsub.ops[b] = st.tile[b];
            p.op.gw[b] = uint64_t(1) << st.tile[b];
        }
        return p;
    }

    int stage128_lds_blocks(const std::vector<ResolvedGate> &gates, int kq)
    {
        const auto plan = plan_greedy(gates, kq);
        int last = 0;
        for (int i = 0; i < (int)plan.size(); ++i)
            for (int gi : plan[i].gates)
            {
                const Kind kd = gates[gi].spec.kind;
                if (gates[gi].mmask != 0 || !(kd == Kind::GlobalPhase || kd == Kind::PhaseOne || kd == Kind::Diag1 || kd == Kind::Diag2 || kd == Kind::Identity))
                    last = i + 1;
            }
        return last;
    }

    void launch_prepared_stage128(hipStream_t s, qsv::c128 *state, int n, PreparedStage128 &p, const double *mats_dev)
    {
        p.op.M = mats_dev;
        p.op.blocks = reinterpret_cast<const qsv::StageBlock *>(mats_dev + p.desc_offset);
        p.op.table = p.table_offset == SIZE_MAX ? nullptr : mats_dev + p.table_offset;
        p.op.trail = p.trail_offset == SIZE_MAX ? nullptr : mats_dev + p.trail_offset;
        qsv::launch_stage128(s, state, n, p.tsub, p.op);
    }

    void FusedEngine::set_stage_stream(bool on)
    {
        flush();
        if (on != stage_stream_)
            clear_trials(); // measured with the other kernel
        stage_stream_ = on;
    }

    void FusedEngine::set_stage_bits(int m)
    {
        if (m != 0 && (m < 11 || m > 14))
            throw std::runtime_error("stage tile bits must be 0 (per-block launches), 11, 12, 13 or 14");
        flush();
        if (m != stage_bits_)
            clear_trials(); // the candidate plans (and the default) change
        stage_bits_ = m;
    }

    void FusedEngine::apply_stages(const std::vector<ResolvedGate> &gates, int m, int k, int tiles)
    {
        apply_planned_stages(gates, plan_stages(gates, sim_.n_, m, k, tiles), m, k);
    }

    void FusedEngine::apply_planned_stages(const std::vector<ResolvedGate> &gates, const std::vector<PlannedStage> &stages, int m, int k)
    {
        const int n = sim_.n_;
        // one matrix buffer for the whole flush: FP32 (complex64 modes) or FP64 entries
        std::vector<float> mats;
        std::vector<double> mats64;
        struct Item
        {
            bool stage;
            PreparedStage st;
            PreparedStage128 st128;
            std::vector<PreparedBlock> blocks;
            std::vector<PreparedBlock128> blocks128;
        };
        std::vector<Item> items;
        for (const auto &s : stages)
        {
            Item it;
            it.stage = !s.tile.empty();
            if (it.stage)
            {
                if (f64_)
                    it.st128 = prepare_stage128(gates, s, n, mats64, k);
                else
                {
                    it.st = prepare_stage(gates, s, k, m, n, mats, x3_, true);
                    attach_stage(it.st, state_, mats);
                }
            }
            else
            {
                std::vector<ResolvedGate> one{gates[s.gates[0]]};
                const auto plan = plan_greedy(one, k_);
                if (f64_)
                {
                    it.blocks128 = prepare_blocks128(one, plan, n, mats64);
                    for (auto &b : it.blocks128)
                        b.gates = 1;
                }
                else
                    it.blocks = prepare_blocks(one, plan, kq_, n, mats, x3_);
            }
            items.push_back(std::move(it));
        }
        const size_t bytes = f64_ ? mats64.size() * sizeof(double) : mats.size() * sizeof(float);
        reserve_matrices(bytes);
        std::memcpy(rhost_, f64_ ? (const void *)mats64.data() : (const void *)mats.data(), bytes);
        qsv::check(hipMemcpyAsync(rdev_, rhost_, bytes, hipMemcpyHostToDevice, sim_.stream_), "matrix upload");
        qsv::check(hipEventRecord(copy_done_, sim_.stream_), "hipEventRecord");
        copy_pending_ = true;
        for (auto &it : items)
        {
            if (it.stage)
            {
                sim_.prof_begin(QCS_EXT_K_FUSED, f64_ ? it.st128.bytes : it.st.bytes);
                launch_mark_begin();
                if (f64_)
                    launch_prepared_stage128(sim_.stream_, state128_, n, it.st128, static_cast<const double *>(rdev_));
                else
                    launch_prepared_stage(sim_.stream_, it.st, static_cast<const float *>(rdev_), stage_stream_);
                launch_mark_end();
                sim_.prof_end();
                blocks_applied_ += 1;
                gates_applied_ += f64_ ? it.st128.gates : it.st.gates;
                continue;
            }
            for (auto &b : it.blocks128)
            {
                sim_.prof_begin(QCS_EXT_K_FUSED, b.bytes);
                launch_mark_begin();
                launch_prepared128(sim_.stream_, state128_, n, b, static_cast<const double *>(rdev_));
                launch_mark_end();
                sim_.prof_end();
                blocks_applied_ += 1;
                gates_applied_ += b.gates;
            }
            for (auto &b : it.blocks)
            {
                sim_.prof_begin(QCS_EXT_K_FUSED, b.bytes);
                launch_mark_begin();
                launch_prepared(sim_.stream_, state_, n, b, static_cast<const float *>(rdev_));
                launch_mark_end();
                sim_.prof_end();
                blocks_applied_ += 1;
                gates_applied_ += b.gates;
            }
        }
    }

    void FusedEngine::apply_blocks(const std::vector<ResolvedGate> &gates, const std::vector<PlannedBlock> &plan)
    {
        if (plan.empty())
            return;
        const int n = sim_.n_;
        if (f64_)
        {
            std::vector<double> mats;
            auto blocks = prepare_blocks128(gates, plan, n, mats);
            const size_t bytes = mats.size() * sizeof(double);
            reserve_matrices(bytes);
            std::memcpy(rhost_, mats.data(), bytes);
            qsv::check(hipMemcpyAsync(rdev_, rhost_, bytes, hipMemcpyHostToDevice, sim_.stream_), "matrix upload");
            qsv::check(hipEventRecord(copy_done_, sim_.stream_), "hipEventRecord");
            copy_pending_ = true;
            for (auto &b : blocks)
            {
                sim_.prof_begin(QCS_EXT_K_FUSED, b.bytes);
                launch_mark_begin();
                launch_prepared128(sim_.stream_, state128_, n, b, static_cast<const double *>(rdev_));
                launch_mark_end();
                sim_.prof_end();
                blocks_applied_ += 1;
                gates_applied_ += b.gates;
            }
            return;
        }
        std::vector<float> mats;
        auto blocks = prepare_blocks(gates, plan, kq_, n, mats, x3_);
        const size_t bytes = mats.size() * sizeof(float);
        reserve_matrices(bytes);
        std::memcpy(rhost_, mats.data(), bytes);
        qsv::check(hipMemcpyAsync(rdev_, rhost_, bytes, hipMemcpyHostToDevice, sim_.stream_), "matrix upload");
        qsv::check(hipEventRecord(copy_done_, sim_.stream_), "hipEventRecord");
        copy_pending_ = true;
        for (auto &b : blocks)
        {
            sim_.prof_begin(QCS_EXT_K_FUSED, b.bytes);
            launch_mark_begin();
            launch_prepared(sim_.stream_, state_, n, b, static_cast<const float *>(rdev_));
            launch_mark_end();
            sim_.prof_end();
            blocks_applied_ += 1;
            gates_applied_ += b.gates;
        }
    }
} // namespace qcore
