// Batching form host side: partition, plan, capture into one hipGraph, autotune.
//
// Circuits of up to `lds_max_qubits` qubits go to the persistent LDS-resident kernel (one
// workgroup per circuit, work-counter driven); larger ones keep a complex64 state in HBM
// and run as fused k-qubit blocks.  Everything -- counter reset, the persistent kernel,
// the block sequences spread over several streams, the <Z> reductions -- is captured once
// into a hipGraph and replayed with a single launch.
#include "batch.hpp"
#include "simulator.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>

namespace qcore
{
    namespace
    {
        using cd = std::complex<double>;

        void mat2(const BatchGate &g, cd m[4])
        {
            GateSpec s = make_gate(g.name, g.params);
            if (s.ntargets != 1)
                throw std::runtime_error("batch: only 1-target gates (with at most one control) are batched: " + g.name);
            auto u = dense_matrix(s);
            for (int i = 0; i < 4; ++i)
                m[i] = u[i];
        }

        qsv::BatchOp to_op(int target, int control, const cd m[4])
        {
            qsv::BatchOp op{};
            op.target = target;
            op.control = control;
            for (int i = 0; i < 4; ++i)
            {
                op.m[2 * i] = float(m[i].real());
                op.m[2 * i + 1] = float(m[i].imag());
            }
            return op;
        }
    } // namespace

    Batch::Batch()
    {
        qsv::check(hipGetDevice(&device_), "hipGetDevice");
        qsv::check(hipEventCreate(&t0_), "hipEventCreate");
        qsv::check(hipEventCreate(&t1_), "hipEventCreate");
    }

    Batch::~Batch()
    {
        try
        {
            release();
        }
        catch (...)
        {
        }
        (void)hipEventDestroy(t0_);
        (void)hipEventDestroy(t1_);
    }

    void Batch::release()
    {
        for (auto s : streams_)
            (void)hipStreamSynchronize(s);
        if (exec_)
            (void)hipGraphExecDestroy(exec_);
        if (graph_)
            (void)hipGraphDestroy(graph_);
        exec_ = nullptr;
        graph_ = nullptr;
        work_ = nullptr;
        for (void *p : {(void *)d_circs_, (void *)d_ops_, (void *)d_out_, (void *)d_counter_, (void *)d_states_, (void *)d_partials_, (void *)d_mats_})
            if (p)
                (void)hipFree(p);
        d_circs_ = nullptr;
        d_ops_ = nullptr;
        d_out_ = nullptr;
        d_counter_ = nullptr;
        d_states_ = nullptr;
        d_partials_ = nullptr;
        d_mats_ = nullptr;
        for (auto e : fork_join_)
            (void)hipEventDestroy(e);
        fork_join_.clear();
        for (auto s : streams_)
            (void)hipStreamDestroy(s);
        streams_.clear();
        prepared_ = false;
    }

    int Batch::add_circuit(int n, const std::vector<BatchGate> &gates)
    {
        if (n < 1 || n > 30)
            throw std::runtime_error("batch: circuits must have 1..30 qubits");
        for (const auto &g : gates)
        {
            if (g.target < 0 || g.target >= n || g.control >= n || g.control == g.target)
                throw std::runtime_error("batch: qubit index out of range or control equal to target");
            cd m[4];
            mat2(g, m); // validates the gate name and arity
        }
        release();
        circuits_.push_back(Circuit{n, gates, total_qubits_});
        total_qubits_ += n;
        return (int)circuits_.size() - 1;
    }

    void Batch::set_tuning(const BatchTuning &t)
    {
        if (t.lds_max_qubits < 0 || t.lds_max_qubits > qsv::kBatchMaxLdsQubits)
            throw std::runtime_error("batch: lds_max_qubits must be in [0, 14]");
        if (t.block_size != 256 && t.block_size != 512 && t.block_size != 1024)
            throw std::runtime_error("batch: block size must be 256, 512 or 1024");
        if (t.block_qubits != 4 && t.block_qubits != 5)
            throw std::runtime_error("batch: block qubits must be 4 or 5");
        if (t.streams < 1 || t.streams > 4)
            throw std::runtime_error("batch: streams must be in [1, 4]");
        if (t.stage_bits != 0 && (t.stage_bits < 11 || t.stage_bits > 13))
            throw std::runtime_error("batch: stage bits must be 0, 11, 12 or 13");
        release();
        tuning_ = t;
    }

    void Batch::prepare()
    {
        if (prepared_)
            return;
        if (circuits_.empty())
            throw std::runtime_error("batch: no circuits");
        const BatchTuning &T = tuning_;
        // ---- small circuits: op lists for the persistent kernel ----
        std::vector<qsv::BatchCircuit> small;
        std::vector<qsv::BatchOp> ops;
        int small_max_n = 1;
        std::vector<int> large;
        for (int ci = 0; ci < (int)circuits_.size(); ++ci)
        {
            const Circuit &C = circuits_[ci];
            if (C.n > T.lds_max_qubits)
            {
                large.push_back(ci);
                continue;
            }
            qsv::BatchCircuit bc{C.n, 0, (int)ops.size(), C.out_offset};
            // adjacent uncontrolled gates on the same qubit are multiplied together (FP64)
            int last_t = -1;
            cd acc[4];
            auto emit = [&]() {
                if (last_t >= 0)
                    ops.push_back(to_op(last_t, -1, acc));
                last_t = -1;
            };
            for (const auto &g : C.gates)
            {
                cd m[4];
                mat2(g, m);
                if (g.control < 0)
                {
                    if (last_t == g.target)
                    {
                        cd r[4] = {m[0] * acc[0] + m[1] * acc[2], m[0] * acc[1] + m[1] * acc[3],
                                   m[2] * acc[0] + m[3] * acc[2], m[2] * acc[1] + m[3] * acc[3]};
                        for (int i = 0; i < 4; ++i)
                            acc[i] = r[i];
                    }
                    else
                    {
                        emit();
                        last_t = g.target;
                        for (int i = 0; i < 4; ++i)
                            acc[i] = m[i];
                    }
                }
                else
                {
                    emit();
                    ops.push_back(to_op(g.target, g.control, m));
                }
            }
            emit();
            bc.nops = (int)ops.size() - bc.op_offset;
            small.push_back(bc);
            small_max_n = std::max(small_max_n, C.n);
        }
        // ---- large circuits: fused blocks on HBM states ----
        struct LargeItem
        {
            bool stage;
            PreparedStage st;
            PreparedBlock b;
        };
        struct LargeRun
        {
            int ci;
            uint64_t state_offset;
            size_t partial_offset;
            std::vector<LargeItem> items;
        };
        std::vector<LargeRun> runs;
        std::vector<float> mats;
        uint64_t state_total = 0;
        size_t partial_total = 0;
        for (int ci : large)
        {
            const Circuit &C = circuits_[ci];
            std::vector<ResolvedGate> rg;
            rg.reserve(C.gates.size());
            for (const auto &g : C.gates)
            {
                ResolvedGate r;
                r.spec = make_gate(g.name, g.params);
                r.tphys = {g.target};
                if (g.control >= 0)
                    r.pphys = {g.control};
                rg.push_back(r);
            }
            const int kq = (C.n >= 5 && T.block_qubits == 5) ? 5 : 4;
            if (C.n < kq)
                throw std::runtime_error("batch: circuit too small for the HBM path");
            LargeRun lr{ci, state_total, partial_total, {}};
            if (T.stage_bits > 0 && C.n >= T.stage_bits + 4)
            {
                for (const auto &st : plan_stages(rg, C.n, T.stage_bits, kq))
                {
                    if (!st.tile.empty())
                    {
                        lr.items.push_back(LargeItem{true, prepare_stage(rg, st, kq, T.stage_bits, C.n, mats, false), {}});
                        continue;
                    }
                    std::vector<ResolvedGate> one{rg[st.gates[0]]};
                    for (auto &b : prepare_blocks(one, plan_greedy(one, kq), kq, C.n, mats))
                        lr.items.push_back(LargeItem{false, {}, b});
                }
            }
            else
            {
                for (auto &b : prepare_blocks(rg, plan_greedy(rg, kq), kq, C.n, mats))
                    lr.items.push_back(LargeItem{false, {}, b});
            }
            state_total += uint64_t(1) << C.n;
            partial_total += size_t(qsv::zsum_blocks(C.n)) * 32;
            runs.push_back(std::move(lr));
        }
        // ---- device buffers ----
        auto upload = [](void **dst, const void *src, size_t bytes) {
            if (bytes == 0)
                return;
            qsv::check(hipMalloc(dst, bytes), "hipMalloc batch");
            qsv::check(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), "batch upload");
        };
        upload((void **)&d_circs_, small.data(), small.size() * sizeof(qsv::BatchCircuit));
        upload((void **)&d_ops_, ops.data(), ops.size() * sizeof(qsv::BatchOp));
        upload((void **)&d_mats_, mats.data(), mats.size() * sizeof(float));
        qsv::check(hipMalloc((void **)&d_out_, std::max(1, total_qubits_) * sizeof(double)), "hipMalloc batch");
        qsv::check(hipMalloc((void **)&d_counter_, sizeof(unsigned)), "hipMalloc batch");
        if (state_total)
            qsv::check(hipMalloc((void **)&d_states_, state_total * sizeof(qsv::c64)), "hipMalloc batch states");
        if (partial_total)
            qsv::check(hipMalloc((void **)&d_partials_, partial_total * sizeof(double)), "hipMalloc batch");
        int grid = 0;
        if (!small.empty())
        {
            qsv::batch_prepare(T.block_size, small_max_n);
            grid = std::min<int>(qsv::batch_lds_grid(T.block_size, small_max_n), (int)small.size());
        }
        // ---- streams and graph capture ----
        const int ns = std::max(1, T.streams);
        for (int i = 0; i < ns; ++i)
        {
            hipStream_t s;
            qsv::check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
            streams_.push_back(s);
        }
        for (int i = 0; i < 2 * ns; ++i)
        {
            hipEvent_t e;
            qsv::check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
            fork_join_.push_back(e);
        }
        hipStream_t s0 = streams_[0];
        const int nsmall = (int)small.size();
        // the whole batch as launches on the streams: captured once into the graph below,
        // or issued directly by run_direct
        work_ = [this, runs = std::move(runs), grid, nsmall, small_max_n, ns, block_size = T.block_size]() {
            hipStream_t s0 = streams_[0];
            qsv::check(hipEventRecord(fork_join_[0], s0), "hipEventRecord");
            for (int i = 1; i < ns; ++i)
                qsv::check(hipStreamWaitEvent(streams_[i], fork_join_[0], 0), "hipStreamWaitEvent");
            if (nsmall > 0)
                qsv::launch_batch_lds(s0, block_size, grid, small_max_n, d_circs_, nsmall, d_ops_, d_out_, d_counter_);
            // large circuits, biggest first, over the streams (the LDS kernel occupies s0)
            std::vector<int> order(runs.size());
            for (size_t i = 0; i < order.size(); ++i)
                order[i] = (int)i;
            std::sort(order.begin(), order.end(), [&](int a, int b) { return circuits_[runs[a].ci].n > circuits_[runs[b].ci].n; });
            for (size_t oi = 0; oi < order.size(); ++oi)
            {
                auto lr = runs[order[oi]];
                const Circuit &C = circuits_[lr.ci];
                hipStream_t s = streams_[ns == 1 ? 0 : 1 + oi % (ns - 1)];
                qsv::c64 *st = d_states_ + lr.state_offset;
                qsv::launch_fill_zero_c64(s, st, uint64_t(1) << C.n);
                for (auto &it : lr.items)
                {
                    if (it.stage)
                        launch_prepared_stage(s, st, C.n, it.st, d_mats_);
                    else
                        launch_prepared(s, st, C.n, it.b, d_mats_);
                }
                qsv::launch_zsum(s, st, C.n, d_partials_ + lr.partial_offset, d_out_ + C.out_offset);
            }
            for (int i = 1; i < ns; ++i)
            {
                qsv::check(hipEventRecord(fork_join_[ns + i], streams_[i]), "hipEventRecord");
                qsv::check(hipStreamWaitEvent(s0, fork_join_[ns + i], 0), "hipStreamWaitEvent");
            }
        };
        qsv::check(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try
        {
            work_();
        }
        catch (...)
        {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s0, &g);
            if (g)
                (void)hipGraphDestroy(g);
            throw;
        }
        qsv::check(hipStreamEndCapture(s0, &graph_), "hipStreamEndCapture");
        size_t nodes = 0;
        qsv::check(hipGraphGetNodes(graph_, nullptr, &nodes), "hipGraphGetNodes");
        graph_nodes_ = nodes;
        qsv::check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
        prepared_ = true;
    }

    double Batch::run(double *expz)
    {
        prepare();
        hipStream_t s0 = streams_[0];
        qsv::check(hipEventRecord(t0_, s0), "hipEventRecord");
        qsv::check(hipGraphLaunch(exec_, s0), "hipGraphLaunch");
        qsv::check(hipEventRecord(t1_, s0), "hipEventRecord");
        qsv::check(hipEventSynchronize(t1_), "hipEventSynchronize");
        float ms = 0;
        qsv::check(hipEventElapsedTime(&ms, t0_, t1_), "hipEventElapsedTime");
        if (expz)
            qsv::check(hipMemcpy(expz, d_out_, total_qubits_ * sizeof(double), hipMemcpyDeviceToHost), "batch copy");
        return ms * 1e-3;
    }

    double Batch::run_direct(double *expz)
    {
        prepare();
        hipStream_t s0 = streams_[0];
        qsv::check(hipEventRecord(t0_, s0), "hipEventRecord");
        work_();
        qsv::check(hipEventRecord(t1_, s0), "hipEventRecord");
        qsv::check(hipEventSynchronize(t1_), "hipEventSynchronize");
        float ms = 0;
        qsv::check(hipEventElapsedTime(&ms, t0_, t1_), "hipEventElapsedTime");
        if (expz)
            qsv::check(hipMemcpy(expz, d_out_, total_qubits_ * sizeof(double), hipMemcpyDeviceToHost), "batch copy");
        return ms * 1e-3;
    }

    double Batch::run_plain(double *expz)
    {
        std::map<int, std::unique_ptr<Simulator>> sims;
        const auto t0 = std::chrono::steady_clock::now();
        for (const auto &c : circuits_)
        {
            auto &sim = sims[c.n];
            if (!sim)
            {
                sim = std::make_unique<Simulator>();
                sim->num_qubits_cfg = c.n;
                sim->allocate();
            }
            else
                sim->set_zero_state();
            for (const auto &g : c.gates)
            {
                std::vector<int> pos;
                if (g.control >= 0)
                    pos.push_back(g.control);
                sim->operate(make_gate(g.name.c_str(), g.params), {g.target}, {}, pos);
            }
            std::vector<double> z(c.n);
            sim->expect_z(z.data(), c.n);
            if (expz)
                std::copy(z.begin(), z.end(), expz + c.out_offset);
        }
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    BatchTuning Batch::autotune()
    {
        auto measure = [&](const BatchTuning &t) {
            set_tuning(t);
            run(nullptr); // warm-up (first replay)
            double best = 1e30;
            for (int r = 0; r < 2; ++r)
                best = std::min(best, run(nullptr));
            return best;
        };
        BatchTuning best = tuning_;
        double best_s = measure(best);
        auto try_set = [&](auto setter, std::initializer_list<int> values) {
            for (int v : values)
            {
                BatchTuning t = best;
                setter(t, v);
                double s = measure(t);
                if (s < best_s)
                {
                    best_s = s;
                    best = t;
                }
            }
        };
        try_set([](BatchTuning &t, int v) { t.lds_max_qubits = v; }, {12, 13, 14});
        try_set([](BatchTuning &t, int v) { t.block_size = v; }, {256, 512, 1024});
        try_set([](BatchTuning &t, int v) { t.block_qubits = v; }, {4, 5});
        try_set([](BatchTuning &t, int v) { t.streams = v; }, {1, 2, 4});
        try_set([](BatchTuning &t, int v) { t.stage_bits = v; }, {0, 11, 12, 13});
        set_tuning(best);
        return best;
    }
} // namespace qcore
