// Batching form host side: partition, plan, capture into one hipGraph, autotune.
//
// Circuits of up to `lds_max_qubits` qubits go to the persistent LDS-resident kernel (one
// workgroup per circuit, work-counter driven); larger ones keep a complex64 state in HBM.
// Those of at least stage_bits + 4 qubits run as stages on LDS tiles, batched across
// circuits: one launch per stage round covers the r-th stage of every such circuit.  The
// rest run as fused k-qubit blocks, one launch each, spread over several streams.
// Everything -- counter reset, the persistent kernel, the stage rounds, the fills and
// block sequences of the smaller HBM circuits, the <Z> reductions -- is captured once into
// a hipGraph and replayed with a single launch.
#include "batch.hpp"
#include "simulator.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
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

    void Batch::set_precision(int mode)
    {
        if (mode != 1 && mode != 3)
            throw std::runtime_error("batch: precision must be 1 (FP32 MFMA) or 3 (BF16x3 MFMA)");
        release();
        x3_ = mode == 3;
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
        // ---- large circuits: complex64 states in HBM ----
        // Circuits of >= stage_bits + 4 qubits run in stage rounds: launch r applies stage r
        // of every such circuit (tiles numbered across the circuits), after one basis-state
        // fill for all of them and before one <Z> reduction for all of them.  The others run
        // their fused blocks one launch each, spread over the streams.
        struct Staged
        {
            int ci;
            std::vector<PreparedStage> stages;
        };
        struct Blocked
        {
            int ci;
            std::vector<PreparedBlock> blocks;
        };
        std::vector<Staged> staged;
        std::vector<Blocked> blocked;
        std::vector<float> mats;
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
            if (T.stage_bits > 0 && C.n >= T.stage_bits + 4 && !rg.empty())
            {
                const auto plan = plan_stages(rg, C.n, T.stage_bits, kq);
                // batched gates span at most 2 qubits, so every stage has a tile
                bool all_tiled = true;
                for (const auto &st : plan)
                    all_tiled = all_tiled && !st.tile.empty();
                if (all_tiled)
                {
                    Staged sc{ci, {}};
                    for (const auto &st : plan)
                        sc.stages.push_back(prepare_stage(rg, st, kq, T.stage_bits, C.n, mats, x3_));
                    staged.push_back(std::move(sc));
                    continue;
                }
            }
            blocked.push_back(Blocked{ci, prepare_blocks(rg, plan_greedy(rg, kq), kq, C.n, mats, x3_)});
        }
        // largest first: round launches hand out the biggest states' tiles first, and with
        // sizes descending every state starts at a multiple of its own size
        std::stable_sort(staged.begin(), staged.end(), [&](const Staged &a, const Staged &b) { return circuits_[a.ci].n > circuits_[b.ci].n; });
        std::stable_sort(blocked.begin(), blocked.end(), [&](const Blocked &a, const Blocked &b) { return circuits_[a.ci].n > circuits_[b.ci].n; });
        uint64_t state_total = 0;
        size_t partial_total = 0;
        std::vector<uint64_t> state_off, partial_off; // staged circuits, then blocked ones
        auto place = [&](int ci) {
            state_off.push_back(state_total);
            partial_off.push_back(partial_total);
            state_total += uint64_t(1) << circuits_[ci].n;
            partial_total += size_t(qsv::zsum_blocks(circuits_[ci].n)) * 32;
        };
        for (const auto &sc : staged)
            place(sc.ci);
        for (const auto &bc : blocked)
            place(bc.ci);
        // ---- device buffers (addresses go into the task arrays below) ----
        auto upload = [](void **dst, const void *src, size_t bytes) {
            if (bytes == 0)
                return;
            qsv::check(hipMalloc(dst, bytes), "hipMalloc batch");
            qsv::check(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), "batch upload");
        };
        upload((void **)&d_circs_, small.data(), small.size() * sizeof(qsv::BatchCircuit));
        upload((void **)&d_ops_, ops.data(), ops.size() * sizeof(qsv::BatchOp));
        qsv::check(hipMalloc((void **)&d_out_, std::max(1, total_qubits_) * sizeof(double)), "hipMalloc batch");
        qsv::check(hipMalloc((void **)&d_counter_, sizeof(unsigned)), "hipMalloc batch");
        if (state_total)
            qsv::check(hipMalloc((void **)&d_states_, state_total * sizeof(qsv::c64)), "hipMalloc batch states");
        if (partial_total)
            qsv::check(hipMalloc((void **)&d_partials_, partial_total * sizeof(double)), "hipMalloc batch");
        // ---- staged circuits: task arrays of the rounds, the fill and the reduction ----
        struct Round
        {
            size_t task_offset;
            int ntasks;
            uint64_t tiles;
        };
        std::vector<Round> rounds;
        size_t zsum_offset = 0;
        uint32_t zsum_blocks_total = 0, zsum_qubits_total = 0;
        const int kq_staged = T.block_qubits;
        if (!staged.empty())
        {
            size_t nrounds = 0;
            for (const auto &sc : staged)
                nrounds = std::max(nrounds, sc.stages.size());
            for (size_t r = 0; r < nrounds; ++r)
            {
                std::vector<qsv::StageTask> tasks;
                uint64_t tiles = 0;
                for (size_t i = 0; i < staged.size(); ++i)
                {
                    if (r >= staged[i].stages.size())
                        continue;
                    const PreparedStage &p = staged[i].stages[r];
                    if (p.kq != kq_staged)
                        throw std::runtime_error("batch: stage round with mixed block sizes");
                    tasks.push_back(stage_task(p, d_states_ + state_off[i], tiles));
                    if (r == 0)
                        tasks.back().flags = qsv::kStageZeroState; // the first stage starts from |0...0>
                    tiles += uint64_t(1) << (p.n - p.m);
                }
                rounds.push_back(Round{append_stage_tasks(mats, tasks), (int)tasks.size(), tiles});
            }
            std::vector<qsv::ZsumTask> zt;
            for (size_t i = 0; i < staged.size(); ++i)
            {
                const int n = circuits_[staged[i].ci].n;
                qsv::ZsumTask z{};
                z.state = d_states_ + state_off[i];
                z.part = d_partials_ + partial_off[i];
                z.out = d_out_ + circuits_[staged[i].ci].out_offset;
                z.block_begin = zsum_blocks_total;
                z.qubit_begin = zsum_qubits_total;
                z.n = n;
                z.nblocks = qsv::zsum_blocks(n);
                zsum_blocks_total += uint32_t(z.nblocks);
                zsum_qubits_total += uint32_t(n);
                zt.push_back(z);
            }
            static_assert(sizeof(qsv::ZsumTask) % 16 == 0, "ZsumTask keeps 16-B alignment");
            auto append = [&](const void *src, size_t bytes) {
                while (mats.size() % 4)
                    mats.push_back(0.f);
                const size_t off = mats.size();
                mats.resize(off + (bytes + sizeof(float) - 1) / sizeof(float));
                std::memcpy(mats.data() + off, src, bytes);
                return off;
            };
            zsum_offset = append(zt.data(), zt.size() * sizeof(qsv::ZsumTask));
        }
        upload((void **)&d_mats_, mats.data(), mats.size() * sizeof(float));
        // algorithmic traffic and launches of one run (see qcs_ext_batch_traffic)
        hbm_bytes_ = 0;
        launches_ = small.empty() ? 0 : 2;
        if (!staged.empty())
        {
            for (const auto &sc : staged)
            {
                // the first stage writes without reading; the reduction reads once
                const double amps = double(uint64_t(1) << circuits_[sc.ci].n);
                hbm_bytes_ += 16.0 * amps * double(sc.stages.size()) - 8.0 * amps + 8.0 * amps;
            }
            launches_ += rounds.size() + 2;
        }
        for (const auto &bc : blocked)
        {
            const double amps = double(uint64_t(1) << circuits_[bc.ci].n);
            hbm_bytes_ += 16.0 * amps * double(bc.blocks.size()) + 8.0 * amps + 8.0 * amps;
            launches_ += 1 + bc.blocks.size() + 2;
        }
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
        const int nstaged = (int)staged.size();
        std::vector<uint64_t> blocked_state(state_off.begin() + nstaged, state_off.end());
        std::vector<uint64_t> blocked_partial(partial_off.begin() + nstaged, partial_off.end());
        // the whole batch as launches on the streams: captured once into the graph below,
        // or issued directly by run_direct
        work_ = [this, blocked = std::move(blocked), blocked_state, blocked_partial, rounds, nstaged, kq_staged,
                 stage_bits = T.stage_bits, zsum_offset, zsum_blocks_total, zsum_qubits_total, grid, nsmall,
                 small_max_n, ns, block_size = T.block_size, x3 = x3_]() mutable {
            hipStream_t s0 = streams_[0];
            qsv::check(hipEventRecord(fork_join_[0], s0), "hipEventRecord");
            for (int i = 1; i < ns; ++i)
                qsv::check(hipStreamWaitEvent(streams_[i], fork_join_[0], 0), "hipStreamWaitEvent");
            if (nsmall > 0)
                qsv::launch_batch_lds(s0, block_size, grid, small_max_n, d_circs_, nsmall, d_ops_, d_out_, d_counter_);
            // the LDS kernel occupies s0; the stage rounds take the next stream, the blocked
            // circuits (biggest first) the ones after it
            int next = 0;
            auto stream = [&]() { return streams_[ns == 1 ? 0 : 1 + (next++) % (ns - 1)]; };
            if (nstaged > 0)
            {
                hipStream_t s = stream();
                for (const auto &r : rounds)
                    launch_stage_tasks(s, kq_staged, stage_bits, x3, d_mats_, r.task_offset, r.ntasks, r.tiles);
                qsv::launch_zsum_tasks(s, reinterpret_cast<const qsv::ZsumTask *>(d_mats_ + zsum_offset), nstaged, zsum_blocks_total, zsum_qubits_total);
            }
            for (size_t i = 0; i < blocked.size(); ++i)
            {
                const Circuit &C = circuits_[blocked[i].ci];
                hipStream_t s = stream();
                qsv::c64 *st = d_states_ + blocked_state[i];
                qsv::launch_fill_zero_c64(s, st, uint64_t(1) << C.n);
                for (auto &b : blocked[i].blocks)
                    launch_prepared(s, st, C.n, b, d_mats_);
                qsv::launch_zsum(s, st, C.n, d_partials_ + blocked_partial[i], d_out_ + C.out_offset);
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
        // block size and stage tile size together: the better block size depends on the
        // tile (FP32 products favour 4-qubit blocks, BF16x3 ones 5-qubit blocks on some tiles)
        try_set([](BatchTuning &t, int v) { t.block_qubits = v / 16; t.stage_bits = v % 16; },
                {4 * 16 + 0, 4 * 16 + 11, 4 * 16 + 12, 4 * 16 + 13, 5 * 16 + 0, 5 * 16 + 11, 5 * 16 + 12, 5 * 16 + 13});
        try_set([](BatchTuning &t, int v) { t.streams = v; }, {1, 2, 4});
        set_tuning(best);
        return best;
    }
} // namespace qcore
