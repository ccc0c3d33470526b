// Host core of libqcs.so.
//
// Semantics follow the reference simulator core (src/qcs.cu:1685-2899):
//  - lazy measurement: measured, non-targeted qubits become implicit controls of their
//    measured value; a contradicting explicit control skips the gate; targets leave the
//    measured lists (src/qcs.cu:2161-2204, 2480-2484, 2679-2703);
//  - measurement: cached result, {p0,p1} over the subspace consistent with earlier
//    measurements, r = std::uniform_real_distribution<double>(0, p0+p1)(std::mt19937_64),
//    result = p0 < r, collapse without renormalisation unless enabled
//    (src/qcs.cu:2545-2677);
//  - mapping perm_p2l / perm_l2p and the swap-in of "global" targets
//    (src/qcs.cu:1860-1885, 2206-2332) -- here the data movement is an in-HBM qubit swap;
//  - capacity check n <= floor(log2(free / 16)) (src/qcs.cu:1895-1908), no swap buffer.
// What differs by design: the operand descriptor is a kernel argument (no per-gate
// __constant__ copy), reductions use the canonical order documented in qcs_ext.h, and all
// work is on one stream owned by the simulator.
#include "simulator.hpp"

#include "fused.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

#include <qcs_ext.h>

namespace qcore
{
    namespace
    {
        std::string fmt(const char *f, int a, int b = 0)
        {
            char buf[256];
            snprintf(buf, sizeof buf, f, a, b);
            return buf;
        }

        bool contains(const std::vector<int> &v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); }

        qsv::c128 to_c(cd z) { return qsv::c128{z.real(), z.imag()}; }

        qsv::Subspace make_sub(std::vector<int> ops, uint64_t set_bits)
        {
            std::sort(ops.begin(), ops.end());
            qsv::Subspace s{};
            s.set_bits = set_bits;
            s.nops = (int)ops.size();
            if (s.nops > qsv::kMaxQubits)
                throw std::runtime_error("too many operands");
            for (int k = 0; k < s.nops; ++k)
                s.ops[k] = ops[k];
            return s;
        }
    } // namespace

    // Runs f on the state pointer of the current precision (complex128 or complex64).
    template <class F>
    void Simulator::with_state(F &&f)
    {
        if (fused_ && fused_->f64())
            f(fused_->state128());
        else if (fused_)
            f(fused_->state());
        else
            f(state_);
    }

    Simulator::Simulator()
    {
        qsv::check(hipGetDevice(&device_), "hipGetDevice");
        qsv::check(hipStreamCreate(&stream_), "hipStreamCreate");
        qsv::check(hipHostMalloc((void **)&pinned_, 64 * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        std::random_device rd; // src/qcs.cu:1845-1852
        seed_ = rd();
        engine_.seed(seed_);
    }

    Simulator::~Simulator()
    {
        try
        {
            free_memory();
        }
        catch (...)
        {
        }
        for (auto e : events_)
            (void)hipEventDestroy(e);
        for (auto &r : recs_)
        {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : free_events_)
            (void)hipEventDestroy(e);
        if (open_a_)
            (void)hipEventDestroy(open_a_);
        if (pinned_)
            (void)hipHostFree(pinned_);
        if (stream_)
            (void)hipStreamDestroy(stream_);
    }

    void Simulator::need_alloc() const
    {
        if (!allocated_)
            throw std::runtime_error("allocate_memory has not been called");
    }

    void Simulator::set_seed(uint64_t s)
    {
        seed_ = s;
        engine_.seed(s);
    }

    void Simulator::set_precision(int mode, int k)
    {
        if (allocated_)
            throw std::runtime_error("set_precision must be called before allocate_memory");
        if (mode < 0 || mode > 3)
            throw std::runtime_error("precision mode must be 0 (complex128 per gate), 1 (complex64 fused, FP32), 2 (complex128 fused) or 3 (complex64 fused, BF16x3)");
        if (mode != 0 && (k < 1 || k > 5))
            throw std::runtime_error("fused block size must be in [1, 5]");
        precision_ = mode;
        block_qubits_ = k;
    }

    void Simulator::set_stage_bits(int m)
    {
        if (m != 0 && (m < 11 || m > 14))
            throw std::runtime_error("stage tile bits must be 0 (one sweep per block), 11, 12, 13 or 14");
        stage_bits_ = m;
        if (fused_)
            fused_->set_stage_bits(m);
    }

    void Simulator::set_stage_stream(int on)
    {
        if (on != 0 && on != 1)
            throw std::runtime_error("stage stream setting must be 0 or 1");
        stage_stream_ = on;
        if (fused_)
            fused_->set_stage_stream(on != 0);
    }

    void Simulator::set_planner_check(int on)
    {
        if (on != 0 && on != 1)
            throw std::runtime_error("planner check setting must be 0 or 1");
        flush();
        planner_check_ = on;
        if (fused_)
            fused_->set_check(on != 0);
    }

    void Simulator::reinitialize_mapping()
    { // src/qcs.cu:1860-1885
        state_changed();
        p2l_.assign(n_, 0);
        l2p_.assign(n_, 0);
        if (initial_p2l.empty())
        {
            for (int q = 0; q < n_; ++q)
                p2l_[q] = l2p_[q] = q;
        }
        else
        {
            if ((int)initial_p2l.size() != n_)
                throw std::runtime_error(fmt("mapping size %d does not match num_qubits %d", (int)initial_p2l.size(), n_));
            p2l_ = initial_p2l;
            for (int q = 0; q < n_; ++q)
                l2p_[p2l_[q]] = q;
        }
    }

    void Simulator::allocate()
    { // src/qcs.cu:1887-1998
        if (allocated_)
            throw std::runtime_error(fmt("num_qubits is already set %d > 0", n_));
        int nq = num_qubits_cfg;
        if (nq < 1 || nq > qsv::kMaxQubits)
            throw std::runtime_error(fmt("num_qubits %d out of range [1, %d]", nq, qsv::kMaxQubits));
        size_t free_b = 0, total_b = 0;
        qsv::check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const uint64_t S = c64_state() ? 8 : 16;
        int cap = 0;
        while ((uint64_t(1) << (cap + 1)) <= free_b / S)
            ++cap;
        if (nq > cap)
            throw std::runtime_error(fmt("num_qubits(%d) > max_num_qubits_local_device(%d)", nq, cap));
        n_ = nq;
        n_local_ = nq;
        reinitialize_mapping();
        const uint64_t count = uint64_t(1) << n_;
        if (precision_ == 0)
        {
            qsv::check(hipMalloc((void **)&state_, count * sizeof(qsv::c128)), "hipMalloc state");
        }
        else
        {
            fused_ = new FusedEngine(*this, block_qubits_, precision_);
            fused_->set_stage_bits(stage_bits_);
            fused_->set_stage_stream(stage_stream_ != 0);
            fused_->set_check(planner_check_ != 0);
        }
        work_len_ = qsv::reduce_workspace(count) + 16;
        qsv::check(hipMalloc((void **)&work_, work_len_ * sizeof(double)), "hipMalloc workspace");
        allocated_ = true;
        set_zero_state();
    }

    void Simulator::free_memory()
    {
        if (fused_)
        {
            delete fused_;
            fused_ = nullptr;
        }
        if (state_)
        {
            (void)hipStreamSynchronize(stream_);
            (void)hipFree(state_);
            state_ = nullptr;
        }
        if (work_)
        {
            (void)hipFree(work_);
            work_ = nullptr;
        }
        for (auto p : levels_)
            (void)hipFree(p);
        levels_.clear();
        if (stage_dev_)
        {
            (void)hipFree(stage_dev_);
            stage_dev_ = nullptr;
        }
        if (stage_host_)
        {
            (void)hipHostFree(stage_host_);
            stage_host_ = nullptr;
        }
        stage_len_ = 0;
        allocated_ = false;
        n_ = 0;
        state_changed();
    }

    void Simulator::set_zero_state()
    { // src/qcs.cu:2065-2079
        need_alloc();
        state_changed();
        clear_measurement_state();
        if (fused_)
            fused_->drop(); // queued gates act on a state that is being overwritten
        prof_begin(QCS_EXT_K_INIT, (c64_state() ? 8.0 : 16.0) * (1ull << n_));
        with_state([&](auto *st) { qsv::launch_fill_basis(stream_, st, uint64_t(1) << n_, 0); });
        prof_end();
    }

    void Simulator::set_sequential_state()
    { // src/qcs.cu:71-76, 2025-2043
        need_alloc();
        state_changed();
        clear_measurement_state();
        if (fused_)
            fused_->drop();
        with_state([&](auto *st) { qsv::launch_fill_sequential(stream_, st, uint64_t(1) << n_); });
    }

    void Simulator::set_flat_state()
    { // src/qcs.cu:78-82, 2045-2063
        need_alloc();
        state_changed();
        clear_measurement_state();
        if (fused_)
            fused_->drop();
        with_state([&](auto *st) { qsv::launch_fill_flat(stream_, st, uint64_t(1) << n_); });
    }

    void Simulator::set_entangled_state()
    { // src/qcs.cu:2081-2095 with one rank: amp[0] = 1, nothing else written
        need_alloc();
        state_changed();
        clear_measurement_state();
        flush(); // amp[0] = 1 on top of the current state
        with_state([&](auto *st) { qsv::launch_set_one(stream_, st, 0); });
    }

    void Simulator::set_random_state()
    { // src/qcs.cu:2097-2119: curand normals there; this build's own generator here.
      // The measured lists are kept, as in the reference.
        need_alloc();
        state_changed();
        if (fused_)
            fused_->drop();
        with_state([&](auto *st) { qsv::launch_fill_random(stream_, st, uint64_t(1) << n_, seed_); });
    }

    void Simulator::clear_measurement_state()
    {
        m0_.clear();
        m1_.clear();
    }

    void Simulator::check_operands(const std::vector<int> &t, const std::vector<int> &ng, const std::vector<int> &ps) const
    {
        std::vector<int> all;
        all.insert(all.end(), t.begin(), t.end());
        all.insert(all.end(), ng.begin(), ng.end());
        all.insert(all.end(), ps.begin(), ps.end());
        for (int q : all)
            if (q < 0 || q >= n_)
                throw std::runtime_error(fmt("qubit %d out of range [0, %d)", q, n_));
        std::sort(all.begin(), all.end());
        if (std::adjacent_find(all.begin(), all.end()) != all.end())
            throw std::runtime_error("a qubit appears more than once among targets and controls");
    }

    // The in-HBM replacement of the reference's NCCL global<->local exchange for one pair:
    // swap the data of physical qubits a and b with the SWAP pair kernel (it reads and writes
    // only the half of the state the swap changes).  complex128 per-gate state only.
    void Simulator::swap_physical(int a, int b)
    {
        if (a == b)
            return;
        qsv::Subspace sub = make_sub({a, b}, 0);
        qsv::PairOp op{};
        op.offA = uint64_t(1) << a;
        op.offB = uint64_t(1) << b;
        op.m[0] = {0, 0};
        op.m[1] = {1, 0};
        op.m[2] = {1, 0};
        op.m[3] = {0, 0};
        prof_begin(QCS_EXT_K_PERMUTE, 2.0 * 16.0 * (1ull << (n_ - 1)));
        qsv::launch_pair(stream_, state_, n_, sub, op);
        prof_end();
    }

    // The mapping after the data of physical qubit q moved to physical qubit dst[q].
    void Simulator::remap_physical(const std::vector<int> &dst)
    {
        std::vector<int> p2l(n_);
        for (int q = 0; q < n_; ++q)
            p2l[dst[q]] = p2l_[q];
        p2l_ = p2l;
        for (int q = 0; q < n_; ++q)
            l2p_[p2l_[q]] = q;
    }

    // Moves the data of every physical qubit q to physical qubit dst[q]: one sweep of the
    // LDS-tiled permutation kernel per pass (a pass moves up to 6 qubits above the low run
    // that every tile keeps contiguous: 64 complex128 or 128 complex64 amplitudes, 1 KiB),
    // then the mapping follows the data.
    void Simulator::move_physical(const std::vector<int> &dst)
    {
        bool any = false;
        for (int q = 0; q < n_; ++q)
            any = any || dst[q] != q;
        if (!any)
            return;
        flush();
        state_changed();
        const int tb = c64_state() ? qsv::permute_tile_bits<qsv::c64>() : qsv::permute_tile_bits<qsv::c128>();
        const int low = c64_state() ? 7 : 6;
        const auto passes = n_ <= tb ? std::vector<std::vector<int>>{dst} : qsv::plan_permute_passes(dst, tb - low, low);
        const double S = c64_state() ? 8.0 : 16.0;
        for (const auto &p : passes)
        {
            prof_begin(QCS_EXT_K_PERMUTE, 2.0 * S * (1ull << n_));
            with_state([&](auto *st) { qsv::launch_permute(stream_, st, n_, p); });
            prof_end();
        }
        remap_physical(dst);
    }

    void Simulator::ensure_local(std::vector<int> &tphys)
    { // src/qcs.cu:2206-2332 (bookkeeping) with the data move done in HBM
        // Slot for the i-th global target: n_local-1-i as in the reference, except that
        // slots holding a local target of the same gate are skipped (the reference would move
        // that target away and apply the gate to the wrong qubit).
        std::vector<int> glob, loc;
        int next_slot = n_local_ - 1;
        for (auto &t : tphys)
        {
            if (t >= n_local_)
            {
                while (std::find(tphys.begin(), tphys.end(), next_slot) != tphys.end())
                    --next_slot;
                glob.push_back(t);
                loc.push_back(next_slot);
                t = next_slot--;
            }
        }
        if (glob.empty())
            return;
        std::vector<int> dst(n_);
        for (int q = 0; q < n_; ++q)
            dst[q] = q;
        for (size_t i = 0; i < glob.size(); ++i)
        {
            dst[glob[i]] = loc[i];
            dst[loc[i]] = glob[i];
        }
        // one swap of two qubits at 3 or above: the SWAP kernel moves half the state,
        // coalesced; anything else takes one permutation sweep for all the swaps
        if (!fused_ && glob.size() == 1 && std::min(glob[0], loc[0]) >= 3)
        {
            state_changed();
            swap_physical(glob[0], loc[0]);
            remap_physical(dst);
        }
        else
        {
            move_physical(dst);
        }
    }

    void Simulator::operate(const GateSpec &g, std::vector<int> targets, std::vector<int> neg, std::vector<int> pos)
    {
        need_alloc();
        state_changed();
        if ((int)targets.size() != g.ntargets)
            throw std::runtime_error(fmt("gate needs %d target qubit(s), %d given", g.ntargets, (int)targets.size()));
        check_operands(targets, neg, pos);
        // lazy measurement (src/qcs.cu:2161-2204)
        std::vector<int> *ctrl[2] = {&neg, &pos};
        const std::vector<int> *meas[2] = {&m0_, &m1_};
        std::vector<int> keep[2];
        bool condition = true;
        for (int v = 0; v < 2; ++v)
        {
            for (int q : *meas[v])
            {
                if (contains(targets, q))
                    continue;
                keep[v].push_back(q);
                bool same = contains(*ctrl[v], q), other = contains(*ctrl[1 - v], q);
                if (!same && !other)
                    ctrl[v]->push_back(q);
                else if (other)
                    condition = false;
            }
        }
        if (!condition)
            return; // src/qcs.cu:2690-2691
        ResolvedGate r;
        r.spec = g;
        r.tphys.resize(targets.size());
        for (size_t j = 0; j < targets.size(); ++j)
            r.tphys[j] = l2p_[targets[j]];
        ensure_local(r.tphys);
        for (int q : pos)
            r.pphys.push_back(l2p_[q]);
        for (int q : neg)
            r.nphys.push_back(l2p_[q]);
        for (size_t j = 0; j < targets.size(); ++j)
        {
            if (contains(m0_, targets[j]))
                r.mmask |= 1u << j;
            else if (contains(m1_, targets[j]))
            {
                r.mmask |= 1u << j;
                r.mval |= 1u << j;
            }
        }
        if (fused_)
            fused_->enqueue(r);
        else
            apply_resolved(r);
        m0_ = keep[0];
        m1_ = keep[1];
    }

    void Simulator::apply_resolved(const ResolvedGate &r)
    {
        const GateSpec &g = r.spec;
        uint64_t setb = 0;
        for (int q : r.pphys)
            setb |= uint64_t(1) << q;
        std::vector<int> ctrl_ops = r.pphys;
        ctrl_ops.insert(ctrl_ops.end(), r.nphys.begin(), r.nphys.end());
        auto readable = [&](unsigned pattern) { return ((pattern ^ r.mval) & r.mmask) == 0; };
        const int nc = (int)ctrl_ops.size();
        switch (g.kind)
        {
        case Kind::Identity:
            return;
        case Kind::GlobalPhase:
        {
            auto sub = make_sub(ctrl_ops, setb);
            prof_begin(QCS_EXT_K_PHASE, 32.0 * (1ull << (n_ - nc)));
            qsv::launch_phase(stream_, state_, n_, sub, to_c(g.phase[0]));
            prof_end();
            return;
        }
        case Kind::PhaseOne:
        {
            int t = r.tphys[0];
            std::vector<int> ops = ctrl_ops;
            ops.push_back(t);
            auto sub = make_sub(ops, setb | (uint64_t(1) << t));
            // a target measured as 0 makes the |1> input read as zero (src/qcs.cu:1153-1171)
            cd ph = readable(1) ? g.phase[1] : cd(0.0, 0.0);
            prof_begin(QCS_EXT_K_PHASE, 32.0 * (1ull << (n_ - nc - 1)));
            qsv::launch_phase(stream_, state_, n_, sub, to_c(ph));
            prof_end();
            return;
        }
        case Kind::Diag1:
        case Kind::Diag2:
        {
            auto sub = make_sub(ctrl_ops, setb);
            qsv::DiagOp op{};
            op.ntargets = g.ntargets;
            for (int j = 0; j < g.ntargets; ++j)
                op.t[j] = r.tphys[j];
            for (int k = 0; k < (1 << g.ntargets); ++k)
                op.phase[k] = readable(k) ? to_c(g.phase[k]) : qsv::c128{0.0, 0.0};
            prof_begin(QCS_EXT_K_DIAG, 32.0 * (1ull << (n_ - nc)));
            qsv::launch_diag(stream_, state_, n_, sub, op);
            prof_end();
            return;
        }
        case Kind::Pair1:
        case Kind::PairSub:
        {
            std::vector<int> ops = ctrl_ops;
            ops.insert(ops.end(), r.tphys.begin(), r.tphys.end());
            auto sub = make_sub(ops, setb);
            qsv::PairOp op{};
            int pa = g.kind == Kind::Pair1 ? 0 : g.pat[0];
            int pb = g.kind == Kind::Pair1 ? 1 : g.pat[1];
            auto off = [&](int pat) {
                uint64_t o = 0;
                for (int j = 0; j < g.ntargets; ++j)
                    if ((pat >> j) & 1)
                        o |= uint64_t(1) << r.tphys[j];
                return o;
            };
            op.offA = off(pa);
            op.offB = off(pb);
            for (int i = 0; i < 4; ++i)
                op.m[i] = to_c(g.m[i]);
            op.zero_mask = (readable(pa) ? 0u : 1u) | (readable(pb) ? 0u : 2u);
            prof_begin(QCS_EXT_K_PAIR, 64.0 * (1ull << (n_ - (int)ops.size())));
            qsv::launch_pair(stream_, state_, n_, sub, op);
            prof_end();
            return;
        }
        case Kind::Dense4:
        case Kind::Mono:
        {
            std::vector<int> ops = ctrl_ops;
            ops.insert(ops.end(), r.tphys.begin(), r.tphys.end());
            auto sub = make_sub(ops, setb);
            const int K = 1 << g.ntargets;
            uint64_t offs[16];
            for (int k = 0; k < K; ++k)
            {
                offs[k] = 0;
                for (int j = 0; j < g.ntargets; ++j)
                    if ((k >> j) & 1)
                        offs[k] |= uint64_t(1) << r.tphys[j];
            }
            unsigned rmask = 0;
            for (int k = 0; k < K; ++k)
                if (readable(k))
                    rmask |= 1u << k;
            const uint64_t items = 1ull << (n_ - (int)ops.size());
            if (g.kind == Kind::Dense4)
            {
                qsv::Dense4Op op{};
                for (int k = 0; k < 4; ++k)
                    op.off[k] = offs[k];
                for (int i = 0; i < 16; ++i)
                    op.m[i] = to_c(g.m[i]);
                // rows equal to the identity row are not written (the reference's apply()
                // skips them too)
                op.write_mask = 0;
                unsigned needed = 0;
                for (int row = 0; row < 4; ++row)
                {
                    bool ident = true;
                    for (int c = 0; c < 4; ++c)
                        if (g.m[4 * row + c] != (row == c ? cd(1.0, 0.0) : cd(0.0, 0.0)))
                            ident = false;
                    if (!ident)
                    {
                        op.write_mask |= 1u << row;
                        for (int c = 0; c < 4; ++c)
                            if (g.m[4 * row + c] != cd(0.0, 0.0))
                                needed |= 1u << c;
                    }
                }
                op.read_mask = rmask & needed;
                double bytes = 16.0 * (__builtin_popcount(op.read_mask) + __builtin_popcount(op.write_mask)) * items;
                prof_begin(QCS_EXT_K_DENSE4, bytes);
                qsv::launch_dense4(stream_, state_, n_, sub, op);
                prof_end();
            }
            else
            {
                qsv::MonoOp op{};
                op.ntargets = g.ntargets;
                op.nrows = g.nrows;
                for (int k = 0; k < K; ++k)
                    op.off[k] = offs[k];
                for (int k = 0; k < g.nrows; ++k)
                {
                    op.row[k] = g.row[k];
                    op.src[k] = g.src[k];
                    op.coef[k] = to_c(g.coef[k]);
                }
                op.read_mask = rmask;
                prof_begin(QCS_EXT_K_MONO, 32.0 * g.nrows * items);
                qsv::launch_mono(stream_, state_, n_, sub, op);
                prof_end();
            }
            return;
        }
        }
    }

    int Simulator::measure(int q)
    { // src/qcs.cu:2545-2677
        need_alloc();
        if (q < 0 || q >= n_)
            throw std::runtime_error(fmt("qubit %d out of range [0, %d)", q, n_));
        if (contains(m0_, q))
            return 0;
        if (contains(m1_, q))
            return 1;
        flush();
        state_changed();
        std::vector<int> tphys = {l2p_[q]};
        ensure_local(tphys);
        const int tq = tphys[0];
        std::vector<int> ops = {tq};
        uint64_t setb = 0;
        for (int x : m1_)
        {
            ops.push_back(l2p_[x]);
            setb |= uint64_t(1) << l2p_[x];
        }
        for (int x : m0_)
            ops.push_back(l2p_[x]);
        auto sub = make_sub(ops, setb);
        const uint64_t items = uint64_t(1) << (n_ - sub.nops);
        const double S = c64_state() ? 8.0 : 16.0;
        prof_begin(QCS_EXT_K_REDUCE, 2.0 * S * items);
        with_state([&](auto *st) { qsv::reduce_pair_probs(stream_, st, sub, uint64_t(1) << tq, items, work_, pinned_); });
        prof_end();
        const double p0 = pinned_[0], p1 = pinned_[1];
        std::uniform_real_distribution<double> dist(0, p0 + p1);
        const double rnd = dist(engine_);
        const int result = p0 < rnd;
        (result ? m1_ : m0_).push_back(q);
        const double pk = result ? p1 : p0;
        const double scale = (renormalize_ && pk > 0) ? 1.0 / std::sqrt(pk) : 1.0;
        prof_begin(QCS_EXT_K_COLLAPSE, S * (1ull << (n_ - 1)) * (scale != 1.0 ? 3.0 : 1.0));
        with_state([&](auto *st) { qsv::launch_collapse(stream_, st, n_, tq, result, scale); });
        prof_end();
        return result;
    }

    uint64_t Simulator::logical_chunk() const
    {
        return std::min<uint64_t>(uint64_t(1) << n_, uint64_t(1) << 22); // 64 MiB of c128
    }

    void Simulator::ensure_stage()
    {
        uint64_t len = logical_chunk();
        if (stage_len_ >= len)
            return;
        if (stage_dev_)
            (void)hipFree(stage_dev_);
        if (stage_host_)
            (void)hipHostFree(stage_host_);
        stage_dev_ = nullptr;
        stage_host_ = nullptr;
        qsv::check(hipMalloc((void **)&stage_dev_, len * sizeof(qsv::c128)), "hipMalloc stage");
        qsv::check(hipHostMalloc(&stage_host_, len * sizeof(qsv::c128), hipHostMallocDefault), "hipHostMalloc stage");
        stage_len_ = len;
    }

    void Simulator::save_statevector(const char *path)
    { // src/qcs.cu:2494-2543: raw little-endian complex<double>, logical order, no header
        need_alloc();
        flush();
        if (path == nullptr)
            throw std::runtime_error("output file name must not be NULL");
        FILE *fp = std::fopen(path, "wb");
        if (!fp)
            throw std::runtime_error("open failed");
        try
        {
            const uint64_t count = uint64_t(1) << n_;
            ensure_stage();
            const uint64_t chunk = logical_chunk();
            for (uint64_t first = 0; first < count; first += chunk)
            {
                with_state([&](auto *st) { qsv::launch_gather_logical(stream_, st, stage_dev_, first, chunk, l2p_.data(), n_); });
                qsv::check(hipMemcpyAsync(stage_host_, stage_dev_, chunk * sizeof(qsv::c128), hipMemcpyDeviceToHost, stream_), "save copy");
                qsv::check(hipStreamSynchronize(stream_), "save sync");
                if (std::fwrite(stage_host_, sizeof(qsv::c128), chunk, fp) != chunk)
                    throw std::runtime_error("fwrite failed");
            }
        }
        catch (...)
        {
            std::fclose(fp);
            throw;
        }
        if (std::fclose(fp) != 0)
            throw std::runtime_error("close failed");
    }

    void Simulator::get_statevector(double *out, uint64_t count)
    {
        need_alloc();
        flush();
        if (out == nullptr || count != (uint64_t(1) << n_))
            throw std::runtime_error("statevector buffer size mismatch");
        ensure_stage();
        const uint64_t chunk = logical_chunk();
        for (uint64_t first = 0; first < count; first += chunk)
        {
            with_state([&](auto *st) { qsv::launch_gather_logical(stream_, st, stage_dev_, first, chunk, l2p_.data(), n_); });
            qsv::check(hipMemcpyAsync(stage_host_, stage_dev_, chunk * sizeof(qsv::c128), hipMemcpyDeviceToHost, stream_), "copy");
            qsv::check(hipStreamSynchronize(stream_), "sync");
            std::memcpy(out + 2 * first, stage_host_, chunk * sizeof(qsv::c128));
        }
    }

    void Simulator::set_statevector(const double *in, uint64_t count)
    {
        need_alloc();
        flush();
        if (in == nullptr || count != (uint64_t(1) << n_))
            throw std::runtime_error("statevector buffer size mismatch");
        state_changed();
        if (fused_)
            fused_->drop();
        ensure_stage();
        const uint64_t chunk = logical_chunk();
        for (uint64_t first = 0; first < count; first += chunk)
        {
            qsv::check(hipStreamSynchronize(stream_), "sync");
            std::memcpy(stage_host_, in + 2 * first, chunk * sizeof(qsv::c128));
            qsv::check(hipMemcpyAsync(stage_dev_, stage_host_, chunk * sizeof(qsv::c128), hipMemcpyHostToDevice, stream_), "copy");
            with_state([&](auto *st) { qsv::launch_scatter_logical(stream_, st, stage_dev_, first, chunk, l2p_.data(), n_); });
        }
        qsv::check(hipStreamSynchronize(stream_), "sync");
        // a new state: earlier measurement results no longer describe it
        clear_measurement_state();
    }

    void Simulator::load_statevector(const char *path)
    { // src/qcs.cu:2121-2159 (reads physical order from a fixed name); here: any path, the
      // logical-order format save_statevector writes, streamed through the pinned stage
        need_alloc();
        flush();
        if (path == nullptr)
            throw std::runtime_error("input file name must not be NULL");
        FILE *fp = std::fopen(path, "rb");
        if (!fp)
            throw std::runtime_error("open failed");
        try
        {
            const uint64_t count = uint64_t(1) << n_;
            if (std::fseek(fp, 0, SEEK_END) != 0)
                throw std::runtime_error("fseek failed");
            const long long size = std::ftell(fp);
            if (size < 0 || uint64_t(size) != count * sizeof(qsv::c128))
                throw std::runtime_error("statevector file holds " + std::to_string(size) + " bytes, expected " + std::to_string(count * sizeof(qsv::c128)));
            std::rewind(fp);
            state_changed();
            if (fused_)
                fused_->drop();
            ensure_stage();
            const uint64_t chunk = logical_chunk();
            for (uint64_t first = 0; first < count; first += chunk)
            {
                qsv::check(hipStreamSynchronize(stream_), "sync");
                if (std::fread(stage_host_, sizeof(qsv::c128), chunk, fp) != chunk)
                    throw std::runtime_error("fread failed");
                qsv::check(hipMemcpyAsync(stage_dev_, stage_host_, chunk * sizeof(qsv::c128), hipMemcpyHostToDevice, stream_), "copy");
                with_state([&](auto *st) { qsv::launch_scatter_logical(stream_, st, stage_dev_, first, chunk, l2p_.data(), n_); });
            }
            qsv::check(hipStreamSynchronize(stream_), "sync");
        }
        catch (...)
        {
            std::fclose(fp);
            throw;
        }
        std::fclose(fp);
        clear_measurement_state();
    }

    void Simulator::ensure_levels()
    {
        auto sz = qsv::fold_level_sizes(uint64_t(1) << n_);
        if (levels_.size() + 1 == sz.size())
            return;
        for (auto p : levels_)
            (void)hipFree(p);
        levels_.clear();
        for (size_t lev = 1; lev < sz.size(); ++lev)
        {
            double *p = nullptr;
            qsv::check(hipMalloc((void **)&p, sz[lev] * sizeof(double)), "hipMalloc levels");
            levels_.push_back(p);
        }
    }

    double Simulator::probability_total()
    {
        need_alloc();
        flush();
        // the levels are built once per state: sample() after probability_total() (or a
        // second probability_total()) on an unchanged state reuses them
        if (levels_valid_)
            return levels_total_;
        ensure_levels();
        prof_begin(QCS_EXT_K_REDUCE, (c64_state() ? 8.0 : 16.0) * (1ull << n_));
        with_state([&](auto *st) { qsv::build_fold_levels(stream_, st, uint64_t(1) << n_, levels_); });
        prof_end();
        qsv::check(hipMemcpyAsync(pinned_, levels_.back(), sizeof(double), hipMemcpyDeviceToHost, stream_), "copy");
        qsv::check(hipStreamSynchronize(stream_), "sync");
        levels_total_ = pinned_[0];
        levels_valid_ = true;
        return levels_total_;
    }

    void Simulator::sample(uint64_t seed, uint64_t shots, uint64_t *out)
    {
        need_alloc();
        if (shots > 0 && out == nullptr)
            throw std::runtime_error("outcome buffer must not be NULL");
        flush();
        const double total = probability_total();
        if (shots == 0)
            return;
        uint64_t *dev = nullptr;
        qsv::check(hipMalloc((void **)&dev, shots * sizeof(uint64_t)), "hipMalloc outcomes");
        try
        {
            prof_begin(QCS_EXT_K_SAMPLE, 0.0);
            with_state([&](auto *st) { qsv::launch_sample_search(stream_, st, uint64_t(1) << n_, levels_, total, seed, shots, p2l_.data(), n_, dev); });
            prof_end();
            qsv::check(hipMemcpyAsync(out, dev, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, stream_), "copy");
            qsv::check(hipStreamSynchronize(stream_), "sync");
        }
        catch (...)
        {
            (void)hipFree(dev);
            throw;
        }
        (void)hipFree(dev);
    }

    void Simulator::sample_noalloc(uint64_t seed, uint64_t shots, uint64_t *out)
    {
        need_alloc();
        if (fused_ && fused_->has_queued())
            throw std::runtime_error("sample_noalloc: fused gates are queued and their flush may allocate; call qcs_ext_flush first");
        with_state([&](auto *st) { qsv::sample_noalloc(stream_, device_, st, n_, seed, shots, p2l_.data(), pinned_, out); });
    }

    double Simulator::norm()
    {
        need_alloc();
        flush();
        prof_begin(QCS_EXT_K_REDUCE, (c64_state() ? 8.0 : 16.0) * (1ull << n_));
        with_state([&](auto *st) { qsv::reduce_probs(stream_, st, uint64_t(1) << n_, 0, work_, pinned_); });
        prof_end();
        return pinned_[0];
    }

    void Simulator::expect_z(double *out, int count)
    {
        need_alloc();
        flush();
        if (count < 0 || count > n_ || (count > 0 && out == nullptr))
            throw std::runtime_error("expect_z count out of range");
        for (int q = 0; q < count; ++q)
        {
            prof_begin(QCS_EXT_K_REDUCE, (c64_state() ? 8.0 : 16.0) * (1ull << n_));
            with_state([&](auto *st) { qsv::reduce_probs(stream_, st, uint64_t(1) << n_, uint64_t(1) << l2p_[q], work_, pinned_); });
            prof_end();
            out[q] = pinned_[0];
        }
    }

    void Simulator::set_local_qubits(int nl)
    {
        need_alloc();
        if (nl < 1 || nl > n_)
            throw std::runtime_error(fmt("num_local_qubits %d out of range [1, %d]", nl, n_));
        n_local_ = nl;
    }

    void Simulator::get_mapping(int16_t *out, int count) const
    {
        need_alloc();
        if (out == nullptr || count != n_)
            throw std::runtime_error("mapping buffer size mismatch");
        for (int q = 0; q < n_; ++q)
            out[q] = (int16_t)p2l_[q];
    }

    void Simulator::permute(const std::vector<int> &want)
    {
        need_alloc();
        flush();
        state_changed();
        if ((int)want.size() != n_)
            throw std::runtime_error("mapping size mismatch");
        std::vector<bool> used(n_, false);
        for (int l : want)
        {
            if (l < 0 || l >= n_ || used[l])
                throw std::runtime_error("not a permutation");
            used[l] = true;
        }
        // logical qubit want[q] goes to physical slot q
        std::vector<int> dst(n_);
        for (int q = 0; q < n_; ++q)
            dst[l2p_[want[q]]] = q;
        move_physical(dst);
    }

    void Simulator::overlap(Simulator &other, double out[3])
    {
        need_alloc();
        other.need_alloc();
        flush();
        other.flush();
        if (n_ != other.n_ || p2l_ != other.p2l_)
            throw std::runtime_error("overlap needs two states of the same size and qubit mapping");
        if (device_ != other.device_)
            throw std::runtime_error("overlap needs both simulators on the same device");
        qsv::check(hipStreamSynchronize(other.stream_), "hipStreamSynchronize");
        const uint64_t count = uint64_t(1) << n_;
        double *work = nullptr;
        qsv::check(hipMalloc((void **)&work, qsv::overlap_workspace(count) * sizeof(double)), "hipMalloc overlap");
        try
        {
            with_state([&](auto *a) { other.with_state([&](auto *b) { qsv::overlap(stream_, a, b, count, work, out); }); });
        }
        catch (...)
        {
            (void)hipFree(work);
            throw;
        }
        (void)hipFree(work);
    }

    void Simulator::set_plan(bool enable, int tile_bits, int k, int tiles, int window, int rule)
    {
        if (!fused_)
            throw std::runtime_error("a fixed fusion plan needs a fused precision mode (1, 2 or 3, allocated)");
        fused_->set_plan(enable, tile_bits, k, tiles, window, rule);
    }

    void Simulator::set_planner_model(CostModel *m)
    {
        if (!fused_)
            throw std::runtime_error("the fusion planner needs a fused precision mode (1, 2 or 3, allocated)");
        flush();
        fused_->set_model(m);
    }

    void Simulator::planner_stats(double out[5])
    {
        if (!fused_)
            throw std::runtime_error("the fusion planner needs a fused precision mode (1, 2 or 3, allocated)");
        flush();
        fused_->planner_stats(out);
    }

    int Simulator::planner_record_plans(int capacity, float *features, float *log2_us)
    {
        if (!fused_)
            throw std::runtime_error("recording plan timings needs a fused precision mode (1, 2 or 3, allocated)");
        state_changed();
        return fused_->record_plans(capacity, features, log2_us);
    }

    void Simulator::planner_last_choice(int out[5])
    {
        if (!fused_)
            throw std::runtime_error("the fusion planner needs a fused precision mode (1, 2 or 3, allocated)");
        const PlanOption o = fused_->last_choice();
        out[0] = o.stage_bits;
        out[1] = o.k;
        out[2] = o.tiles;
        out[3] = o.window;
        out[4] = o.rule;
    }

    void Simulator::planner_record(int count, uint64_t seed, float *features, float *log2_us)
    {
        if (!fused_)
            throw std::runtime_error("recording block timings needs a fused precision mode (1, 2 or 3, allocated)");
        if (count < 0 || (count > 0 && (!features || !log2_us)))
            throw std::runtime_error("planner_record: bad output buffers");
        state_changed();
        fused_->record(count, seed, features, log2_us);
    }

    void Simulator::fusion_stats(uint64_t *gates, uint64_t *blocks)
    {
        flush();
        *gates = fused_ ? fused_->gates_applied() : 0;
        *blocks = fused_ ? fused_->blocks_applied() : 0;
    }

    void Simulator::flush()
    {
        if (fused_)
            fused_->flush();
    }

    int Simulator::event_create()
    {
        hipEvent_t e;
        qsv::check(hipEventCreate(&e), "hipEventCreate");
        events_.push_back(e);
        return (int)events_.size() - 1;
    }

    void Simulator::event_record(int e)
    {
        if (e < 0 || e >= (int)events_.size())
            throw std::runtime_error(fmt("event %d does not exist", e));
        flush();
        qsv::check(hipEventRecord(events_[e], stream_), "hipEventRecord");
    }

    double Simulator::event_elapsed(int a, int b)
    { // src/qcs.cu:2862-2897: seconds from the float milliseconds of the runtime
        if (a < 0 || a >= (int)events_.size() || b < 0 || b >= (int)events_.size())
            throw std::runtime_error("event does not exist");
        qsv::check(hipEventSynchronize(events_[a]), "hipEventSynchronize");
        qsv::check(hipEventSynchronize(events_[b]), "hipEventSynchronize");
        float ms = 0.0f;
        qsv::check(hipEventElapsedTime(&ms, events_[a], events_[b]), "hipEventElapsedTime");
        return ms * 1e-3;
    }

    // ---- profiling ----
    hipEvent_t Simulator::pool_event()
    {
        if (!free_events_.empty())
        {
            hipEvent_t e = free_events_.back();
            free_events_.pop_back();
            return e;
        }
        hipEvent_t e;
        qsv::check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }

    void Simulator::prof_begin(int cls, double bytes)
    {
        if (!profile_)
            return;
        open_cls_ = cls;
        open_bytes_ = bytes;
        open_a_ = pool_event();
        qsv::check(hipEventRecord(open_a_, stream_), "hipEventRecord");
    }

    void Simulator::prof_end()
    {
        if (!profile_ || open_cls_ < 0)
            return;
        hipEvent_t b = pool_event();
        qsv::check(hipEventRecord(b, stream_), "hipEventRecord");
        recs_.push_back(Rec{open_cls_, open_bytes_, open_a_, b});
        open_cls_ = -1;
        open_a_ = nullptr;
        if (recs_.size() > 4096)
            prof_resolve();
    }

    void Simulator::prof_resolve()
    {
        for (auto &r : recs_)
        {
            qsv::check(hipEventSynchronize(r.b), "hipEventSynchronize");
            float ms = 0.0f;
            qsv::check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
            prof_launches_[r.cls] += 1;
            prof_seconds_[r.cls] += ms * 1e-3;
            prof_bytes_[r.cls] += r.bytes;
            free_events_.push_back(r.a);
            free_events_.push_back(r.b);
        }
        recs_.clear();
    }

    void Simulator::profile_enable(bool on)
    {
        if (!on)
            prof_resolve();
        profile_ = on;
    }

    void Simulator::profile_get(int cls, uint64_t *launches, double *seconds, double *bytes)
    {
        if (cls < 0 || cls >= QCS_EXT_K_COUNT)
            throw std::runtime_error("kernel class out of range");
        flush();
        prof_resolve();
        if (!launches || !seconds || !bytes)
            throw std::runtime_error("result pointer must not be NULL");
        *launches = prof_launches_[cls];
        *seconds = prof_seconds_[cls];
        *bytes = prof_bytes_[cls];
    }

    void Simulator::profile_reset()
    {
        prof_resolve();
        for (int i = 0; i < 16; ++i)
        {
            prof_launches_[i] = 0;
            prof_seconds_[i] = 0;
            prof_bytes_[i] = 0;
        }
    }
} // namespace qcore
