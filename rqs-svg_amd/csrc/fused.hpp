// Precision form: complex64 state with queued gates fused into k-qubit dense blocks.
#pragma once
#include "fused_kernels.hpp"
#include "simulator.hpp"

#include <vector>

namespace qcore
{
    // A block of consecutive-in-dependency-order gates whose qubits (targets and controls)
    // all lie in `qubits`, or a single gate wider than the block size (`wide`).
    struct PlannedBlock
    {
        std::vector<int> qubits; // physical qubits touched by the block's gates
        std::vector<int> gates;  // indices into the gate list, in application order
        bool wide = false;
    };

    // Greedy fusion: each gate joins the block (at or after the last block that touches one
    // of its qubits) that grows least while staying within k qubits; otherwise it opens a
    // new block.  Gates on disjoint qubits commute, so a gate may join a later block.
    std::vector<PlannedBlock> plan_greedy(const std::vector<ResolvedGate> &gates, int k);

    // How one planned block is launched: local bit j of the block matrix is physical qubit
    // qubits[j] (block qubits, then padding); out_ctrls stay outside as subspace controls.
    struct BlockLayout
    {
        std::vector<int> qubits;    // size kq
        std::vector<int> pos_out;   // positive controls handled by the subspace
        std::vector<int> neg_out;   // negative controls handled by the subspace
    };
    BlockLayout layout_block(const std::vector<ResolvedGate> &gates, const PlannedBlock &b, int kq, int n);

    // 2^kq x 2^kq matrix (row-major) of the block's gates composed in FP64, in the local
    // basis of `lay.qubits`; gates' controls that are in lay.qubits act inside the matrix.
    std::vector<cd> block_matrix(const std::vector<ResolvedGate> &gates, const PlannedBlock &b, const BlockLayout &lay);

    // A composed block ready to launch: matrix at matrix_offset (floats) of a matrix buffer.
    struct PreparedBlock
    {
        int kq = 5;
        int gates = 0;
        size_t matrix_offset = 0;
        double bytes = 0; // algorithmic HBM bytes of one launch
        bool lds = true;  // LDS-staged kernel (needs 5 qubits outside the block)
        qsv::Subspace sub{};
        qsv::BlockLdsOp lop{};
        qsv::BlockOp op{};
    };
    // Lays out and composes every planned block; appends the real (2^(kq+1))^2 matrices to
    // `mats` (host) in launch order.
    std::vector<PreparedBlock> prepare_blocks(const std::vector<ResolvedGate> &gates, const std::vector<PlannedBlock> &plan, int kq, int n, std::vector<float> &mats);
    // Launches one prepared block whose matrices were uploaded to mats_dev.
    void launch_prepared(hipStream_t s, qsv::c64 *state, int n, PreparedBlock &b, const float *mats_dev);

    class FusedEngine
    {
    public:
        FusedEngine(Simulator &sim, int max_block_qubits);
        ~FusedEngine();
        FusedEngine(const FusedEngine &) = delete;
        FusedEngine &operator=(const FusedEngine &) = delete;

        qsv::c64 *state() const { return state_; }
        void enqueue(const ResolvedGate &g);
        void flush();
        void drop() { queue_.clear(); }
        void swap_physical(int a, int b);
        int block_qubits() const { return k_; }
        uint64_t blocks_applied() const { return blocks_applied_; }
        uint64_t gates_applied() const { return gates_applied_; }

    private:
        void apply_blocks(const std::vector<ResolvedGate> &gates, const std::vector<PlannedBlock> &plan);
        void reserve_matrices(size_t floats);

        Simulator &sim_;
        int k_;  // planner limit (qubits per block)
        int kq_; // kernel block size (4 or 5)
        qsv::c64 *state_ = nullptr;
        std::vector<ResolvedGate> queue_;
        float *rdev_ = nullptr;
        float *rhost_ = nullptr;
        size_t rcap_ = 0;
        hipEvent_t copy_done_ = nullptr;
        bool copy_pending_ = false;
        uint64_t blocks_applied_ = 0;
        uint64_t gates_applied_ = 0;
    };
} // namespace qcore
