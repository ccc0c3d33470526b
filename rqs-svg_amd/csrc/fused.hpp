// Precision form: complex64 state with queued gates fused into k-qubit dense blocks.
#pragma once
#include "fused_kernels.hpp"
#include "simulator.hpp"

#include <algorithm>
#include <unordered_map>
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
    // The same with the search window and the choice rule as parameters (see fused.cpp).
    std::vector<PlannedBlock> plan_rule(const std::vector<ResolvedGate> &gates, int k, int window, int rule);

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
    // x3: store each matrix as bfloat16 high parts followed by low parts (BF16x3 kernels).
    std::vector<PreparedBlock> prepare_blocks(const std::vector<ResolvedGate> &gates, const std::vector<PlannedBlock> &plan, int kq, int n, std::vector<float> &mats, bool x3 = false);
    // hi[i] + lo[i] ~ R[i] with both parts bfloat16 bits rounded to nearest even.
    void split_bf16(const double *R, size_t count, uint16_t *hi, uint16_t *lo);
    // Launches one prepared block whose matrices were uploaded to mats_dev.
    void launch_prepared(hipStream_t s, qsv::c64 *state, int n, PreparedBlock &b, const float *mats_dev);

    // complex128 blocks (4 qubits, FP64): the 16x16 complex matrices are appended to `mats`
    // as interleaved doubles.
    struct PreparedBlock128
    {
        int gates = 0;
        size_t matrix_offset = 0;
        double bytes = 0;
        bool staged = true;
        qsv::Subspace sub{};
        qsv::Block128Op op{};
    };
    std::vector<PreparedBlock128> prepare_blocks128(const std::vector<ResolvedGate> &gates, const std::vector<PlannedBlock> &plan, int n, std::vector<double> &mats);
    void launch_prepared128(hipStream_t s, qsv::c128 *state, int n, PreparedBlock128 &b, const double *mats_dev);

    // Stages: runs of gates whose qubits all lie in one set of m tile qubits (qubits 0..3
    // always among them), applied as several blocks to LDS-resident tiles in one sweep.
    // Gates are taken in order; a gate that does not fit (or is wider than k, or shares a
    // qubit with an earlier deferred gate it does not commute with -- diagonal gates
    // commute with each other) is deferred to a later stage, so every gate taken commutes
    // with every deferred one before it.  A stage with an empty tile holds one gate
    // wider than k, run by the per-block path.
    struct PlannedStage
    {
        std::vector<int> tile;  // m tile qubits, ascending (empty: one wide gate)
        std::vector<int> gates; // indices into the gate list, in application order
    };
    // How a stage's tile is chosen.  Grow: qubits 0..3, then the qubits of the gates taken
    // in order until m are reached (free places padded with the lowest free qubits).
    // Window: per stage, the candidate among Grow and every {0..3} + m-4 consecutive qubits
    // that takes the most gates.  Auto: both plans, the one with fewer stages (Grow on a tie).
    enum StageTiles
    {
        kTilesGrow = 0,
        kTilesWindow = 1,
        kTilesAuto = 2
    };
    std::vector<PlannedStage> plan_stages(const std::vector<ResolvedGate> &gates, int n, int m, int k, int tiles = kTilesAuto);

    // A stage ready to launch: its blocks' matrices and descriptors in a matrix buffer.
    struct PreparedStage
    {
        int gates = 0;
        int nblocks = 0;
        int n = 0, kq = 0, m = 0;
        bool x3 = false;
        double bytes = 0;       // one sweep
        size_t desc_offset = 0; // floats: StageBlock array
        size_t task_offset = 0; // floats: its one-state task (attach_stage), 0: none yet
        // floats: the product table of an all-diagonal stage (2^m complex), or none
        size_t table_offset = SIZE_MAX;
        // floats: the product table of the diagonal blocks after the last other block, or none
        size_t trail_offset = SIZE_MAX;
        qsv::c64 *state = nullptr; // the state of attach_stage
        qsv::Subspace tsub{};
    };
    // Plans the stage's gates into blocks of at most kq qubits inside the tile and appends
    // their matrices (FP32, or bfloat16 halves when x3) and descriptors to `mats`.  With
    // table, the diagonal blocks after the last other block become the product of their
    // diagonals over the tile indices (FP64 products rounded to FP32): a stage of diagonal
    // blocks only then runs as that table alone, another applies it as its tiles are stored.
    PreparedStage prepare_stage(const std::vector<ResolvedGate> &gates, const PlannedStage &st, int kq, int m, int n, std::vector<float> &mats, bool x3, bool table = false);
    // Blocks a complex64 stage with tables runs through LDS (as stage128_lds_blocks).
    int stage64_lds_blocks(const std::vector<ResolvedGate> &gates, int kq);
    // The stage's task on `state` as the launch's tiles from tile_begin on (checked).
    qsv::StageTask stage_task(const PreparedStage &p, qsv::c64 *state, uint64_t tile_begin);
    // Appends tasks to `mats` 16-B aligned; returns their offset in floats (never 0).
    size_t append_stage_tasks(std::vector<float> &mats, const std::vector<qsv::StageTask> &tasks);
    // Appends the stage's one-state task on `state` (a device address) to `mats`.
    void attach_stage(PreparedStage &p, qsv::c64 *state, std::vector<float> &mats);
    // Launches an attached stage; mats_dev holds `mats` in device memory.
    // stream: the persistent streaming kernel where it applies (2^13 tiles of a state of 22
    // to 33 qubits), else one workgroup per tile; both give the same bits
    void launch_prepared_stage(hipStream_t s, const PreparedStage &p, const float *mats_dev, bool stream);
    // Launches ntasks tasks (stages of one kq, m and x3 on several states) at task_offset.
    void launch_stage_tasks(hipStream_t s, int kq, int m, bool x3, const float *mats_dev, size_t task_offset, int ntasks, uint64_t total_tiles);

    // complex128 stages (tiles of 2^11, 2^12 or 2^13, FP64 blocks of kq = 3 or 4 qubits): matrices and
    // descriptors appended to `mats` as doubles.
    struct PreparedStage128
    {
        int gates = 0;
        int nblocks = 0;
        double bytes = 0;
        size_t desc_offset = 0; // doubles: StageBlock array
        // doubles: the product table of an all-diagonal stage (2^12 complex), or none
        size_t table_offset = SIZE_MAX;
        // doubles: the product table of the diagonal blocks after the last other block, or none
        size_t trail_offset = SIZE_MAX;
        qsv::Subspace tsub{};
        qsv::Stage128Op op{};
    };
    PreparedStage128 prepare_stage128(const std::vector<ResolvedGate> &gates, const PlannedStage &st, int n, std::vector<double> &mats, int kq = 4);
    // Blocks a complex128 stage over these gates runs through LDS: its greedy blocks up to
    // the last one with a non-diagonal gate (the diagonal blocks after it fold into a table).
    int stage128_lds_blocks(const std::vector<ResolvedGate> &gates, int kq);
    void launch_prepared_stage128(hipStream_t s, qsv::c128 *state, int n, PreparedStage128 &p, const double *mats_dev);

    class CostModel;

    // One fusion plan of a flush, as the learned planner enumerates them.
    struct PlanOption
    {
        int stage_bits; // > 0: stages on tiles of 2^stage_bits amplitudes; 0: one launch per block
        int k;          // qubits per block
        int window;     // blocks examined after the dependency frontier (block plans)
        int rule;       // 0: least growth, 1: fullest block, 2: earliest block that fits
        int tiles = kTilesAuto; // stage plans: tile choice (StageTiles)
    };

    class FusedEngine
    {
    public:
        // mode 1: complex64 state, FP32 MFMA blocks; 2: complex128 state, FP64 blocks of 4
        // qubits; 3: complex64 state, BF16x3 split-precision MFMA blocks.
        FusedEngine(Simulator &sim, int max_block_qubits, int mode);
        ~FusedEngine();
        FusedEngine(const FusedEngine &) = delete;
        FusedEngine &operator=(const FusedEngine &) = delete;

        qsv::c64 *state() const { return state_; }
        qsv::c128 *state128() const { return state128_; }
        bool f64() const { return f64_; }
        bool x3() const { return x3_; }
        // planner feature 2: 1 complex128, 0.5 complex64 BF16x3, 0 complex64 FP32
        float math_feature() const { return f64_ ? 1.f : (x3_ ? 0.5f : 0.f); }
        void enqueue(const ResolvedGate &g);
        void flush();
        void drop() { queue_.clear(); }
        bool has_queued() const { return !queue_.empty(); }
        // LDS tile bits of the stage kernel (11, 12 or 13), or 0 for one launch per block.
        void set_stage_bits(int m);
        void set_stage_stream(bool on);
        int stage_bits() const { return stage_bits_; }
        // tile bits of the default stage plan for this state (complex128: 2^12), 0 if
        // per-block launches are used; stages need at least 16 tiles
        int tile_bits() const
        {
            const int m = f64_ ? 12 : stage_bits_; // complex128 default tiles are 2^12
            return (stage_bits_ > 0 && sim_.n_ >= m + 4) ? m : 0;
        }
        bool uses_stages() const { return tile_bits() > 0; }
        // block size of the default stage plan: the engine's block limit, within what the
        // stage kernels take (complex128: 3 or 4, complex64: 4 or 5)
        // block size of the default stage plan; complex128 stages take 3-qubit FP64 blocks
        // (33-qubit QFT 4.31 s against 4.87 s with 4-qubit blocks, brick depth 40 4.65 s
        // against 5.44 s: half the FP64 work per block outweighs the extra blocks)
        int stage_k() const { return f64_ ? 3 : kq_; }
        // block sizes the stage kernels take for this arithmetic (the planner's choices)
        std::vector<int> stage_block_sizes() const
        {
            std::vector<int> ks;
            for (int k : f64_ ? std::vector<int>{3, 4} : std::vector<int>{4, 5})
                if (k <= kq_)
                    ks.push_back(k);
            return ks;
        }
        int block_qubits() const { return k_; }
        uint64_t blocks_applied() const { return blocks_applied_; }
        uint64_t gates_applied() const { return gates_applied_; }

        // Learned planner: with a model attached, each flush picks the candidate plan with
        // the lowest predicted time unless it does not beat the default plan by the model's
        // margin (CostModel::choice_factor).
        void set_model(CostModel *m);
        // Measured check of the model's overrides (on by default): per flush structure (the
        // gates' names, kinds and qubits; parameters ignored), the first flush for which the
        // model picks a non-default plan runs that plan with events around it, the next flush
        // of the same structure runs the default plan timed, and later ones run whichever of
        // the two was faster.  Off: every flush runs the model's choice.
        void set_check(bool on);
        // A fixed plan for every flush without a model: stages on tiles of 2^tile_bits
        // amplitudes with k-qubit blocks and tile choice `tiles` (StageTiles), or, with
        // tile_bits 0, one launch per block planned with k, window and rule (plan_rule).
        // enable false restores the default plan.
        void set_plan(bool enable, int tile_bits, int k, int tiles, int window, int rule);
        // flushes, learned choices, predicted greedy us, predicted chosen us, blocks
        void planner_stats(double out[5]) const;
        // Time `count` random candidate blocks on this state; features + log2(microseconds).
        void record(int count, uint64_t seed, float *features, float *log2_us);
        // Times the launches of every stage plan the learned planner would consider for the
        // queued gates, each plan applied in turn (the state afterwards is not the circuit's
        // result); at most `capacity` rows of features + log2(microseconds); returns the count.
        int record_plans(int capacity, float *features, float *log2_us);
        // the plan the learned planner ran at the last flush
        PlanOption last_choice() const { return last_choice_; }

    private:
        // the plan of one flush: the model's choice, the fixed plan, or the default
        void apply_plan(const std::vector<ResolvedGate> &gates);
        // candidate plans of the learned planner (index 0: the default plan)
        std::vector<PlanOption> plan_options() const;
        // A candidate plan as planned for its features, kept to be applied if chosen: the
        // stages of a stage option, the blocks of a block option.
        struct PlannedOption
        {
            std::vector<PlannedStage> stages;
            std::vector<PlannedBlock> blocks;
        };
        // feature rows of a plan's launches, in the order apply_option issues them; the plan
        // itself into *keep when given
        void option_features(const std::vector<ResolvedGate> &gates, const PlanOption &o, std::vector<float> &X, PlannedOption *keep = nullptr) const;
        void apply_option(const std::vector<ResolvedGate> &gates, const PlanOption &o);
        // FP64 norm and rescale of a complex64 state (no-op for complex128)
        void renormalize();
        // events around each launch while launch_log_ is set (plan recording)
        void launch_mark_begin();
        void launch_mark_end();
        void apply_blocks(const std::vector<ResolvedGate> &gates, const std::vector<PlannedBlock> &plan);
        void apply_stages(const std::vector<ResolvedGate> &gates, int m, int k, int tiles = kTilesAuto);
        void apply_planned_stages(const std::vector<ResolvedGate> &gates, const std::vector<PlannedStage> &stages, int m, int k);
        void reserve_matrices(size_t bytes);

        Simulator &sim_;
        int k_;  // planner limit (qubits per block)
        int kq_; // kernel block size (4 or 5)
        int stage_bits_ = 13;
        bool stage_stream_ = true;
        bool f64_ = false;
        bool x3_ = false;
        qsv::c64 *state_ = nullptr;
        qsv::c128 *state128_ = nullptr;
        std::vector<ResolvedGate> queue_;
        void *rdev_ = nullptr;  // block matrices on the device
        void *rhost_ = nullptr; // pinned staging for their upload
        size_t rcap_ = 0;       // bytes
        hipEvent_t copy_done_ = nullptr;
        bool copy_pending_ = false;
        uint64_t blocks_applied_ = 0;
        uint64_t gates_applied_ = 0;
        CostModel *model_ = nullptr;
        PlanOption last_choice_{0, 0, 0, 0, kTilesAuto};
        // one flush structure's measured check: stage 1 after the timed run of the model's
        // choice (option), 2 after the timed run of the default plan, 3 once decided
        struct PlanTrial
        {
            int stage = 0;
            size_t option = 0;
            bool learned = false; // decided: the model's choice was faster
            hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
        };
        static constexpr size_t kMaxTrials = 1024;
        static uint64_t structure_key(const std::vector<ResolvedGate> &gates);
        void clear_trials();
        bool check_ = true;
        std::unordered_map<uint64_t, PlanTrial> trials_;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> *launch_log_ = nullptr;
        std::vector<hipEvent_t> ev_pool_;
        size_t ev_used_ = 0;
        bool forced_ = false;
        int forced_bits_ = 0, forced_k_ = 0, forced_tiles_ = kTilesAuto, forced_window_ = 64, forced_rule_ = 0;
        double stat_flushes_ = 0, stat_learned_ = 0, stat_pred_greedy_ = 0, stat_pred_chosen_ = 0;
    };
} // namespace qcore
