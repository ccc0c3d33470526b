// Host core of libqcs.so: one simulator = one state vector in HBM + one HIP stream.
#pragma once
#include "gates.hpp"
#include "sv_kernels.hpp"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <random>
#include <string>
#include <vector>

namespace qcore
{
    class FusedEngine;
    class CostModel;

    struct ResolvedGate
    {
        GateSpec spec;
        std::vector<int> tphys, pphys, nphys; // physical targets / positive / negative controls
        unsigned mmask = 0, mval = 0;         // measured targets (bit j = target j) and values
    };

    class Simulator
    {
    public:
        Simulator();
        ~Simulator();
        Simulator(const Simulator &) = delete;
        Simulator &operator=(const Simulator &) = delete;

        // configuration (before allocate)
        int num_qubits_cfg = 0;
        int num_clbits = 0;
        std::vector<bool> clbits;
        std::vector<int> initial_p2l;

        void allocate();
        void free_memory();
        bool allocated() const { return allocated_; }
        int num_qubits() const { return n_; }

        void reinitialize_mapping();
        void set_zero_state();
        void set_sequential_state();
        void set_flat_state();
        void set_entangled_state();
        void set_random_state();
        void clear_measurement_state();

        void operate(const GateSpec &g, std::vector<int> targets, std::vector<int> neg, std::vector<int> pos);
        int measure(int q);

        void save_statevector(const char *path);
        void get_statevector(double *out, uint64_t count);
        void set_statevector(const double *in, uint64_t count);
        void load_statevector(const char *path);
        void sample(uint64_t seed, uint64_t shots, uint64_t *out);
        // sample() with the prefix levels and outcomes in the library's scratch array: no
        // allocation; refuses queued fused gates (their flush may allocate)
        void sample_noalloc(uint64_t seed, uint64_t shots, uint64_t *out);
        double probability_total();
        double norm();
        void expect_z(double *out, int count);
        void set_local_qubits(int nl);
        void get_mapping(int16_t *out, int count) const;
        void permute(const std::vector<int> &want_p2l);
        void set_precision(int mode, int k);
        void set_stage_bits(int m);
        void set_stage_stream(int on);
        void set_planner_check(int on);
        void flush();
        void overlap(Simulator &other, double out[3]);
        void set_planner_model(CostModel *m);
        void set_plan(bool enable, int tile_bits, int k, int tiles, int window, int rule);
        void planner_stats(double out[5]);
        void planner_record(int count, uint64_t seed, float *features, float *log2_us);
        int planner_record_plans(int capacity, float *features, float *log2_us);
        void planner_last_choice(int out[5]);
        void fusion_stats(uint64_t *gates, uint64_t *blocks);

        int event_create();
        void event_record(int e);
        double event_elapsed(int a, int b);

        void set_seed(uint64_t s);
        void set_renormalize(bool on) { renormalize_ = on; }

        void profile_enable(bool on);
        void profile_get(int cls, uint64_t *launches, double *seconds, double *bytes);
        void profile_reset();

        hipStream_t stream() const { return stream_; }

    private:
        template <class F>
        void with_state(F &&f);
        void need_alloc() const;
        void check_operands(const std::vector<int> &t, const std::vector<int> &ng, const std::vector<int> &ps) const;
        void ensure_local(std::vector<int> &tphys);
        void swap_physical(int a, int b);
        void move_physical(const std::vector<int> &dst);
        void remap_physical(const std::vector<int> &dst);
        void apply_resolved(const ResolvedGate &r);
        void ensure_levels();
        uint64_t logical_chunk() const;
        void ensure_stage();

        // profiling: events around each launch
        struct Rec
        {
            int cls;
            double bytes;
            hipEvent_t a, b;
        };
        hipEvent_t pool_event();
        void prof_begin(int cls, double bytes);
        void prof_end();
        void prof_resolve();

        int device_ = 0;
        hipStream_t stream_ = nullptr;
        bool allocated_ = false;
        int n_ = 0;
        int n_local_ = 0;
        int precision_ = 0;
        int stage_bits_ = 13;
        int stage_stream_ = 1;
        int planner_check_ = 1;
        bool c64_state() const { return precision_ == 1 || precision_ == 3; }
        int block_qubits_ = 5;
        qsv::c128 *state_ = nullptr; // complex128 state (precision 0)
        std::vector<int> p2l_, l2p_;
        std::vector<int> m0_, m1_;
        uint64_t seed_ = 0;
        std::mt19937_64 engine_;
        bool renormalize_ = false;
        double *work_ = nullptr;
        uint64_t work_len_ = 0;
        std::vector<double *> levels_;
        // the sampling prefix levels (and their total) describe the current state; cleared
        // by every call that may change the state
        bool levels_valid_ = false;
        double levels_total_ = 0.0;
        void state_changed() { levels_valid_ = false; }
        double *pinned_ = nullptr; // 64 doubles of pinned host memory
        qsv::c128 *stage_dev_ = nullptr;
        void *stage_host_ = nullptr;
        uint64_t stage_len_ = 0;
        std::vector<hipEvent_t> events_;

        bool profile_ = false;
        std::vector<Rec> recs_;
        std::vector<hipEvent_t> free_events_;
        int open_cls_ = -1;
        double open_bytes_ = 0;
        hipEvent_t open_a_ = nullptr;
        uint64_t prof_launches_[16] = {};
        double prof_seconds_[16] = {};
        double prof_bytes_[16] = {};

        FusedEngine *fused_ = nullptr;
        friend class FusedEngine;
    };
} // namespace qcore
