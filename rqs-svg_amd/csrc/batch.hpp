// Batching form: many small variational circuits, one hipGraph launch.
#pragma once
#include "batch_kernels.hpp"
#include "fused.hpp"

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

namespace qcore
{
    // One gate of a batched circuit: a 1-target gate with at most one positive control.
    struct BatchGate
    {
        std::string name;
        int target;
        int control; // -1: none
        double params[4];
    };

    struct BatchTuning
    {
        int lds_max_qubits = 14; // circuits up to this size run in the LDS-resident kernel
        int block_size = 512;    // threads per workgroup of that kernel
        int block_qubits = 5;    // fused block size for the circuits in HBM
        int streams = 4;         // graph branches the HBM circuits are spread over
        int stage_bits = 12;     // HBM circuits of >= stage_bits + 4 qubits run as stages on
                                 // LDS tiles of 2^stage_bits amplitudes (0: one launch per block)
    };

    class Batch
    {
    public:
        Batch();
        ~Batch();
        Batch(const Batch &) = delete;
        Batch &operator=(const Batch &) = delete;

        int add_circuit(int n, const std::vector<BatchGate> &gates);
        int num_circuits() const { return (int)circuits_.size(); }
        int total_qubits() const { return total_qubits_; }
        void set_tuning(const BatchTuning &t);
        const BatchTuning &tuning() const { return tuning_; }
        // Block products of the circuits in HBM: 1 = FP32 MFMA (default), 3 = BF16x3
        // split-precision MFMA (three bf16 products, FP32 accumulation).  States stay
        // complex64 and <Z> FP64 either way; the LDS-resident circuits are unaffected.
        void set_precision(int mode);
        int precision() const { return x3_ ? 3 : 1; }
        // Runs every circuit from |0...0>, writes <Z_q> of circuit c at offset
        // sum_{c' < c} n_c' + q; returns seconds between graph launch and completion.
        double run(double *expz);
        // The same launches issued directly on the streams, without the graph (for tools that
        // do not follow graph launches); returns seconds between the first and last event.
        double run_direct(double *expz);
        // The plain method per circuit (complex128 simulators, one kernel per gate); returns
        // host wall seconds of the loop.
        double run_plain(double *expz);
        // Coordinate search over the tuning parameters on this batch; keeps the best.
        BatchTuning autotune();
        uint64_t graph_nodes() const { return graph_nodes_; }
        // algorithmic HBM bytes and kernel launches of one run (prepares the batch)
        double hbm_bytes()
        {
            prepare();
            return hbm_bytes_;
        }
        uint64_t kernel_launches()
        {
            prepare();
            return launches_;
        }

    private:
        struct Circuit
        {
            int n;
            std::vector<BatchGate> gates;
            int out_offset;
        };
        void prepare();
        std::function<void()> work_; // records the batch's launches on the streams
        void release();

        std::vector<Circuit> circuits_;
        int total_qubits_ = 0;
        BatchTuning tuning_;
        bool x3_ = false;
        bool prepared_ = false;
        int device_ = 0;
        std::vector<hipStream_t> streams_;
        std::vector<hipEvent_t> fork_join_;
        hipEvent_t t0_ = nullptr, t1_ = nullptr;
        hipGraph_t graph_ = nullptr;
        hipGraphExec_t exec_ = nullptr;
        uint64_t graph_nodes_ = 0;
        double hbm_bytes_ = 0;
        uint64_t launches_ = 0;
        // device buffers
        qsv::BatchCircuit *d_circs_ = nullptr;
        qsv::BatchOp *d_ops_ = nullptr;
        double *d_out_ = nullptr;
        unsigned *d_counter_ = nullptr;
        qsv::c64 *d_states_ = nullptr;
        double *d_partials_ = nullptr;
        float *d_mats_ = nullptr;
    };
} // namespace qcore
