// Launch interface of the batching-form kernels (batch_kernels.hip).
#pragma once
#include "sv_kernels.hpp"

namespace qsv
{
    constexpr int kBatchMaxLdsQubits = 14; // 2^14 complex64 = 128 KiB of LDS

    // One 2x2 op of a batched circuit: m = {re00, im00, re01, im01, re10, im10, re11, im11},
    // acting on `target` when `control` (>= 0) is 1, or always (control < 0).
    struct BatchOp
    {
        int target;
        int control;
        float m[8];
    };

    struct BatchCircuit
    {
        int n;          // qubits
        int nops;       // ops
        int op_offset;  // first op in the op array
        int out_offset; // first <Z_q> slot in the output array
    };

    size_t batch_lds_bytes(int max_n);
    // Raise the dynamic-LDS limit of the persistent kernel (call once, outside graph capture).
    void batch_prepare(int block_size, int max_n);
    // Workgroups that fit at once (persistent grid).
    int batch_lds_grid(int block_size, int max_n);
    // Zero the work counter, then run the persistent kernel over all circuits.
    void launch_batch_lds(hipStream_t s, int block_size, int grid, int max_n, const BatchCircuit *circs, int ncirc, const BatchOp *ops, double *out, unsigned *counter);

    // <Z_q> of a complex64 state in HBM for q < n: partials needs zsum_blocks(n) * 32 doubles.
    int zsum_blocks(int n);
    void launch_zsum(hipStream_t s, const c64 *state, int n, double *partials, double *out);
    void launch_fill_zero_c64(hipStream_t s, c64 *state, uint64_t count);

    // <Z_q> of several complex64 states in two launches, with the same sums in the same order
    // as launch_zsum per state.  Task t owns workgroups [block_begin, block_begin + nblocks)
    // of the partial launch (nblocks = zsum_blocks(n)) and [qubit_begin, qubit_begin + n) of
    // the finish launch; part holds nblocks * 32 doubles, out n.  Tasks are in device memory
    // with both begins ascending from 0.
    struct alignas(16) ZsumTask
    {
        const c64 *state;
        double *part;
        double *out;
        uint32_t block_begin;
        uint32_t qubit_begin;
        int n;
        int nblocks;
    };
    void launch_zsum_tasks(hipStream_t s, const ZsumTask *tasks, int ntasks, uint32_t total_blocks, uint32_t total_qubits);
} // namespace qsv
