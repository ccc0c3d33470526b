// Launch interface of the precision-form block kernel (fused_kernels.hip).
#pragma once
#include "sv_kernels.hpp"

namespace qsv
{
    // One dense block on kq qubits: off[e] is the amplitude offset of local basis state e
    // (bit j of e = the block's j-th qubit); R points to the (2^(kq+1))^2 real matrix on
    // interleaved (re, im) vectors, row-major, in device memory.
    struct BlockOp
    {
        int kq;
        uint64_t off[32];
        const float *R;
    };

    // Applies the block to every column of the subspace `sub` (block qubits plus any
    // controls, ascending; set_bits = positive controls).
    void launch_block(hipStream_t s, c64 *state, int n, const Subspace &sub, const BlockOp &op);
} // namespace qsv
