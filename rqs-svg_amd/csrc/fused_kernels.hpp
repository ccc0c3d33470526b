// Launch interface of the precision-form block kernel (fused_kernels.hip).
#pragma once
#include "sv_kernels.hpp"

namespace qsv
{
    // One dense block on kq qubits: off[e] is the amplitude offset of local basis state e
    // (bit j of e = the block's j-th qubit); R points to the (2^(kq+1))^2 real matrix on
    // interleaved (re, im) vectors, row-major, in device memory.  x3 = 0: R is FP32 and the
    // product runs on FP32 MFMA (kq = 5: R holds Re M, Im M and Re M + Im M instead, each
    // 32 x 32 row-major, for three real products in place of four); x3 = 1: R holds the
    // bfloat16 high parts of the D*D entries followed by their bfloat16 low parts (same byte
    // size), BF16x3 split-precision MFMA.
    struct BlockOp
    {
        int kq;
        int x3;
        uint64_t off[32];
        const float *R;
    };

    // Applies the block to every column of the subspace `sub` (block qubits plus any
    // controls, ascending; set_bits = positive controls).  Lanes read their own columns;
    // used when fewer than 5 qubits lie outside the block.
    void launch_block(hipStream_t s, c64 *state, int n, const Subspace &sub, const BlockOp &op);

    // LDS-staged form.  Tile bits TB[0..kq+4] = the block qubits plus the 5 lowest other
    // non-control qubits, ascending.  gw[b] = 1 << TB[b] (amplitude offset of tile bit b);
    // lw[b] = LDS index weight of tile bit b: 1 << j if TB[b] is the block's j-th qubit,
    // ((1 << kq) + 2) << i if TB[b] is the i-th of the 5 column
    // qubits.  tsub = tile bits plus controls (ascending), set_bits = positive controls.
    struct BlockLdsOp
    {
        int kq;
        int x3; // as in BlockOp
        uint64_t gw[10];
        int lw[10];
        const float *R;
    };
    void launch_block_lds(hipStream_t s, c64 *state, int n, const Subspace &tsub, const BlockLdsOp &op);

    // Stage: several blocks applied to LDS-resident tiles in one sweep.  The tile qubits T
    // (tsub.ops, ascending, m of them, T[0..3] = 0..3) are the stage's; gw[b] = 1 << T[b].
    // Block i's matrix is at R + blocks[i].matrix (layout of BlockOp for x3); w[j] (j < kq)
    // is the LDS index weight (1 << tile bit position) of the block's j-th qubit and
    // w[kq + j] that of the j-th column bit (the other tile bits, ascending).  blocks
    // points to device memory.  Complex64 stages also carry lds[j] = stage_lds_offset(w[j]),
    // the byte offset of that weight in the swizzled LDS tile, so that the kernels read the
    // offsets ready-made instead of swizzling them at every block.
    struct StageBlock
    {
        int w[16];
        uint32_t matrix; // offset in floats from the launch's matrix base (complex128 stages: doubles from M)
        uint32_t pad[3];
        uint32_t lds[16];
    };
    // LDS slot of tile index i in the complex64 stage kernels, as a byte offset (8-B slots).
    // Blocks gather their columns at strides 2^b (b = the tile bits below the block's lowest
    // qubit) and would hit one bank from all lanes; XOR-ing index bits 1..4 with bits 5..8
    // and 9..12 spreads such strides over the banks.  Bit 0 is kept, so amplitudes 2j and
    // 2j + 1 stay one aligned 16-B pair for the tile's global loads and stores.  The map is
    // linear over GF(2), so the offset of a | b (disjoint bits) is the XOR of theirs.
    constexpr uint32_t stage_lds_offset(uint32_t i) { return (i ^ ((((i >> 5) ^ (i >> 9)) & 15u) << 1)) << 3; }
    // One state's part of a stage launch (device memory, read by the kernel): its tiles are
    // the launch's tiles [tile_begin, tile_begin + 2^(n - m)), its nblocks StageBlock
    // descriptors start `blocks` floats after the launch's matrix base (a multiple of 4).
    struct alignas(16) StageTask
    {
        c64 *state;
        uint64_t tile_begin;
        uint64_t gw[16]; // global stride of tile bit b: 2^tile qubit b
        Subspace tsub;   // the tile qubits, ascending (tiles enumerate the other qubits)
        uint32_t blocks;
        int nblocks;
        uint32_t flags; // kStageZeroState: the state is |0...0>; the tiles are made, not loaded
        uint32_t trail; // kStageTrailTable: floats from the matrix base to a table of 2^m
                        // complex floats (tile index bit b = tile qubit b) applied at the store
    };
    constexpr uint32_t kStageZeroState = 1, kStageTrailTable = 2;
    // Host-side check of one task before upload (the kernel trusts the tasks).
    void check_stage_task(const StageTask &t, int n, int kq, int m);
    // tasks: device array of ntasks tasks with ascending tile_begin, the first 0, all checked
    // with check_stage_task; total_tiles = the tiles of all tasks.
    void launch_stage(hipStream_t s, int kq, int m, bool x3, const float *base, const StageTask *tasks, int ntasks, uint64_t total_tiles);
    // The same stage on one task by a persistent kernel (one workgroup per CU, two tile
    // buffers, the next tile loaded while the blocks run).  Same products as launch_stage,
    // so the same bits.  The task must not carry kStageZeroState.
    bool stage_stream_supported(int kq, int m, int n);
    void launch_stage_stream(hipStream_t s, int kq, int m, bool x3, const float *base, const StageTask *task, uint64_t total_tiles);
    // complex64 stage whose blocks are all diagonal: every amplitude times table[tile index],
    // the tile index made of the amplitude's bits at tsub.ops[0..m-1] (m = tsub.nops, 11..14,
    // ops[0..3] = 0..3); table holds 2^m complex floats in device memory; 2^n >= 2^13.
    void launch_stage_table64(hipStream_t s, c64 *state, int n, const Subspace &tsub, const float *table);

    // complex128 stage: tile of 2^12 amplitudes, blocks of kq = 3 or 4 qubits in FP64;
    // blocks[i].matrix is the offset in doubles of the 2^kq x 2^kq complex matrix (row-major,
    // interleaved) from M; w[0..kq-1] block bits, w[kq..11] column bits; blocks[i].pad[0] = 1
    // marks a diagonal matrix (only its diagonal is applied).
    // table != nullptr: every block of the stage is diagonal and table holds their product,
    // 2^12 complex values indexed by the tile index (bit b = tile qubit b); the kernel then
    // only multiplies each amplitude by its entry (no blocks, no LDS).
    // trail != nullptr: a table of the same form for diagonal blocks that follow the
    // nblocks blocks, applied to each amplitude as the tile is stored.
    struct Stage128Op
    {
        int m;
        int kq;
        int nblocks;
        uint64_t gw[16];
        const StageBlock *blocks;
        const double *M;
        const double *table;
        const double *trail;
    };
    constexpr uint32_t kStageBlockDiagonal = 1;
    void launch_stage128(hipStream_t s, c128 *state, int n, const Subspace &tsub, const Stage128Op &op);

    // complex128 block on 4 qubits in FP64.  M: 16x16 complex, row-major, interleaved
    // doubles, in device memory.  Staged form: gw/lw as for BlockLdsOp with 4 block bits and
    // the 6 lowest other qubits as column bits (LDS weight of column bit i = 17 << i);
    // direct form: off[e] = offset of local basis state e.
    struct Block128Op
    {
        uint64_t gw[10];
        int lw[10];
        uint64_t off[16];
        const double *M;
    };
    void launch_block128(hipStream_t s, c128 *state, int n, const Subspace &sub, const Block128Op &op, bool staged);
} // namespace qsv
