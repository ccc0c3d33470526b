// Launch interface of the state-vector kernels (sv_kernels.hip).
//
// All functions enqueue work on the given stream and return immediately; hipError_t
// results are turned into std::runtime_error by the caller-side check in the launchers.
// Shapes are validated by the host core (simulator.cpp) before these are called; the
// launchers re-check what the kernels' indexing relies on.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace qsv
{
    constexpr int kMaxQubits = 40;

    // Sorted operand positions (controls and targets) of one gate and the bits that
    // positive controls force to 1.  A work item t maps to the amplitude index obtained by
    // inserting a 0 at each operand position of t (ascending) and OR-ing set_bits.
    struct Subspace
    {
        uint64_t set_bits;
        int nops;
        int ops[kMaxQubits];
    };

    struct c128
    {
        double re, im;
    };
    struct c64
    {
        float re, im;
    };

    // 2x2 on the amplitude pair (base|offA, base|offB): dense 1-qubit gates (offA = 0,
    // offB = 1<<t) and the swap-like 2-qubit gates.  zero_mask bit 0/1: input A/B reads as
    // zero (branch of a measured target that was not observed).
    struct PairOp
    {
        uint64_t offA, offB;
        c128 m[4]; // row-major m00 m01 m10 m11
        unsigned zero_mask;
    };

    // Diagonal on up to 2 targets over the control subspace: amplitude i gets
    // phase[bit(i,t0) | bit(i,t1)<<1].
    struct DiagOp
    {
        int ntargets;
        int t[2];
        c128 phase[4];
    };

    // Dense 4x4 on targets (t0,t1); rows with write_mask bit clear are left untouched.
    struct Dense4Op
    {
        uint64_t off[4];
        c128 m[16];
        unsigned write_mask;
        unsigned read_mask; // inputs that may be non-zero (measured-target rule)
    };

    // Up to 16 rows out[row[k]] = coef[k] * in[src[k]] on T <= 4 targets.
    struct MonoOp
    {
        int ntargets;
        int nrows;
        uint64_t off[16];
        int row[16];
        int src[16];
        c128 coef[16];
        unsigned read_mask;
    };

    enum class Launch
    {
        Pair,
        Phase,
        Diag,
        Dense4,
        Mono
    };

    // ---- gate kernels (complex128 state) ----
    void launch_pair(hipStream_t s, c128 *state, int n, const Subspace &sub, const PairOp &op);
    void launch_phase(hipStream_t s, c128 *state, int n, const Subspace &sub, c128 phase);
    void launch_diag(hipStream_t s, c128 *state, int n, const Subspace &sub, const DiagOp &op);
    void launch_dense4(hipStream_t s, c128 *state, int n, const Subspace &sub, const Dense4Op &op);
    void launch_mono(hipStream_t s, c128 *state, int n, const Subspace &sub, const MonoOp &op);

    // ---- state preparation (A = c128 or c64) ----
    template <class A> void launch_fill_basis(hipStream_t s, A *state, uint64_t count, uint64_t index); // |index>
    template <class A> void launch_fill_sequential(hipStream_t s, A *state, uint64_t count);
    template <class A> void launch_fill_flat(hipStream_t s, A *state, uint64_t count);
    template <class A> void launch_fill_random(hipStream_t s, A *state, uint64_t count, uint64_t seed);
    template <class A> void launch_set_one(hipStream_t s, A *state, uint64_t index);

    // ---- reductions (FP64, canonical order of include/qcs_ext.h) ----
    // Workspace size in doubles for reduce_* on `nitems` items.
    uint64_t reduce_workspace(uint64_t nitems);
    // {p0,p1} over the subspace `sub` (p1 at index | tbit); host out[2], stream synchronised.
    template <class A>
    void reduce_pair_probs(hipStream_t s, const A *state, const Subspace &sub, uint64_t tbit, uint64_t nitems, double *work, double out[2]);
    // canonical sum of p(i) (negated where bit `zbit` is set, if zbit != 0) over all amplitudes
    template <class A>
    void reduce_probs(hipStream_t s, const A *state, uint64_t count, uint64_t zbit, double *work, double *out);
    // out[0..1] = <a|b> (canonical tree per component), out[2] = max |a_i - b_i|
    uint64_t overlap_workspace(uint64_t count);
    template <class A, class B>
    void overlap(hipStream_t s, const A *a, const B *b, uint64_t count, double *work, double out[3]);

    // Scales the state to unit norm without a host round trip: the canonical FP64 sum of
    // p(i) lands in work[0] on the device, then every amplitude is multiplied by
    // 1/sqrt(work[0]) in FP64 (left as is when the sum is zero).
    template <class A>
    void renormalize_async(hipStream_t s, A *state, uint64_t count, double *work);

    // ---- measurement collapse ----
    template <class A> void launch_collapse(hipStream_t s, A *state, int n, int tq, int result, double scale);

    // ---- sampling ----
    // Level sizes of the 64-ary fold for 2^n entries; level 0 is the state itself.
    std::vector<uint64_t> fold_level_sizes(uint64_t count);
    template <class A>
    void build_fold_levels(hipStream_t s, const A *state, uint64_t count, const std::vector<double *> &levels);
    template <class A>
    void launch_sample_search(hipStream_t s, const A *state, uint64_t count, const std::vector<double *> &levels, double total, uint64_t seed, uint64_t nshots, const int *p2l, int n, uint64_t *out_dev);

    // ---- sampling without allocation ----
    // The prefix levels and the outcomes live in a file-scope device array of
    // kNoAllocScratchBytes, which each GPU holds from the library's load until the process
    // ends: states of at most kNoAllocMaxQubits qubits, at most kNoAllocMaxShots shots.
    constexpr int kMaxFoldLevels = 12; // 64-ary levels of up to 2^63 entries, level 0 included
    constexpr int kNoAllocMaxQubits = 30;
    constexpr uint64_t kNoAllocMaxShots = uint64_t(1) << 16;
    constexpr int kNoAllocMaxDevices = 64;
    // bytes of the fold levels of 2^n entries, each level's start aligned to 256 B
    constexpr uint64_t fold_level_bytes(int n)
    {
        uint64_t c = uint64_t(1) << n, total = 0;
        while (c > 1)
        {
            c = (c + 63) / 64;
            total += (c * 8 + 255) / 256 * 256;
        }
        return total;
    }
    constexpr uint64_t kNoAllocLevelBytes = fold_level_bytes(kNoAllocMaxQubits);
    constexpr uint64_t kNoAllocScratchBytes = kNoAllocLevelBytes + kNoAllocMaxShots * 8;
    // qcs_ext_sample's result on `state` (2^n amplitudes on `device`, queue `s`), built in
    // the scratch array: no device or pinned memory is allocated.  pinned: one double of
    // pinned host memory; out: nshots outcomes (host), logical indices.  Synchronises `s`;
    // returns the probability total.  Calls that overlap on one device race on the array.
    template <class A>
    double sample_noalloc(hipStream_t s, int device, const A *state, int n, uint64_t seed, uint64_t nshots, const int *p2l, double *pinned, uint64_t *out);

    // ---- qubit permutation (the in-HBM replacement of the reference's rank exchange) ----
    // Moves the data of every physical qubit q to physical qubit dst[q] in one sweep: the
    // amplitude at index i goes to the index whose bit dst[q] is bit q of i.  The moved
    // qubits (dst[q] != q) and the lowest unmoved ones form the tile bits (12 for
    // complex128, 13 for complex64: 64 KiB of LDS); a workgroup loads a tile, then writes
    // it back permuted.  At most permute_tile_bits(A) qubits may move (checked).
    template <class A> int permute_tile_bits();
    template <class A> void launch_permute(hipStream_t s, A *state, int n, const std::vector<int> &dst);
    // Splits a permutation (qubit q -> dst[q]) into passes for launch_permute, applied in
    // order, each moving at most `budget` qubits at or above `free_below` (qubits below it
    // are in every tile anyway).  budget >= 2.
    std::vector<std::vector<int>> plan_permute_passes(const std::vector<int> &dst, int budget, int free_below);

    // ---- layout: logical-order staging is always complex128 ----
    template <class A>
    void launch_gather_logical(hipStream_t s, const A *state, c128 *out, uint64_t first, uint64_t len, const int *l2p, int n);
    template <class A>
    void launch_scatter_logical(hipStream_t s, A *state, const c128 *in, uint64_t first, uint64_t len, const int *l2p, int n);

    void check(hipError_t e, const char *what);
} // namespace qsv
