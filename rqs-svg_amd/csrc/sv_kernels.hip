// State-vector kernels for gfx950 (MI355X), complex128 plain method.
//
// Layout: one contiguous array of 2^n amplitudes, interleaved (re, im) doubles, 16 B per
// amplitude; bit q of the array index is physical qubit q.
//
// Every gate kernel walks the "work items" of its gate: item t is mapped to an amplitude
// index by inserting a 0 bit at each sorted operand position and OR-ing the positive
// control bits (the same mapping as src/qcs.cu:156-376 of the reference, done here with
// kernel-argument operands instead of a __constant__ descriptor).  Consecutive lanes take
// consecutive items, so every operand above bit 2 leaves the access fully coalesced.
//
// Streaming choices measured on MI355X (in-place 1-qubit update of a 16 GiB state):
//  - 2 items per thread, 256-thread workgroups;
//  - 16-byte loads/stores, non-temporal except on states of 2^22-2^24 amplitudes
//    (64-256 MiB), which stay in the memory-side cache from one gate to the next;
//  - workgroup ids remapped so that each of the 8 XCDs streams one contiguous eighth of
//    the tiles of a launch (consecutive workgroups are dispatched round-robin over XCDs).
#include "sv_kernels.hpp"

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>

namespace qsv
{
    typedef double v2d __attribute__((ext_vector_type(2)));

    void check(hipError_t e, const char *what)
    {
        if (e != hipSuccess)
            throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
    }

    namespace
    {
        constexpr int kBlock = 256;
        constexpr uint64_t kMaxGrid = uint64_t(1) << 22;

        __device__ __forceinline__ v2d ld(const c128 *p) { return __builtin_nontemporal_load(reinterpret_cast<const v2d *>(p)); }
        __device__ __forceinline__ void st(c128 *p, v2d v) { __builtin_nontemporal_store(v, reinterpret_cast<v2d *>(p)); }

        __device__ __forceinline__ v2d cmul(c128 m, v2d a)
        {
            v2d r;
            r.x = m.re * a.x - m.im * a.y;
            r.y = m.re * a.y + m.im * a.x;
            return r;
        }
        __device__ __forceinline__ v2d cfma(c128 m, v2d a, v2d acc)
        {
            acc.x += m.re * a.x - m.im * a.y;
            acc.y += m.re * a.y + m.im * a.x;
            return acc;
        }

        __device__ __forceinline__ uint64_t expand(uint64_t t, const Subspace &sub)
        {
            for (int k = 0; k < sub.nops; ++k)
            {
                uint64_t lo = t & ((uint64_t(1) << sub.ops[k]) - 1);
                t = lo | ((t ^ lo) << 1);
            }
            return t | sub.set_bits;
        }

        // Workgroup id with the 8 XCDs each given a contiguous range.
        __device__ __forceinline__ uint64_t xcd_block()
        {
            uint64_t nb = gridDim.x, b = blockIdx.x;
            if ((nb & 7) == 0)
                b = (b & 7) * (nb >> 3) + (b >> 3);
            return b;
        }

        // p(i) = fma(re, re, im*im), the probability definition shared with the oracle.
        __device__ __forceinline__ double prob(v2d a) { return __builtin_fma(a.x, a.x, a.y * a.y); }

        // Orders this wave's LDS accesses against each other (lanes exchange data through LDS).
        __device__ __forceinline__ void wave_sync()
        {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }

        typedef float v2f __attribute__((ext_vector_type(2)));
        typedef float v4f __attribute__((ext_vector_type(4)));

        // Amplitude load/store in FP64 registers for both storage types; complex64 values are
        // widened exactly (float -> double) and rounded to nearest on the way back.
        __device__ __forceinline__ v2d lda(const c128 *p) { return ld(p); }
        __device__ __forceinline__ v2d lda(const c64 *p)
        {
            v2f f = __builtin_nontemporal_load(reinterpret_cast<const v2f *>(p));
            return v2d{double(f.x), double(f.y)};
        }
        __device__ __forceinline__ void sta(c128 *p, v2d v) { st(p, v); }
        __device__ __forceinline__ void sta(c64 *p, v2d v)
        {
            __builtin_nontemporal_store(v2f{float(v.x), float(v.y)}, reinterpret_cast<v2f *>(p));
        }
        __device__ __forceinline__ v2d ldca(const c128 *p) { return *reinterpret_cast<const v2d *>(p); }
        __device__ __forceinline__ v2d ldca(const c64 *p)
        {
            v2f f = *reinterpret_cast<const v2f *>(p);
            return v2d{double(f.x), double(f.y)};
        }
        // Gate-kernel amplitude access: cached on states of 64-256 MiB, which then stay in the
        // 256 MB memory-side cache from gate to gate (measured on 24 qubits, 256 MiB: pair
        // kernel 95 -> 77 us), non-temporal on larger states (faster on 30 qubits) and on
        // launch-bound small ones (see use_nt)
        template <bool NT>
        __device__ __forceinline__ v2d ldx(const c128 *p)
        {
            if constexpr (NT)
                return ld(p);
            else
                return *reinterpret_cast<const v2d *>(p);
        }
        template <bool NT>
        __device__ __forceinline__ void stx(c128 *p, v2d v)
        {
            if constexpr (NT)
                st(p, v);
            else
                *reinterpret_cast<v2d *>(p) = v;
        }

        template <int PER, bool NT>
        __global__ __launch_bounds__(kBlock) void k_pair(c128 *__restrict__ s, uint64_t nitems, Subspace sub, PairOp op)
        {
            const uint64_t per_tile = uint64_t(kBlock) * PER;
            const uint64_t ntiles = (nitems + per_tile - 1) / per_tile;
            for (uint64_t tile = xcd_block(); tile < ntiles; tile += gridDim.x)
            {
                uint64_t idx[PER];
                v2d a[PER], b[PER];
                bool ok[PER];
#pragma unroll
                for (int j = 0; j < PER; ++j)
                {
                    uint64_t item = tile * per_tile + uint64_t(j) * kBlock + threadIdx.x;
                    ok[j] = item < nitems;
                    idx[j] = ok[j] ? expand(item, sub) : 0;
                    a[j] = (ok[j] && !(op.zero_mask & 1)) ? ldx<NT>(s + (idx[j] | op.offA)) : v2d{0.0, 0.0};
                    b[j] = (ok[j] && !(op.zero_mask & 2)) ? ldx<NT>(s + (idx[j] | op.offB)) : v2d{0.0, 0.0};
                }
#pragma unroll
                for (int j = 0; j < PER; ++j)
                {
                    if (!ok[j])
                        continue;
                    v2d oa = cfma(op.m[1], b[j], cmul(op.m[0], a[j]));
                    v2d ob = cfma(op.m[3], b[j], cmul(op.m[2], a[j]));
                    stx<NT>(s + (idx[j] | op.offA), oa);
                    stx<NT>(s + (idx[j] | op.offB), ob);
                }
            }
        }

        // 2x2 on a low target t < 6 with every control at qubit 6 or above: a wave takes
        // 64 consecutive amplitudes (one per lane, fully coalesced) and each lane gets its
        // partner amplitude (lane ^ 2^t) through a cross-lane shuffle, instead of loading
        // pairs 2^t apart.  `csub` enumerates the 64-amplitude chunks: the controls' positions
        // minus 6 and their set bits shifted down by 6.
        template <int PER, bool NT>
        __global__ __launch_bounds__(kBlock) void k_pair_low(c128 *__restrict__ s, uint64_t nchunks, Subspace csub, int t, PairOp op)
        {
            constexpr int kWavesPerBlock = kBlock / 64;
            const int lane = threadIdx.x & 63;
            const int wave = threadIdx.x >> 6;
            const bool hi = (lane >> t) & 1;
            // this lane's amplitude reads as zero on a measured target's other branch
            const bool zero = hi ? (op.zero_mask & 2) : (op.zero_mask & 1);
            const uint64_t per_tile = uint64_t(kWavesPerBlock) * PER;
            const uint64_t ntiles = (nchunks + per_tile - 1) / per_tile;
            for (uint64_t tile = xcd_block(); tile < ntiles; tile += gridDim.x)
            {
                uint64_t idx[PER];
                v2d a[PER];
                bool ok[PER];
#pragma unroll
                for (int j = 0; j < PER; ++j)
                {
                    const uint64_t chunk = tile * per_tile + uint64_t(j) * kWavesPerBlock + wave;
                    ok[j] = chunk < nchunks; // uniform per wave
                    idx[j] = ok[j] ? (expand(chunk, csub) << 6) + lane : 0;
                    a[j] = (ok[j] && !zero) ? ldx<NT>(s + idx[j]) : v2d{0.0, 0.0};
                }
#pragma unroll
                for (int j = 0; j < PER; ++j)
                {
                    if (!ok[j])
                        continue;
                    const v2d p{__shfl_xor(a[j].x, 1 << t), __shfl_xor(a[j].y, 1 << t)};
                    // bit t clear: a is the |0> amplitude, p the |1> one; set: the reverse
                    const v2d out = hi ? cfma(op.m[3], a[j], cmul(op.m[2], p)) : cfma(op.m[1], p, cmul(op.m[0], a[j]));
                    stx<NT>(s + idx[j], out);
                }
            }
        }

        template <int PER, bool NT>
        __global__ __launch_bounds__(kBlock) void k_phase(c128 *__restrict__ s, uint64_t nitems, Subspace sub, c128 phase)
        {
            const uint64_t per_tile = uint64_t(kBlock) * PER;
            const uint64_t ntiles = (nitems + per_tile - 1) / per_tile;
            for (uint64_t tile = xcd_block(); tile < ntiles; tile += gridDim.x)
            {
                uint64_t idx[PER];
                v2d a[PER];
                bool ok[PER];
#pragma unroll
                for (int j = 0; j < PER; ++j)
                {
                    uint64_t item = tile * per_tile + uint64_t(j) * kBlock + threadIdx.x;
                    ok[j] = item < nitems;
                    idx[j] = ok[j] ? expand(item, sub) : 0;
                    a[j] = ok[j] ? ldx<NT>(s + idx[j]) : v2d{0.0, 0.0};
                }
#pragma unroll
                for (int j = 0; j < PER; ++j)
                    if (ok[j])
                        stx<NT>(s + idx[j], cmul(phase, a[j]));
            }
        }

        // Phase on a subspace with operand bits below 3 (amplitudes of interest 16-64 B
        // apart): a wave takes 64 consecutive amplitudes, and the lanes whose low 6 index bits
        // match the subspace's low pattern load, multiply and store; `csub` enumerates the
        // 64-amplitude chunks through the operands at 6 and above.  Each wave instruction
        // then covers 8 whole lines instead of spreading over 32-64.
        template <int PER, bool NT>
        __global__ __launch_bounds__(kBlock) void k_phase_low(c128 *__restrict__ s, uint64_t nchunks, Subspace csub, unsigned low_mask, unsigned low_set, c128 phase)
        {
            constexpr int kWavesPerBlock = kBlock / 64;
            const int lane = threadIdx.x & 63;
            const int wave = threadIdx.x >> 6;
            const bool mine = (unsigned(lane) & low_mask) == low_set;
            // an operand at qubit 0 or 1: every 64-B run holds changed and unchanged amplitudes,
            // and writing part of it costs more than writing it whole (measured on 30 qubits:
            // CZ(0,1) 8.2 -> 5.4 ms, CZ(0,7) 4.1 -> 2.7 ms), so all lanes load and store
            const bool all = (low_mask & 3u) != 0;
            const uint64_t per_tile = uint64_t(kWavesPerBlock) * PER;
            const uint64_t ntiles = (nchunks + per_tile - 1) / per_tile;
            for (uint64_t tile = xcd_block(); tile < ntiles; tile += gridDim.x)
            {
                uint64_t idx[PER];
                v2d a[PER];
                bool ok[PER];
#pragma unroll
                for (int j = 0; j < PER; ++j)
                {
                    const uint64_t chunk = tile * per_tile + uint64_t(j) * kWavesPerBlock + wave;
                    ok[j] = chunk < nchunks && (mine || all);
                    idx[j] = ok[j] ? (expand(chunk, csub) << 6) + lane : 0;
                    a[j] = ok[j] ? ldx<NT>(s + idx[j]) : v2d{0.0, 0.0};
                }
#pragma unroll
                for (int j = 0; j < PER; ++j)
                    if (ok[j])
                        stx<NT>(s + idx[j], mine ? cmul(phase, a[j]) : a[j]);
            }
        }

        template <int PER, bool NT>
        __global__ __launch_bounds__(kBlock) void k_diag(c128 *__restrict__ s, uint64_t nitems, Subspace sub, DiagOp op)
        {
            const uint64_t per_tile = uint64_t(kBlock) * PER;
            const uint64_t ntiles = (nitems + per_tile - 1) / per_tile;
            for (uint64_t tile = xcd_block(); tile < ntiles; tile += gridDim.x)
            {
                uint64_t idx[PER];
                v2d a[PER];
                bool ok[PER];
#pragma unroll
                for (int j = 0; j < PER; ++j)
                {
                    uint64_t item = tile * per_tile + uint64_t(j) * kBlock + threadIdx.x;
                    ok[j] = item < nitems;
                    idx[j] = ok[j] ? expand(item, sub) : 0;
                    a[j] = ok[j] ? ldx<NT>(s + idx[j]) : v2d{0.0, 0.0};
                }
#pragma unroll
                for (int j = 0; j < PER; ++j)
                {
                    if (!ok[j])
                        continue;
                    const bool b0 = (idx[j] >> op.t[0]) & 1;
                    const bool b1 = op.ntargets == 2 && ((idx[j] >> op.t[1]) & 1);
                    const c128 ph = b1 ? (b0 ? op.phase[3] : op.phase[2]) : (b0 ? op.phase[1] : op.phase[0]);
                    stx<NT>(s + idx[j], cmul(ph, a[j]));
                }
            }
        }

        __global__ __launch_bounds__(kBlock) void k_dense4(c128 *__restrict__ s, uint64_t nitems, Subspace sub, Dense4Op op)
        {
            const uint64_t ntiles = (nitems + kBlock - 1) / kBlock;
            for (uint64_t tile = xcd_block(); tile < ntiles; tile += gridDim.x)
            {
                uint64_t item = tile * kBlock + threadIdx.x;
                if (item >= nitems)
                    continue;
                uint64_t base = expand(item, sub);
                v2d in[4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    in[k] = ((op.read_mask >> k) & 1) ? ld(s + (base | op.off[k])) : v2d{0.0, 0.0};
#pragma unroll
                for (int r = 0; r < 4; ++r)
                {
                    if (!((op.write_mask >> r) & 1))
                        continue;
                    v2d acc = cmul(op.m[4 * r], in[0]);
                    acc = cfma(op.m[4 * r + 1], in[1], acc);
                    acc = cfma(op.m[4 * r + 2], in[2], acc);
                    acc = cfma(op.m[4 * r + 3], in[3], acc);
                    st(s + (base | op.off[r]), acc);
                }
            }
        }

        __global__ __launch_bounds__(kBlock) void k_mono(c128 *__restrict__ s, uint64_t nitems, Subspace sub, MonoOp op)
        {
            const uint64_t ntiles = (nitems + kBlock - 1) / kBlock;
            for (uint64_t tile = xcd_block(); tile < ntiles; tile += gridDim.x)
            {
                uint64_t item = tile * kBlock + threadIdx.x;
                if (item >= nitems)
                    continue;
                uint64_t base = expand(item, sub);
                v2d in[16];
#pragma unroll
                for (int k = 0; k < 16; ++k)
                {
                    in[k] = v2d{0.0, 0.0};
                    if (k < op.nrows && ((op.read_mask >> op.src[k]) & 1))
                        in[k] = ld(s + (base | op.off[op.src[k]]));
                }
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    if (k < op.nrows)
                        st(s + (base | op.off[op.row[k]]), cmul(op.coef[k], in[k]));
            }
        }

        // ---------------- gates with an operand on qubit 0, 1 or 2 ----------------
        // Amplitudes 1-4 apart (16-64 B) leave per-amplitude accesses covering parts of
        // lines.  Here a wave takes the 64 consecutive amplitudes of a chunk, one per lane
        // and per subset of the NH targets at qubit 6 or above (one register each), so every
        // load and store covers whole lines; the TL = T - NH other targets are lane bits.
        // The wave puts its amplitudes in LDS and each lane reads the inputs of its output
        // rows from there.  Lanes outside the pattern of the controls below qubit 6 store
        // their amplitudes unchanged (writing a line whole costs less than writing part of
        // it).  Local index of the gate: the lane-bit targets in bits 0..TL-1, the others
        // above them.
        struct LowOp
        {
            unsigned ctl_mask, ctl_set; // lane bits of the controls below qubit 6, their values
            unsigned tmask;             // lane bits of the targets below qubit 6
            unsigned read_mask;         // local inputs that may be non-zero (measured-target rule)
            unsigned keep_mask;         // local rows left unchanged
            int dep[16];                // lane bits of low local index kl < 2^TL
            uint64_t hoff[4];           // address offset of high subset h < 2^NH
            c128 m[16];                 // dense (T <= 2): row-major 2^T x 2^T
            int src[16];                // one nonzero per row: out[r] = coef[r] * in[src[r]]
            c128 coef[16];
        };

        constexpr int kLowPer = 2; // chunks per wave and tile

        template <int T, int NH, bool MONO, bool NT>
        __global__ __launch_bounds__(kBlock) void k_low(c128 *__restrict__ s, uint64_t nchunks, Subspace csub, LowOp op)
        {
            constexpr int TL = T - NH;
            constexpr int K = 1 << T;
            constexpr int H = 1 << NH;
            constexpr int W = kBlock / 64;
            static_assert(TL >= 0 && NH <= 2 && (MONO || T <= 2), "shape");
            __shared__ v2d buf[W][kLowPer][H][64];
            const int lane = threadIdx.x & 63;
            const int wave = threadIdx.x >> 6;
            const bool active = (unsigned(lane) & op.ctl_mask) == op.ctl_set;
            const int lbase = lane & ~int(op.tmask);
            int rl = 0; // this lane's low local index
#pragma unroll
            for (int i = 0; i < TL; ++i)
                rl |= ((lane & op.dep[1 << i]) != 0 ? 1 : 0) << i;
            // per register h: output row r = rl | h << TL, its coefficients and where its
            // inputs sit in LDS
            c128 cf[H][MONO ? 1 : K];
            int s_reg[H], s_lane[H];
            bool keep[H], readable[H];
#pragma unroll
            for (int h = 0; h < H; ++h)
            {
                const int r = rl | (h << TL);
                keep[h] = !active || ((op.keep_mask >> r) & 1);
                readable[h] = (op.read_mask >> r) & 1;
                if constexpr (MONO)
                {
                    const int sr = op.src[r];
                    cf[h][0] = op.coef[r];
                    s_reg[h] = sr >> TL;
                    s_lane[h] = lbase | op.dep[sr & ((1 << TL) - 1)];
                }
                else
                {
#pragma unroll
                    for (int k = 0; k < K; ++k)
                        cf[h][k] = op.m[r * K + k];
                    s_reg[h] = 0;
                    s_lane[h] = 0;
                }
            }
            const uint64_t per_tile = uint64_t(W) * kLowPer;
            const uint64_t ntiles = (nchunks + per_tile - 1) / per_tile;
            for (uint64_t tile = xcd_block(); tile < ntiles; tile += gridDim.x)
            {
                v2d x[kLowPer][H];
                uint64_t base[kLowPer];
                bool ok[kLowPer];
#pragma unroll
                for (int j = 0; j < kLowPer; ++j)
                {
                    const uint64_t chunk = tile * per_tile + uint64_t(j) * W + wave;
                    ok[j] = chunk < nchunks; // uniform per wave
                    base[j] = ok[j] ? (expand(chunk, csub) << 6) + lane : 0;
#pragma unroll
                    for (int h = 0; h < H; ++h)
                        x[j][h] = ok[j] ? ldx<NT>(s + (base[j] + op.hoff[h])) : v2d{0.0, 0.0};
                }
#pragma unroll
                for (int j = 0; j < kLowPer; ++j)
#pragma unroll
                    for (int h = 0; h < H; ++h)
                        buf[wave][j][h][lane] = readable[h] ? x[j][h] : v2d{0.0, 0.0};
                wave_sync();
#pragma unroll
                for (int j = 0; j < kLowPer; ++j)
                {
                    if (!ok[j])
                        continue;
#pragma unroll
                    for (int h = 0; h < H; ++h)
                    {
                        v2d out;
                        if constexpr (MONO)
                        {
                            out = cmul(cf[h][0], buf[wave][j][s_reg[h]][s_lane[h]]);
                        }
                        else
                        {
                            out = v2d{0.0, 0.0};
#pragma unroll
                            for (int k = 0; k < K; ++k)
                                out = cfma(cf[h][k], buf[wave][j][k >> TL][lbase | op.dep[k & ((1 << TL) - 1)]], out);
                        }
                        stx<NT>(s + (base[j] + op.hoff[h]), keep[h] ? x[j][h] : out);
                    }
                }
                // the next tile overwrites this wave's LDS rows
                wave_sync();
            }
        }

        // ---------------- state preparation ----------------
        template <class A>
        __global__ void k_fill_basis(A *s, uint64_t count, uint64_t index)
        {
            uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            for (; i < count; i += uint64_t(gridDim.x) * blockDim.x)
                sta(s + i, v2d{i == index ? 1.0 : 0.0, 0.0});
        }
        template <class A>
        __global__ void k_fill_sequential(A *s, uint64_t count)
        {
            uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            for (; i < count; i += uint64_t(gridDim.x) * blockDim.x)
                sta(s + i, v2d{double(i), 0.0});
        }
        template <class A>
        __global__ void k_fill_flat(A *s, uint64_t count)
        {
            uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            for (; i < count; i += uint64_t(gridDim.x) * blockDim.x)
                sta(s + i, v2d{1.0, 0.0});
        }
        __device__ __forceinline__ uint64_t splitmix64(uint64_t z)
        {
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        }
        __device__ __forceinline__ double uniform(uint64_t seed, uint64_t j)
        {
            return double(splitmix64(seed + (j + 1) * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
        }
        template <class A>
        __global__ void k_fill_random(A *s, uint64_t count, uint64_t seed)
        {
            uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            for (; i < count; i += uint64_t(gridDim.x) * blockDim.x)
            {
                double u1 = uniform(seed, 2 * i), u2 = uniform(seed, 2 * i + 1);
                double rad = sqrt(-2.0 * log(1.0 - u1));
                double ang = 2.0 * M_PI * u2;
                sta(s + i, v2d{rad * cos(ang), rad * sin(ang)});
            }
        }
        template <class A>
        __global__ void k_set_one(A *s, uint64_t index) { sta(s + index, v2d{1.0, 0.0}); }

        // ---------------- canonical reductions ----------------
        // Tile = 4096 items, 256 threads x 16 items at stride 256.  The register tree, the
        // two LDS steps and the wave shuffles together form the stride-halving tree over
        // the tile (strides 2048 .. 1).
        constexpr int kTile = 4096;

        template <int NV>
        __device__ __forceinline__ void tile_tree(double (&y)[NV][16], double *out_block, uint64_t out_stride)
        {
            __shared__ double buf[NV][kBlock];
#pragma unroll
            for (int v = 0; v < NV; ++v)
            {
#pragma unroll
                for (int st_ = 8; st_ >= 1; st_ /= 2)
#pragma unroll
                    for (int j = 0; j < st_; ++j)
                        y[v][j] = y[v][j] + y[v][j + st_];
                buf[v][threadIdx.x] = y[v][0];
            }
            __syncthreads();
            if (threadIdx.x < 128)
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    buf[v][threadIdx.x] = buf[v][threadIdx.x] + buf[v][threadIdx.x + 128];
            __syncthreads();
            if (threadIdx.x < 64)
            {
#pragma unroll
                for (int v = 0; v < NV; ++v)
                {
                    double x = buf[v][threadIdx.x] + buf[v][threadIdx.x + 64];
#pragma unroll
                    for (int sh = 32; sh >= 1; sh /= 2)
                        x = x + __shfl_down(x, sh, 64);
                    if (threadIdx.x == 0)
                        out_block[v * out_stride] = x;
                }
            }
        }

        template <class A>
        __global__ __launch_bounds__(kBlock) void k_reduce_pair(const A *__restrict__ s, uint64_t nitems, Subspace sub, uint64_t tbit, double *__restrict__ out, uint64_t out_stride)
        {
            double y[2][16];
            uint64_t base = uint64_t(blockIdx.x) * kTile + threadIdx.x;
#pragma unroll
            for (int j = 0; j < 16; ++j)
            {
                uint64_t item = base + uint64_t(j) * kBlock;
                if (item < nitems)
                {
                    uint64_t i0 = expand(item, sub);
                    y[0][j] = prob(lda(s + i0));
                    y[1][j] = prob(lda(s + (i0 | tbit)));
                }
                else
                {
                    y[0][j] = 0.0;
                    y[1][j] = 0.0;
                }
            }
            tile_tree<2>(y, out + blockIdx.x, out_stride);
        }

        // Reduction units: one amplitude (complex128), or an aligned pair of amplitudes
        // (complex64: 16 B per lane, the pair's two probabilities added first), so that every
        // lane load moves 16 B.
        template <class A>
        constexpr uint64_t reduce_per_unit() { return sizeof(A) == 8 ? 2 : 1; }

        template <class A>
        __global__ __launch_bounds__(kBlock) void k_reduce_probs(const A *__restrict__ s, uint64_t count, uint64_t zbit, double *__restrict__ out)
        {
            double y[1][16];
            uint64_t base = uint64_t(blockIdx.x) * kTile + threadIdx.x;
#pragma unroll
            for (int j = 0; j < 16; ++j)
            {
                uint64_t u = base + uint64_t(j) * kBlock;
                if constexpr (reduce_per_unit<A>() == 2)
                {
                    // amplitudes 2u and 2u + 1 (count is even: 2^n, n >= 1)
                    const uint64_t i = 2 * u;
                    double p = 0.0;
                    if (i < count)
                    {
                        const v4f f = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(s + i));
                        const double p0 = prob(v2d{double(f.x), double(f.y)}), p1 = prob(v2d{double(f.z), double(f.w)});
                        p = ((zbit && (i & zbit)) ? -p0 : p0) + ((zbit && ((i + 1) & zbit)) ? -p1 : p1);
                    }
                    y[0][j] = p;
                }
                else
                {
                    double p = u < count ? prob(lda(s + u)) : 0.0;
                    y[0][j] = (zbit && (u & zbit)) ? -p : p;
                }
            }
            tile_tree<1>(y, out + blockIdx.x, 0);
        }

        // in: NV arrays of length m at stride in_stride; out: NV arrays at out_stride.
        template <int NV>
        __global__ __launch_bounds__(kBlock) void k_reduce_tiles(const double *__restrict__ in, uint64_t m, uint64_t in_stride, double *__restrict__ out, uint64_t out_stride)
        {
            double y[NV][16];
            uint64_t base = uint64_t(blockIdx.x) * kTile + threadIdx.x;
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int j = 0; j < 16; ++j)
                {
                    uint64_t i = base + uint64_t(j) * kBlock;
                    y[v][j] = i < m ? in[v * in_stride + i] : 0.0;
                }
            tile_tree<NV>(y, out + blockIdx.x, out_stride);
        }

        // ---------------- collapse ----------------
        template <class A>
        __global__ __launch_bounds__(kBlock) void k_collapse(A *__restrict__ s, uint64_t half, int tq, int result, double scale, int rescale)
        {
            uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
            const uint64_t bit = uint64_t(1) << tq;
            for (; i < half; i += uint64_t(gridDim.x) * kBlock)
            {
                uint64_t lo = i & (bit - 1);
                uint64_t i0 = lo | ((i ^ lo) << 1);
                uint64_t kill = result ? i0 : (i0 | bit);
                uint64_t keep = result ? (i0 | bit) : i0;
                sta(s + kill, v2d{0.0, 0.0});
                if (rescale)
                {
                    v2d a = lda(s + keep);
                    sta(s + keep, v2d{a.x * scale, a.y * scale});
                }
            }
        }

        // Collapse on a qubit whose runs are shorter than a line (2^tq amplitudes of S bytes,
        // fewer than 128 B): every amplitude is read and written, zeroed or scaled by its bit,
        // so that all accesses cover whole lines.
        template <class A>
        __global__ __launch_bounds__(kBlock) void k_collapse_lines(A *__restrict__ s, uint64_t count, int tq, int result, double scale)
        {
            uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
            for (; i < count; i += uint64_t(gridDim.x) * kBlock)
            {
                const v2d a = lda(s + i);
                const bool kept = int((i >> tq) & 1) == result;
                sta(s + i, kept ? v2d{a.x * scale, a.y * scale} : v2d{0.0, 0.0});
            }
        }

        // multiply every amplitude by 1/sqrt(*sum) (the norm computed on the device)
        template <class A>
        __global__ __launch_bounds__(kBlock) void k_scale_to_unit(A *__restrict__ s, uint64_t count, const double *__restrict__ sum)
        {
            const double t = *sum;
            const double f = t > 0.0 ? 1.0 / sqrt(t) : 1.0;
            uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
            for (; i < count; i += uint64_t(gridDim.x) * kBlock)
            {
                const v2d a = lda(s + i);
                sta(s + i, v2d{a.x * f, a.y * f});
            }
        }

        // ---------------- sampling ----------------
        // One wave folds 64 consecutive blocks of 64 entries, each left to right: output
        // entry 64*g + L = (((x0 + x1) + x2) + ... + x63) over entries 4096*g + 64*L + 0..63.
        // The 4096 entries go through LDS in 4 chunks of 16 columns, so that every load
        // instruction covers runs of 16 consecutive entries (256 B of complex128 state) and
        // lane L then continues its own row's sum over the chunk in order.  The loads of the
        // next chunk are issued before the current chunk is summed, so each wave keeps one
        // chunk (16 KiB) in flight while it works.  4 waves per workgroup, 8.5 KiB of LDS
        // per wave.
#ifndef QSV_FOLD_WAVES
#define QSV_FOLD_WAVES 4
#endif
#ifndef QSV_FOLD_CHUNK
#define QSV_FOLD_CHUNK 16
#endif
        constexpr int kFoldWaves = QSV_FOLD_WAVES;
        constexpr int kFoldChunk = QSV_FOLD_CHUNK;   // columns per chunk (8 or 16)
        constexpr int kFoldChunks = 64 / kFoldChunk;
        constexpr int kFoldRow = kFoldChunk + 1; // doubles per LDS row (a chunk + 1 of padding)

        template <bool FROM_STATE, class A>
        __global__ __launch_bounds__(64 * kFoldWaves) void k_fold64(const void *__restrict__ src, uint64_t count, double *__restrict__ out, uint64_t nout, uint64_t ngroups)
        {
            __shared__ double tiles[kFoldWaves][64 * kFoldRow];
            const int lane = threadIdx.x & 63;
            const int wave = threadIdx.x >> 6;
            const uint64_t group = uint64_t(blockIdx.x) * kFoldWaves + wave;
            if (group >= ngroups) // whole wave leaves together; no workgroup barrier below
                return;
            double *tl = tiles[wave];
            // 8-byte entries (complex64 amplitudes, level sums) go two per lane so that each
            // load moves 16 B: lane = (row, column pair); 16-byte amplitudes one per lane
            constexpr bool PAIRS = !(FROM_STATE && sizeof(A) == 16);
            constexpr int NL = PAIRS ? kFoldChunk / 2 : kFoldChunk;    // 16-byte loads per lane and chunk
            constexpr int RUNS = PAIRS ? 128 / kFoldChunk : 64 / kFoldChunk; // rows covered by one load instruction
            constexpr int ES = FROM_STATE ? int(sizeof(A)) : 8; // bytes per entry
            const int run = PAIRS ? (lane / (kFoldChunk / 2)) : (lane / kFoldChunk);
            const int col = PAIRS ? 2 * (lane % (kFoldChunk / 2)) : (lane % kFoldChunk); // the lane's first column
            const char *bytes = static_cast<const char *>(src);
            // raw 16 bytes per load: one complex128, two complex64 or two level sums; entries
            // at or past `count` read as zero
            auto load = [&](int chunk, v2d(&r)[NL]) {
#pragma unroll
                for (int j = 0; j < NL; ++j)
                {
                    const uint64_t e = group * 4096 + uint64_t(RUNS * j + run) * 64 + kFoldChunk * chunk + col;
                    const v2d *p = reinterpret_cast<const v2d *>(bytes + e * ES);
                    if (e + (PAIRS ? 1 : 0) < count)
                        r[j] = FROM_STATE ? __builtin_nontemporal_load(p) : *p;
                    else if (PAIRS && e < count)
                    {
                        if (FROM_STATE)
                        {
                            const v2f f = *reinterpret_cast<const v2f *>(p);
                            r[j] = __builtin_bit_cast(v2d, v4f{f.x, f.y, 0.0f, 0.0f});
                        }
                        else
                            r[j] = v2d{*reinterpret_cast<const double *>(p), 0.0};
                    }
                    else
                        r[j] = v2d{0.0, 0.0};
                }
            };
            auto put = [&](const v2d(&r)[NL]) {
#pragma unroll
                for (int j = 0; j < NL; ++j)
                {
                    const int row = RUNS * j + run;
                    if (PAIRS)
                    {
                        double v0 = r[j].x, v1 = r[j].y;
                        if (FROM_STATE)
                        {
                            const v4f z = __builtin_bit_cast(v4f, r[j]);
                            v0 = prob(v2d{double(z.x), double(z.y)});
                            v1 = prob(v2d{double(z.z), double(z.w)});
                        }
                        tl[row * kFoldRow + col] = v0;
                        tl[row * kFoldRow + col + 1] = v1;
                    }
                    else
                    {
                        tl[row * kFoldRow + col] = prob(r[j]);
                    }
                }
            };
            v2d ra[NL], rb[NL];
            load(0, ra);
            double acc = 0.0;
#pragma unroll
            for (int chunk = 0; chunk < kFoldChunks; ++chunk)
            {
                v2d(&cur)[NL] = (chunk & 1) ? rb : ra;
                v2d(&nxt)[NL] = (chunk & 1) ? ra : rb;
                if (chunk < kFoldChunks - 1)
                    load(chunk + 1, nxt);
                put(cur);
                wave_sync();
#pragma unroll
                for (int c = 0; c < kFoldChunk; ++c)
                {
                    const double x = tl[lane * kFoldRow + c];
                    acc = (chunk == 0 && c == 0) ? x : acc + x;
                }
                wave_sync();
            }
            const uint64_t ob = group * 64 + lane;
            if (ob < nout)
                out[ob] = acc;
        }

        struct Levels
        {
            int nlev;               // number of levels including level 0 (the state)
            uint64_t size[kMaxFoldLevels];
            const double *ptr[kMaxFoldLevels]; // ptr[0] unused (state)
        };

        struct P2L
        {
            int n;
            int p2l[kMaxQubits];
        };

        template <class A>
        __device__ __forceinline__ double level_value(const A *s, const Levels &L, int lev, uint64_t j)
        {
            return lev == 0 ? prob(ldca(s + j)) : L.ptr[lev][j];
        }

        template <class A>
        __global__ void k_sample_search(const A *__restrict__ s, Levels L, double total, uint64_t seed, uint64_t nshots, P2L map, uint64_t *__restrict__ out)
        {
            uint64_t shot = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            if (shot >= nshots)
                return;
            const double r = uniform(seed, shot) * total;
            double offs[kMaxFoldLevels];
            int depth = 0;
            uint64_t block = 0;
            for (int lev = L.nlev - 2; lev >= 0; --lev)
            {
                uint64_t b0 = block * 64, b1 = L.size[lev] < b0 + 64 ? L.size[lev] : b0 + 64;
                double E = 0.0;
                uint64_t pick = b1 - 1;
                for (uint64_t j = b0; j < b1; ++j)
                {
                    double x = level_value(s, L, lev, j);
                    double incl = (j == b0) ? x : E + x;
                    double v = incl;
                    for (int d = depth - 1; d >= 0; --d)
                        v = offs[d] + v;
                    if (r <= v)
                    {
                        pick = j;
                        break;
                    }
                    E = incl;
                }
                double Ep = 0.0;
                for (uint64_t j = b0; j < pick; ++j)
                {
                    double x = level_value(s, L, lev, j);
                    Ep = (j == b0) ? x : Ep + x;
                }
                offs[depth++] = Ep;
                block = pick;
            }
            uint64_t logical = 0;
            for (int q = 0; q < map.n; ++q)
                logical |= ((block >> q) & 1) << map.p2l[q];
            out[shot] = logical;
        }

        // ---------------- qubit permutation ----------------
        // A tile is 2^m amplitudes whose index bits at the tile qubits T[0..m-1] vary (the
        // moved qubits plus the lowest unmoved ones, so T holds a run of low qubits and every
        // load and store instruction covers contiguous runs).  The workgroup loads its tile
        // into LDS in tile order, then writes output tile position y from tile position
        // x(y) = OR over b of bit b of y shifted to src[b], the input tile bit that lands at
        // output bit b.  Input and output use the same address map, so the tile is updated in
        // place and workgroups own disjoint addresses.
        constexpr int kPermThreads = 512;
#ifndef QSV_PERM_LDS
#define QSV_PERM_LDS 65536
#endif
        constexpr int kPermLds = QSV_PERM_LDS; // bytes

        struct PermuteOp
        {
            int m;           // tile bits
            int src[16];     // for output tile bit b: the input tile bit that moves there
            uint64_t gw[16]; // global stride of tile bit b
        };

        // LDS slot of tile position x: the entries of one 128-B LDS row (8 complex128 or 16
        // complex64) XOR-ed with every higher field of the same width, so that entries 2^h
        // apart (a moved low qubit) fall in different banks.
        template <class A>
        __device__ __forceinline__ int perm_slot(int x)
        {
            if constexpr (sizeof(A) == 16)
                return x ^ (((x >> 3) ^ (x >> 6) ^ (x >> 9)) & 7);
            else
                return x ^ (((x >> 4) ^ (x >> 8) ^ (x >> 12)) & 15);
        }

        template <class A>
        __device__ __forceinline__ A ld_amp(const A *p)
        {
            if constexpr (sizeof(A) == 16)
                return __builtin_bit_cast(A, __builtin_nontemporal_load(reinterpret_cast<const v2d *>(p)));
            else
                return __builtin_bit_cast(A, __builtin_nontemporal_load(reinterpret_cast<const v2f *>(p)));
        }
        template <class A>
        __device__ __forceinline__ void st_amp(A *p, A v)
        {
            if constexpr (sizeof(A) == 16)
                __builtin_nontemporal_store(__builtin_bit_cast(v2d, v), reinterpret_cast<v2d *>(p));
            else
                __builtin_nontemporal_store(__builtin_bit_cast(v2f, v), reinterpret_cast<v2f *>(p));
        }

        template <class A>
        __global__ __launch_bounds__(kPermThreads) void k_permute(A *__restrict__ s, uint64_t ntiles, Subspace tsub, PermuteOp op)
        {
            constexpr int CAP = kPermLds / int(sizeof(A));
            constexpr int PER = CAP / kPermThreads; // tile positions per thread at most
            __shared__ __attribute__((aligned(16))) A tile[CAP];
            const int nt = 1 << op.m;
            const int t = threadIdx.x;
            // the thread's low 9 position bits: global offset, and source position on output
            uint64_t g_t = 0;
            int x_t = 0;
            for (int b = 0; b < 9 && b < op.m; ++b)
                if ((t >> b) & 1)
                {
                    g_t += op.gw[b];
                    x_t |= 1 << op.src[b];
                }
            for (uint64_t tl = xcd_block(); tl < ntiles; tl += gridDim.x)
            {
                const uint64_t base = expand(tl, tsub);
                A a[PER];
#pragma unroll
                for (int i = 0; i < PER; ++i)
                {
                    const int x = i * kPermThreads + t;
                    if (x < nt)
                    {
                        uint64_t g = base + g_t;
                        for (int b = 9; b < op.m; ++b) // uniform
                            if (((i * kPermThreads) >> b) & 1)
                                g += op.gw[b];
                        a[i] = ld_amp(s + g);
                    }
                }
#pragma unroll
                for (int i = 0; i < PER; ++i)
                {
                    const int x = i * kPermThreads + t;
                    if (x < nt)
                        tile[perm_slot<A>(x)] = a[i];
                }
                __syncthreads();
#pragma unroll
                for (int i = 0; i < PER; ++i)
                {
                    const int y = i * kPermThreads + t;
                    if (y < nt)
                    {
                        uint64_t g = base + g_t;
                        int x = x_t;
                        for (int b = 9; b < op.m; ++b) // uniform
                            if (((i * kPermThreads) >> b) & 1)
                            {
                                g += op.gw[b];
                                x |= 1 << op.src[b];
                            }
                        st_amp(s + g, tile[perm_slot<A>(x)]);
                    }
                }
                __syncthreads(); // the next tile's positions are written to LDS
            }
        }

        // ---------------- layout ----------------
        struct L2P
        {
            int n;
            int l2p[kMaxQubits];
        };
        __device__ __forceinline__ uint64_t to_physical(uint64_t l, const L2P &m)
        {
            uint64_t p = 0;
            for (int q = 0; q < m.n; ++q)
                p |= ((l >> q) & 1) << m.l2p[q];
            return p;
        }
        template <class A>
        __global__ void k_gather_logical(const A *__restrict__ s, c128 *__restrict__ out, uint64_t first, uint64_t len, L2P m)
        {
            uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            for (; j < len; j += uint64_t(gridDim.x) * blockDim.x)
            {
                v2d a = ldca(s + to_physical(first + j, m));
                out[j] = c128{a.x, a.y};
            }
        }
        template <class A>
        __global__ void k_scatter_logical(A *__restrict__ s, const c128 *__restrict__ in, uint64_t first, uint64_t len, L2P m)
        {
            uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            for (; j < len; j += uint64_t(gridDim.x) * blockDim.x)
                sta(s + to_physical(first + j, m), v2d{in[j].re, in[j].im});
        }

        // <a|b> over two states of equal size, FP64, canonical tree per component; also the
        // largest |a_i - b_i| (as a max-reduction, order independent).
        template <class A, class B>
        __global__ __launch_bounds__(kBlock) void k_overlap(const A *__restrict__ a, const B *__restrict__ b, uint64_t count, double *__restrict__ out, uint64_t out_stride, double *__restrict__ maxdiff)
        {
            double y[2][16];
            double md = 0.0;
            uint64_t base = uint64_t(blockIdx.x) * kTile + threadIdx.x;
#pragma unroll
            for (int j = 0; j < 16; ++j)
            {
                uint64_t i = base + uint64_t(j) * kBlock;
                y[0][j] = 0.0;
                y[1][j] = 0.0;
                if (i < count)
                {
                    v2d x = ldca(a + i), z = ldca(b + i);
                    y[0][j] = x.x * z.x + x.y * z.y; // Re(conj(x) z)
                    y[1][j] = x.x * z.y - x.y * z.x; // Im(conj(x) z)
                    double dx = x.x - z.x, dy = x.y - z.y;
                    md = fmax(md, sqrt(dx * dx + dy * dy));
                }
            }
            tile_tree<2>(y, out + blockIdx.x, out_stride);
            // max over the workgroup with an LDS tree (max is exact in any order)
            __shared__ double mbuf[kBlock];
            mbuf[threadIdx.x] = md;
            __syncthreads();
            for (int st_ = kBlock / 2; st_ >= 1; st_ /= 2)
            {
                if (threadIdx.x < st_)
                    mbuf[threadIdx.x] = fmax(mbuf[threadIdx.x], mbuf[threadIdx.x + st_]);
                __syncthreads();
            }
            if (threadIdx.x == 0)
                maxdiff[blockIdx.x] = mbuf[0];
        }
        __global__ void k_max_tiles(const double *__restrict__ in, uint64_t m, double *__restrict__ out)
        {
            __shared__ double mbuf[kBlock];
            double md = 0.0;
            for (uint64_t i = threadIdx.x; i < m; i += kBlock)
                md = fmax(md, in[i]);
            mbuf[threadIdx.x] = md;
            __syncthreads();
            for (int st_ = kBlock / 2; st_ >= 1; st_ /= 2)
            {
                if (threadIdx.x < st_)
                    mbuf[threadIdx.x] = fmax(mbuf[threadIdx.x], mbuf[threadIdx.x + st_]);
                __syncthreads();
            }
            if (threadIdx.x == 0)
                out[0] = mbuf[0];
        }

        uint64_t grid_for(uint64_t items, uint64_t per_block)
        {
            uint64_t g = (items + per_block - 1) / per_block;
            g = std::max<uint64_t>(1, std::min(g, kMaxGrid));
            return g;
        }

        void after_launch(const char *what) { check(hipGetLastError(), what); }

        // complex128 states of 2^22 to 2^24 amplitudes (64-256 MiB) use cached accesses (see
        // ldx); at 2^20 (16 MiB) a launch-bound pair kernel measured 8.0 us non-temporal and
        // 8.9 us cached (the cached writes are written back at the end of the kernel)
        constexpr int kCachedMinQubits = 22, kCachedMaxQubits = 24;
        bool use_nt(int n) { return n > kCachedMaxQubits || n < kCachedMinQubits; }

        uint64_t items_of(int n, const Subspace &sub, const char *what)
        {
            if (sub.nops < 0 || sub.nops > n || n > kMaxQubits)
                throw std::runtime_error(std::string(what) + ": operand count out of range");
            for (int k = 0; k < sub.nops; ++k)
                if (sub.ops[k] < 0 || sub.ops[k] >= n || (k > 0 && sub.ops[k] <= sub.ops[k - 1]))
                    throw std::runtime_error(std::string(what) + ": operands must be ascending and < n");
            if (n < 64 && (sub.set_bits >> n) != 0)
                throw std::runtime_error(std::string(what) + ": control bit outside the state");
            return uint64_t(1) << (n - sub.nops);
        }

        void check_offset(int n, uint64_t off, const char *what)
        {
            if (n < 64 && (off >> n) != 0)
                throw std::runtime_error(std::string(what) + ": target offset outside the state");
        }
    } // namespace

    namespace
    {
        template <int T, int NH, bool MONO>
        void low_kernel(hipStream_t s, c128 *state, uint64_t nchunks, const Subspace &csub, const LowOp &op, bool nt)
        {
            const dim3 grid(grid_for(nchunks, (kBlock / 64) * kLowPer));
            if (nt)
                hipLaunchKernelGGL((k_low<T, NH, MONO, true>), grid, dim3(kBlock), 0, s, state, nchunks, csub, op);
            else
                hipLaunchKernelGGL((k_low<T, NH, MONO, false>), grid, dim3(kBlock), 0, s, state, nchunks, csub, op);
        }

        // The wave-chunk kernel (k_low) for a gate on T targets tq[0..T-1] (local bit j =
        // target j) whose operands (`sub`: targets and controls) include qubit 0, 1 or 2, on a
        // state of at least 12 qubits with at most 2 targets at qubit 6 or above.  dense:
        // 2^T x 2^T row-major (T <= 2), or null with the rows src/coef (T <= 4).  Returns
        // false when the gate does not have that shape.
        bool launch_low(hipStream_t s, c128 *state, int n, const Subspace &sub, int T, const int *tq, const c128 *dense, const int *src, const c128 *coef, unsigned read_mask, unsigned keep_mask)
        {
            if (n < 12 || sub.nops < 1 || sub.ops[0] >= 3 || T < 1 || T > 4 || (dense && T > 2))
                return false;
            int low[4], high[4], tl = 0, nh = 0;
            for (int j = 0; j < T; ++j)
                (tq[j] < 6 ? low[tl++] : high[nh++]) = j;
            if (nh > 2)
                return false;
            LowOp op{};
            int newbit[4];
            for (int i = 0; i < tl; ++i)
            {
                newbit[low[i]] = i;
                op.tmask |= 1u << tq[low[i]];
            }
            for (int i = 0; i < nh; ++i)
                newbit[high[i]] = tl + i;
            const int K = 1 << T;
            auto perm = [&](int k) {
                int r = 0;
                for (int j = 0; j < T; ++j)
                    r |= ((k >> j) & 1) << newbit[j];
                return r;
            };
            for (int kl = 0; kl < (1 << tl); ++kl)
                for (int i = 0; i < tl; ++i)
                    if ((kl >> i) & 1)
                        op.dep[kl] |= 1 << tq[low[i]];
            for (int h = 0; h < (1 << nh); ++h)
                for (int i = 0; i < nh; ++i)
                    if ((h >> i) & 1)
                        op.hoff[h] |= uint64_t(1) << tq[high[i]];
            // chunks: every operand at 6 or above is a chunk-index operand; controls below 6
            // are lane bits
            Subspace csub{};
            for (int k = 0; k < sub.nops; ++k)
            {
                const int q = sub.ops[k];
                bool is_target = false;
                for (int j = 0; j < T; ++j)
                    is_target |= tq[j] == q;
                const unsigned set = unsigned((sub.set_bits >> q) & 1);
                if (q < 6 && !is_target)
                {
                    op.ctl_mask |= 1u << q;
                    op.ctl_set |= set << q;
                }
                else if (q >= 6)
                {
                    csub.ops[csub.nops++] = q - 6;
                    if (!is_target)
                        csub.set_bits |= uint64_t(set) << (q - 6);
                }
            }
            if (n - 6 - csub.nops < 0)
                return false;
            for (int k = 0; k < K; ++k)
            {
                if ((read_mask >> k) & 1)
                    op.read_mask |= 1u << perm(k);
                if ((keep_mask >> k) & 1)
                    op.keep_mask |= 1u << perm(k);
                if (dense)
                {
                    for (int c = 0; c < K; ++c)
                        op.m[perm(k) * K + perm(c)] = dense[k * K + c];
                }
                else
                {
                    op.src[perm(k)] = perm(src[k]);
                    op.coef[perm(k)] = coef[k];
                }
            }
            const uint64_t nchunks = uint64_t(1) << (n - 6 - csub.nops);
            const bool nt = use_nt(n);
            const int key = (dense ? 0 : 100) + 10 * T + nh;
            switch (key)
            {
            case 10: low_kernel<1, 0, false>(s, state, nchunks, csub, op, nt); break;
            case 11: low_kernel<1, 1, false>(s, state, nchunks, csub, op, nt); break;
            case 20: low_kernel<2, 0, false>(s, state, nchunks, csub, op, nt); break;
            case 21: low_kernel<2, 1, false>(s, state, nchunks, csub, op, nt); break;
            case 22: low_kernel<2, 2, false>(s, state, nchunks, csub, op, nt); break;
            case 120: low_kernel<2, 0, true>(s, state, nchunks, csub, op, nt); break;
            case 121: low_kernel<2, 1, true>(s, state, nchunks, csub, op, nt); break;
            case 122: low_kernel<2, 2, true>(s, state, nchunks, csub, op, nt); break;
            case 130: low_kernel<3, 0, true>(s, state, nchunks, csub, op, nt); break;
            case 131: low_kernel<3, 1, true>(s, state, nchunks, csub, op, nt); break;
            case 132: low_kernel<3, 2, true>(s, state, nchunks, csub, op, nt); break;
            case 140: low_kernel<4, 0, true>(s, state, nchunks, csub, op, nt); break;
            case 141: low_kernel<4, 1, true>(s, state, nchunks, csub, op, nt); break;
            case 142: low_kernel<4, 2, true>(s, state, nchunks, csub, op, nt); break;
            default: return false;
            }
            after_launch("gate kernel (operands below qubit 3)");
            return true;
        }
    } // namespace

    void launch_pair(hipStream_t s, c128 *state, int n, const Subspace &sub, const PairOp &op)
    {
        uint64_t items = items_of(n, sub, "pair");
        check_offset(n, op.offA | op.offB, "pair");
        // one target t < 6 (the pattern |0>, |1> on bit t) and all controls at 6 or above
        int t = -1;
        if (op.offA == 0 && op.offB != 0 && (op.offB & (op.offB - 1)) == 0 && op.offB < 64 && n >= 12)
        {
            t = __builtin_ctzll(op.offB);
            Subspace csub{};
            csub.set_bits = sub.set_bits >> 6;
            bool fits = (sub.set_bits & 63) == 0;
            bool has_t = false;
            for (int k = 0; k < sub.nops && fits; ++k)
            {
                if (sub.ops[k] == t)
                    has_t = true;
                else if (sub.ops[k] < 6)
                    fits = false;
                else
                    csub.ops[csub.nops++] = sub.ops[k] - 6;
            }
            if (fits && has_t && n - 6 - csub.nops >= 0)
            {
                const uint64_t nchunks = uint64_t(1) << (n - 6 - csub.nops);
                if (use_nt(n))
                    hipLaunchKernelGGL((k_pair_low<2, true>), dim3(grid_for(nchunks, (kBlock / 64) * 2)), dim3(kBlock), 0, s, state, nchunks, csub, t, op);
                else
                    hipLaunchKernelGGL((k_pair_low<2, false>), dim3(grid_for(nchunks, (kBlock / 64) * 2)), dim3(kBlock), 0, s, state, nchunks, csub, t, op);
                after_launch("pair kernel (low target)");
                return;
            }
        }
        {
            // an operand below qubit 3 (a control, or a target with a control below 6): the
            // 2x2 on the patterns offA, offB as a gate on the targets, the bits of offA | offB
            const uint64_t tbits = op.offA | op.offB;
            const int T = __builtin_popcountll(tbits);
            if (T >= 1 && T <= 2)
            {
                int tq[2], nt_ = 0;
                for (uint64_t b = tbits; b; b &= b - 1)
                    tq[nt_++] = __builtin_ctzll(b);
                auto local = [&](uint64_t off) {
                    int k = 0;
                    for (int j = 0; j < T; ++j)
                        k |= int((off >> tq[j]) & 1) << j;
                    return k;
                };
                const int pa = local(op.offA), pb = local(op.offB), K = 1 << T;
                c128 m[16] = {};
                for (int k = 0; k < K; ++k)
                    m[k * K + k] = c128{1.0, 0.0};
                m[pa * K + pa] = op.m[0];
                m[pa * K + pb] = op.m[1];
                m[pb * K + pa] = op.m[2];
                m[pb * K + pb] = op.m[3];
                unsigned rmask = (1u << K) - 1, keep = (1u << K) - 1;
                keep &= ~((1u << pa) | (1u << pb));
                if (op.zero_mask & 1)
                    rmask &= ~(1u << pa);
                if (op.zero_mask & 2)
                    rmask &= ~(1u << pb);
                if (pa != pb && launch_low(s, state, n, sub, T, tq, m, nullptr, nullptr, rmask, keep))
                    return;
            }
        }
        const bool nt = use_nt(n);
        // states beyond the memory-side cache: one pair per thread (30 qubits, every target
        // position from 3 to 29: 3-4 % faster than two pairs per thread, e.g. target 20
        // 6.28 -> 6.05 ms, target 6 5.76 -> 5.52 ms)
        if (items >= (uint64_t(1) << 16) && nt && n > kCachedMaxQubits)
            hipLaunchKernelGGL((k_pair<1, true>), dim3(grid_for(items, kBlock)), dim3(kBlock), 0, s, state, items, sub, op);
        else if (items >= (uint64_t(1) << 16) && nt)
            hipLaunchKernelGGL((k_pair<2, true>), dim3(grid_for(items, kBlock * 2)), dim3(kBlock), 0, s, state, items, sub, op);
        else if (items >= (uint64_t(1) << 16))
            hipLaunchKernelGGL((k_pair<2, false>), dim3(grid_for(items, kBlock * 2)), dim3(kBlock), 0, s, state, items, sub, op);
        else if (nt)
            hipLaunchKernelGGL((k_pair<1, true>), dim3(grid_for(items, kBlock)), dim3(kBlock), 0, s, state, items, sub, op);
        else
            hipLaunchKernelGGL((k_pair<1, false>), dim3(grid_for(items, kBlock)), dim3(kBlock), 0, s, state, items, sub, op);
        after_launch("pair kernel");
    }

    void launch_phase(hipStream_t s, c128 *state, int n, const Subspace &sub, c128 phase)
    {
        uint64_t items = items_of(n, sub, "phase");
        // an operand below qubit 3: chunks of 64 amplitudes with a per-lane low-bit pattern
        if (n >= 12 && sub.nops > 0 && sub.ops[0] < 3)
        {
            Subspace csub{};
            csub.set_bits = sub.set_bits >> 6;
            unsigned low_mask = 0;
            for (int k = 0; k < sub.nops; ++k)
            {
                if (sub.ops[k] < 6)
                    low_mask |= 1u << sub.ops[k];
                else
                    csub.ops[csub.nops++] = sub.ops[k] - 6;
            }
            const unsigned low_set = unsigned(sub.set_bits & 63);
            if (n - 6 - csub.nops >= 0 && (low_set & ~low_mask) == 0)
            {
                const uint64_t nchunks = uint64_t(1) << (n - 6 - csub.nops);
                if (use_nt(n))
                    hipLaunchKernelGGL((k_phase_low<2, true>), dim3(grid_for(nchunks, (kBlock / 64) * 2)), dim3(kBlock), 0, s, state, nchunks, csub, low_mask, low_set, phase);
                else
                    hipLaunchKernelGGL((k_phase_low<2, false>), dim3(grid_for(nchunks, (kBlock / 64) * 2)), dim3(kBlock), 0, s, state, nchunks, csub, low_mask, low_set, phase);
                after_launch("phase kernel (low operands)");
                return;
            }
        }
        const bool nt = use_nt(n);
        if (items >= (uint64_t(1) << 16) && nt)
            hipLaunchKernelGGL((k_phase<2, true>), dim3(grid_for(items, kBlock * 2)), dim3(kBlock), 0, s, state, items, sub, phase);
        else if (items >= (uint64_t(1) << 16))
            hipLaunchKernelGGL((k_phase<2, false>), dim3(grid_for(items, kBlock * 2)), dim3(kBlock), 0, s, state, items, sub, phase);
        else if (nt)
            hipLaunchKernelGGL((k_phase<1, true>), dim3(grid_for(items, kBlock)), dim3(kBlock), 0, s, state, items, sub, phase);
        else
            hipLaunchKernelGGL((k_phase<1, false>), dim3(grid_for(items, kBlock)), dim3(kBlock), 0, s, state, items, sub, phase);
        after_launch("phase kernel");
    }

    void launch_diag(hipStream_t s, c128 *state, int n, const Subspace &sub, const DiagOp &op)
    {
        uint64_t items = items_of(n, sub, "diag");
        if (op.ntargets < 1 || op.ntargets > 2 || op.t[0] < 0 || op.t[0] >= n || (op.ntargets == 2 && (op.t[1] < 0 || op.t[1] >= n)))
            throw std::runtime_error("diag: bad targets");
        const bool nt = use_nt(n);
        if (items >= (uint64_t(1) << 16) && nt)
            hipLaunchKernelGGL((k_diag<2, true>), dim3(grid_for(items, kBlock * 2)), dim3(kBlock), 0, s, state, items, sub, op);
        else if (items >= (uint64_t(1) << 16))
            hipLaunchKernelGGL((k_diag<2, false>), dim3(grid_for(items, kBlock * 2)), dim3(kBlock), 0, s, state, items, sub, op);
        else if (nt)
            hipLaunchKernelGGL((k_diag<1, true>), dim3(grid_for(items, kBlock)), dim3(kBlock), 0, s, state, items, sub, op);
        else
            hipLaunchKernelGGL((k_diag<1, false>), dim3(grid_for(items, kBlock)), dim3(kBlock), 0, s, state, items, sub, op);
        after_launch("diag kernel");
    }

    void launch_dense4(hipStream_t s, c128 *state, int n, const Subspace &sub, const Dense4Op &op)
    {
        uint64_t items = items_of(n, sub, "dense4");
        for (int k = 0; k < 4; ++k)
            check_offset(n, op.off[k], "dense4");
        {
            const int tq[2] = {__builtin_ctzll(op.off[1] | (uint64_t(1) << 63)), __builtin_ctzll(op.off[2] | (uint64_t(1) << 63))};
            if (op.off[1] && op.off[2] && launch_low(s, state, n, sub, 2, tq, op.m, nullptr, nullptr, op.read_mask, ~op.write_mask & 15u))
                return;
        }
        hipLaunchKernelGGL(k_dense4, dim3(grid_for(items, kBlock)), dim3(kBlock), 0, s, state, items, sub, op);
        after_launch("dense4 kernel");
    }

    void launch_mono(hipStream_t s, c128 *state, int n, const Subspace &sub, const MonoOp &op)
    {
        uint64_t items = items_of(n, sub, "mono");
        if (op.ntargets < 1 || op.ntargets > 4 || op.nrows < 1 || op.nrows > 16)
            throw std::runtime_error("mono: bad shape");
        for (int k = 0; k < (1 << op.ntargets); ++k)
            check_offset(n, op.off[k], "mono");
        for (int k = 0; k < op.nrows; ++k)
            if (op.row[k] < 0 || op.row[k] >= (1 << op.ntargets) || op.src[k] < 0 || op.src[k] >= (1 << op.ntargets))
                throw std::runtime_error("mono: row index out of range");
        {
            // rows not listed are left unchanged
            const int K = 1 << op.ntargets;
            int tq[4], src[16];
            c128 coef[16];
            unsigned keep = (1u << K) - 1;
            bool ok = true;
            for (int j = 0; j < op.ntargets; ++j)
            {
                ok &= op.off[1 << j] != 0;
                tq[j] = __builtin_ctzll(op.off[1 << j] | (uint64_t(1) << 63));
            }
            for (int k = 0; k < K; ++k)
            {
                src[k] = k;
                coef[k] = c128{1.0, 0.0};
            }
            for (int k = 0; k < op.nrows; ++k)
            {
                src[op.row[k]] = op.src[k];
                coef[op.row[k]] = op.coef[k];
                keep &= ~(1u << op.row[k]);
            }
            if (ok && op.ntargets >= 2 && launch_low(s, state, n, sub, op.ntargets, tq, nullptr, src, coef, op.read_mask, keep))
                return;
        }
        hipLaunchKernelGGL(k_mono, dim3(grid_for(items, kBlock)), dim3(kBlock), 0, s, state, items, sub, op);
        after_launch("mono kernel");
    }

    // ---- support launchers for both amplitude types ----
    template <class A>
    void launch_fill_basis(hipStream_t s, A *state, uint64_t count, uint64_t index)
    {
        hipLaunchKernelGGL(k_fill_basis<A>, dim3(grid_for(count, kBlock)), dim3(kBlock), 0, s, state, count, index);
        after_launch("fill kernel");
    }
    template <class A>
    void launch_fill_sequential(hipStream_t s, A *state, uint64_t count)
    {
        hipLaunchKernelGGL(k_fill_sequential<A>, dim3(grid_for(count, kBlock)), dim3(kBlock), 0, s, state, count);
        after_launch("fill kernel");
    }
    template <class A>
    void launch_fill_flat(hipStream_t s, A *state, uint64_t count)
    {
        hipLaunchKernelGGL(k_fill_flat<A>, dim3(grid_for(count, kBlock)), dim3(kBlock), 0, s, state, count);
        after_launch("fill kernel");
    }
    template <class A>
    void launch_fill_random(hipStream_t s, A *state, uint64_t count, uint64_t seed)
    {
        hipLaunchKernelGGL(k_fill_random<A>, dim3(grid_for(count, kBlock)), dim3(kBlock), 0, s, state, count, seed);
        after_launch("fill kernel");
    }
    template <class A>
    void launch_set_one(hipStream_t s, A *state, uint64_t index)
    {
        hipLaunchKernelGGL(k_set_one<A>, dim3(1), dim3(1), 0, s, state, index);
        after_launch("set kernel");
    }

    uint64_t reduce_workspace(uint64_t nitems)
    {
        uint64_t total = 0, m = nitems;
        while (m > 1)
        {
            m = (m + kTile - 1) / kTile;
            total += m;
        }
        return 2 * (total + 2) + 2;
    }

    namespace
    {
        // Reduce NV arrays (length m each, at stride cur_stride) down to one value each,
        // written to dev_out[v].
        template <int NV>
        void finish_tiles(hipStream_t s, double *cur, uint64_t m, uint64_t cur_stride, double *next, double *dev_out)
        {
            while (m > 1)
            {
                uint64_t g = (m + kTile - 1) / kTile;
                uint64_t out_stride = g;
                double *dst = (g == 1) ? dev_out : next;
                if (g == 1)
                    out_stride = 1;
                hipLaunchKernelGGL(k_reduce_tiles<NV>, dim3(g), dim3(kBlock), 0, s, cur, m, cur_stride, dst, out_stride);
                after_launch("reduce kernel");
                cur = dst;
                cur_stride = out_stride;
                next = dst + NV * g;
                m = g;
            }
        }
    } // namespace

    template <class A>
    void reduce_pair_probs(hipStream_t s, const A *state, const Subspace &sub, uint64_t tbit, uint64_t nitems, double *work, double out[2])
    {
        // work layout: [result(2)] [level arrays ...]
        double *dev_out = work;
        uint64_t g = (nitems + kTile - 1) / kTile;
        double *lvl = work + 2;
        if (g == 1)
        {
            hipLaunchKernelGGL(k_reduce_pair<A>, dim3(1), dim3(kBlock), 0, s, state, nitems, sub, tbit, dev_out, uint64_t(1));
            after_launch("reduce kernel");
        }
        else
        {
            hipLaunchKernelGGL(k_reduce_pair<A>, dim3(g), dim3(kBlock), 0, s, state, nitems, sub, tbit, lvl, g);
            after_launch("reduce kernel");
            finish_tiles<2>(s, lvl, g, g, lvl + 2 * g, dev_out);
        }
        check(hipMemcpyAsync(out, dev_out, 2 * sizeof(double), hipMemcpyDeviceToHost, s), "reduce copy");
        check(hipStreamSynchronize(s), "reduce sync");
    }

    template <class A>
    void reduce_probs(hipStream_t s, const A *state, uint64_t count, uint64_t zbit, double *work, double *out)
    {
        double *dev_out = work;
        if (reduce_per_unit<A>() == 2 && (count & 1))
            throw std::runtime_error("reduce: complex64 states have an even amplitude count");
        uint64_t g = (count / reduce_per_unit<A>() + kTile - 1) / kTile;
        double *lvl = work + 2;
        if (g == 1)
        {
            hipLaunchKernelGGL(k_reduce_probs<A>, dim3(1), dim3(kBlock), 0, s, state, count, zbit, dev_out);
            after_launch("reduce kernel");
        }
        else
        {
            hipLaunchKernelGGL(k_reduce_probs<A>, dim3(g), dim3(kBlock), 0, s, state, count, zbit, lvl);
            after_launch("reduce kernel");
            finish_tiles<1>(s, lvl, g, g, lvl + g, dev_out);
        }
        check(hipMemcpyAsync(out, dev_out, sizeof(double), hipMemcpyDeviceToHost, s), "reduce copy");
        check(hipStreamSynchronize(s), "reduce sync");
    }

    template <class A>
    void renormalize_async(hipStream_t s, A *state, uint64_t count, double *work)
    {
        double *dev_out = work;
        if (reduce_per_unit<A>() == 2 && (count & 1))
            throw std::runtime_error("renormalize: complex64 states have an even amplitude count");
        const uint64_t g = (count / reduce_per_unit<A>() + kTile - 1) / kTile;
        double *lvl = work + 2;
        if (g == 1)
        {
            hipLaunchKernelGGL(k_reduce_probs<A>, dim3(1), dim3(kBlock), 0, s, state, count, uint64_t(0), dev_out);
            after_launch("reduce kernel");
        }
        else
        {
            hipLaunchKernelGGL(k_reduce_probs<A>, dim3(g), dim3(kBlock), 0, s, state, count, uint64_t(0), lvl);
            after_launch("reduce kernel");
            finish_tiles<1>(s, lvl, g, g, lvl + g, dev_out);
        }
        hipLaunchKernelGGL(k_scale_to_unit<A>, dim3(grid_for(count, kBlock * 4)), dim3(kBlock), 0, s, state, count, static_cast<const double *>(dev_out));
        after_launch("scale kernel");
    }

    template <class A, class B>
    void overlap(hipStream_t s, const A *a, const B *b, uint64_t count, double *work, double out[3])
    {
        // work: [result(2)] [maxdiff result(1)] [per-tile maxdiff (g)] [level arrays ...]
        uint64_t g = (count + kTile - 1) / kTile;
        double *dev_out = work;
        double *md_out = work + 2;
        double *md_tiles = work + 3;
        double *lvl = md_tiles + g;
        if (g == 1)
            hipLaunchKernelGGL((k_overlap<A, B>), dim3(1), dim3(kBlock), 0, s, a, b, count, dev_out, uint64_t(1), md_tiles);
        else
            hipLaunchKernelGGL((k_overlap<A, B>), dim3(g), dim3(kBlock), 0, s, a, b, count, lvl, g, md_tiles);
        after_launch("overlap kernel");
        if (g > 1)
            finish_tiles<2>(s, lvl, g, g, lvl + 2 * g, dev_out);
        hipLaunchKernelGGL(k_max_tiles, dim3(1), dim3(kBlock), 0, s, md_tiles, g, md_out);
        after_launch("max kernel");
        check(hipMemcpyAsync(out, dev_out, 3 * sizeof(double), hipMemcpyDeviceToHost, s), "overlap copy");
        check(hipStreamSynchronize(s), "overlap sync");
    }

    uint64_t overlap_workspace(uint64_t count)
    {
        uint64_t g = (count + kTile - 1) / kTile;
        return reduce_workspace(count) + g + 4;
    }

    template <class A>
    void launch_collapse(hipStream_t s, A *state, int n, int tq, int result, double scale)
    {
        if (tq < 0 || tq >= n || n < 1)
            throw std::runtime_error("collapse: qubit out of range");
        if ((uint64_t(sizeof(A)) << tq) < 128 && n >= 8)
        {
            // runs of 2^tq amplitudes shorter than a line: whole lines read and written
            const uint64_t count = uint64_t(1) << n;
            hipLaunchKernelGGL(k_collapse_lines<A>, dim3(grid_for(count, kBlock * 4)), dim3(kBlock), 0, s, state, count, tq, result, scale);
            after_launch("collapse kernel (whole lines)");
            return;
        }
        uint64_t half = uint64_t(1) << (n - 1);
        hipLaunchKernelGGL(k_collapse<A>, dim3(grid_for(half, kBlock)), dim3(kBlock), 0, s, state, half, tq, result, scale, scale != 1.0 ? 1 : 0);
        after_launch("collapse kernel");
    }

    std::vector<uint64_t> fold_level_sizes(uint64_t count)
    {
        std::vector<uint64_t> sz{count};
        while (sz.back() > 1)
            sz.push_back((sz.back() + 63) / 64);
        return sz;
    }

    namespace
    {
        // sz[0] = count, sz[l] = ceil(sz[l-1] / 64) down to 1; returns the number of sizes
        int fold_sizes(uint64_t count, uint64_t (&sz)[kMaxFoldLevels])
        {
            int ns = 0;
            sz[ns++] = count;
            while (sz[ns - 1] > 1)
            {
                if (ns == kMaxFoldLevels)
                    throw std::runtime_error("fold: too many levels");
                sz[ns] = (sz[ns - 1] + 63) / 64;
                ++ns;
            }
            return ns;
        }

        // levels[l - 1] receives level l of the fold (sizes sz[1 .. ns-1])
        template <class A>
        void fold_into(hipStream_t s, const A *state, const uint64_t *sz, int ns, double *const *levels)
        {
            for (int lev = 1; lev < ns; ++lev)
            {
                const uint64_t nout = sz[lev];
                const uint64_t groups = (nout + 63) / 64;
                const uint64_t blocks = (groups + kFoldWaves - 1) / kFoldWaves;
                if (lev == 1)
                    hipLaunchKernelGGL((k_fold64<true, A>), dim3(blocks), dim3(64 * kFoldWaves), 0, s, static_cast<const void *>(state), sz[0], levels[0], nout, groups);
                else
                    hipLaunchKernelGGL((k_fold64<false, A>), dim3(blocks), dim3(64 * kFoldWaves), 0, s, static_cast<const void *>(levels[lev - 2]), sz[lev - 1], levels[lev - 1], nout, groups);
                after_launch("fold kernel");
            }
        }

        template <class A>
        void search_into(hipStream_t s, const A *state, const uint64_t *sz, int ns, const double *const *levels, double total, uint64_t seed, uint64_t nshots, const int *p2l, int n, uint64_t *out_dev)
        {
            if (ns > kMaxFoldLevels || n > kMaxQubits)
                throw std::runtime_error("sample: bad level layout");
            Levels L{};
            L.nlev = ns;
            for (int i = 0; i < ns; ++i)
            {
                L.size[i] = sz[i];
                L.ptr[i] = i == 0 ? nullptr : levels[i - 1];
            }
            P2L m{};
            m.n = n;
            for (int q = 0; q < n; ++q)
                m.p2l[q] = p2l[q];
            if (nshots == 0)
                return;
            hipLaunchKernelGGL(k_sample_search<A>, dim3((nshots + 63) / 64), dim3(64), 0, s, state, L, total, seed, nshots, m, out_dev);
            after_launch("sample kernel");
        }
    } // namespace

    template <class A>
    void build_fold_levels(hipStream_t s, const A *state, uint64_t count, const std::vector<double *> &levels)
    {
        uint64_t sz[kMaxFoldLevels];
        const int ns = fold_sizes(count, sz);
        if ((int)levels.size() + 1 != ns)
            throw std::runtime_error("fold: level buffer count mismatch");
        fold_into(s, state, sz, ns, levels.data());
    }

    template <class A>
    void launch_sample_search(hipStream_t s, const A *state, uint64_t count, const std::vector<double *> &levels, double total, uint64_t seed, uint64_t nshots, const int *p2l, int n, uint64_t *out_dev)
    {
        uint64_t sz[kMaxFoldLevels];
        const int ns = fold_sizes(count, sz);
        if ((int)levels.size() + 1 != ns)
            throw std::runtime_error("sample: bad level layout");
        search_into(s, state, sz, ns, levels.data(), total, seed, nshots, p2l, n, out_dev);
    }

    // ---- sampling without allocation ----
    // The prefix levels (each level's start 256-B aligned) and then the outcomes.  Every
    // byte that a call reads it has written first in the same call (the fold writes each
    // level entry, the search each outcome), so nothing is cleared between calls.
    static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

    namespace
    {
        std::mutex scratch_mu;
        void *scratch_ptr[kNoAllocMaxDevices] = {};

        // this device's address of `scratch`, looked up once per device with it current
        unsigned char *scratch_on(int device)
        {
            if (device < 0 || device >= kNoAllocMaxDevices)
                throw std::runtime_error("sample_noalloc: device index out of range");
            std::lock_guard<std::mutex> lock(scratch_mu);
            if (!scratch_ptr[device])
            {
                int prev = -1;
                check(hipGetDevice(&prev), "hipGetDevice");
                if (prev != device)
                    check(hipSetDevice(device), "hipSetDevice");
                void *p = nullptr;
                const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
                if (prev != device)
                    check(hipSetDevice(prev), "hipSetDevice");
                check(e, "hipGetSymbolAddress(scratch)");
                if (!p)
                    throw std::runtime_error("hipGetSymbolAddress(scratch) returned NULL");
                scratch_ptr[device] = p;
            }
            return static_cast<unsigned char *>(scratch_ptr[device]);
        }
    } // namespace

    template <class A>
    double sample_noalloc(hipStream_t s, int device, const A *state, int n, uint64_t seed, uint64_t nshots, const int *p2l, double *pinned, uint64_t *out)
    {
        if (n < 1 || n > kNoAllocMaxQubits)
            throw std::runtime_error("sample_noalloc: " + std::to_string(n) + " qubits; the scratch array holds the prefix of at most " + std::to_string(kNoAllocMaxQubits));
        if (nshots > kNoAllocMaxShots)
            throw std::runtime_error("sample_noalloc: " + std::to_string(nshots) + " shots; the scratch array holds at most " + std::to_string(kNoAllocMaxShots));
        if (nshots > 0 && out == nullptr)
            throw std::runtime_error("outcome buffer must not be NULL");
        unsigned char *base = scratch_on(device);
        uint64_t sz[kMaxFoldLevels];
        const int ns = fold_sizes(uint64_t(1) << n, sz);
        double *levels[kMaxFoldLevels] = {};
        uint64_t off = 0;
        for (int lev = 1; lev < ns; ++lev)
        {
            levels[lev - 1] = reinterpret_cast<double *>(base + off);
            off += (sz[lev] * sizeof(double) + 255) / 256 * 256;
        }
        if (off > kNoAllocLevelBytes)
            throw std::runtime_error("sample_noalloc: prefix levels exceed the scratch array");
        uint64_t *outcomes = reinterpret_cast<uint64_t *>(base + kNoAllocLevelBytes);
        fold_into(s, state, sz, ns, levels);
        check(hipMemcpyAsync(pinned, levels[ns - 2], sizeof(double), hipMemcpyDeviceToHost, s), "copy");
        check(hipStreamSynchronize(s), "sync");
        const double total = pinned[0];
        if (nshots == 0)
            return total;
        search_into(s, state, sz, ns, levels, total, seed, nshots, p2l, n, outcomes);
        check(hipMemcpyAsync(out, outcomes, nshots * sizeof(uint64_t), hipMemcpyDeviceToHost, s), "copy");
        check(hipStreamSynchronize(s), "sync");
        return total;
    }

    template <class A>
    int permute_tile_bits() { return __builtin_ctz(unsigned(kPermLds / int(sizeof(A)))); }

    template <class A>
    void launch_permute(hipStream_t s, A *state, int n, const std::vector<int> &dst)
    {
        if (n < 1 || n > kMaxQubits || (int)dst.size() != n)
            throw std::runtime_error("permute: permutation size does not match the state");
        std::vector<int> seen(n, 0), moved;
        for (int q = 0; q < n; ++q)
        {
            if (dst[q] < 0 || dst[q] >= n || seen[dst[q]]++)
                throw std::runtime_error("permute: not a permutation of the qubits");
            if (dst[q] != q)
                moved.push_back(q);
        }
        if (moved.empty())
            return;
        const int m = std::min(n, permute_tile_bits<A>());
        if ((int)moved.size() > m)
            throw std::runtime_error("permute: more moved qubits than tile bits");
        // tile bits: the moved qubits and the lowest unmoved ones
        std::vector<int> tb = moved;
        for (int q = 0; (int)tb.size() < m; ++q)
            if (dst[q] == q)
                tb.push_back(q);
        std::sort(tb.begin(), tb.end());
        std::vector<int> pos(n, -1);
        for (int b = 0; b < m; ++b)
            pos[tb[b]] = b;
        PermuteOp op{};
        op.m = m;
        Subspace tsub{};
        tsub.nops = m;
        for (int b = 0; b < m; ++b)
        {
            op.gw[b] = uint64_t(1) << tb[b];
            tsub.ops[b] = tb[b];
            const int out = pos[dst[tb[b]]];
            if (out < 0)
                throw std::runtime_error("permute: tile bits not closed under the permutation");
            op.src[out] = b;
        }
        const uint64_t ntiles = uint64_t(1) << (n - m);
        hipLaunchKernelGGL(k_permute<A>, dim3(grid_for(ntiles, 1)), dim3(kPermThreads), 0, s, state, ntiles, tsub, op);
        after_launch("permute kernel");
    }

    std::vector<std::vector<int>> plan_permute_passes(const std::vector<int> &dst, int budget, int free_below)
    {
        const int n = (int)dst.size();
        if (budget < 2)
            throw std::runtime_error("permute: pass budget below 2");
        std::vector<int> seen(n, 0);
        for (int q = 0; q < n; ++q)
            if (dst[q] < 0 || dst[q] >= n || seen[dst[q]]++)
                throw std::runtime_error("permute: not a permutation of the qubits");
        auto cost = [&](int q) { return q >= free_below ? 1 : 0; };
        std::vector<int> rest = dst; // still to do: data now at q goes to rest[q]
        std::vector<std::vector<int>> passes;
        for (;;)
        {
            std::vector<int> pass(n);
            for (int q = 0; q < n; ++q)
                pass[q] = q;
            int used = 0;
            bool any = false;
            std::vector<char> done(n, 0);
            for (int q0 = 0; q0 < n; ++q0)
            {
                if (done[q0] || rest[q0] == q0)
                    continue;
                std::vector<int> cyc; // q0 -> rest[q0] -> ...
                for (int q = q0; !done[q]; q = rest[q])
                {
                    done[q] = 1;
                    cyc.push_back(q);
                }
                int c = 0;
                for (int q : cyc)
                    c += cost(q);
                if (used + c <= budget)
                {
                    for (int q : cyc)
                        pass[q] = rest[q];
                    used += c;
                    any = true;
                    continue;
                }
                // the longest prefix c1..cr that fits: c_i -> c_{i+1}, c_r -> c1; what
                // remains of the cycle is (c1 c_{r+1} ... c_L)
                int r = 0, cr = 0;
                while (r < (int)cyc.size() && used + cr + cost(cyc[r]) <= budget)
                    cr += cost(cyc[r++]);
                if (r < 2)
                    continue;
                for (int i = 0; i + 1 < r; ++i)
                    pass[cyc[i]] = cyc[i + 1];
                pass[cyc[r - 1]] = cyc[0];
                used += cr;
                any = true;
            }
            if (!any)
                break;
            std::vector<int> hit(n, 0);
            for (int q = 0; q < n; ++q)
                if (hit[pass[q]]++)
                    throw std::runtime_error("permute: a planned pass is not a permutation");
            if ((int)passes.size() >= n)
                throw std::runtime_error("permute: pass planning does not converge");
            // rest' o pass = rest: the data now at pass[q] still goes to rest[q]
            std::vector<int> next(n);
            for (int q = 0; q < n; ++q)
                next[pass[q]] = rest[q];
            rest = next;
            passes.push_back(pass);
        }
        for (int q = 0; q < n; ++q)
            if (rest[q] != q)
                throw std::runtime_error("permute: pass planning did not finish");
        return passes;
    }

    template <class A>
    void launch_gather_logical(hipStream_t s, const A *state, c128 *out, uint64_t first, uint64_t len, const int *l2p, int n)
    {
        L2P m{};
        m.n = n;
        for (int q = 0; q < n; ++q)
            m.l2p[q] = l2p[q];
        hipLaunchKernelGGL(k_gather_logical<A>, dim3(grid_for(len, kBlock)), dim3(kBlock), 0, s, state, out, first, len, m);
        after_launch("gather kernel");
    }

    template <class A>
    void launch_scatter_logical(hipStream_t s, A *state, const c128 *in, uint64_t first, uint64_t len, const int *l2p, int n)
    {
        L2P m{};
        m.n = n;
        for (int q = 0; q < n; ++q)
            m.l2p[q] = l2p[q];
        hipLaunchKernelGGL(k_scatter_logical<A>, dim3(grid_for(len, kBlock)), dim3(kBlock), 0, s, state, in, first, len, m);
        after_launch("scatter kernel");
    }

#define QSV_INSTANTIATE(A)                                                                                                                              \
    template void launch_fill_basis<A>(hipStream_t, A *, uint64_t, uint64_t);                                                                           \
    template void launch_fill_sequential<A>(hipStream_t, A *, uint64_t);                                                                                \
    template void launch_fill_flat<A>(hipStream_t, A *, uint64_t);                                                                                      \
    template void launch_fill_random<A>(hipStream_t, A *, uint64_t, uint64_t);                                                                          \
    template void launch_set_one<A>(hipStream_t, A *, uint64_t);                                                                                        \
    template void reduce_pair_probs<A>(hipStream_t, const A *, const Subspace &, uint64_t, uint64_t, double *, double *);                               \
    template void reduce_probs<A>(hipStream_t, const A *, uint64_t, uint64_t, double *, double *);                                                      \
    template void launch_collapse<A>(hipStream_t, A *, int, int, int, double);                                                                          \
    template void renormalize_async<A>(hipStream_t, A *, uint64_t, double *);                                                                           \
    template void build_fold_levels<A>(hipStream_t, const A *, uint64_t, const std::vector<double *> &);                                                \
    template void launch_sample_search<A>(hipStream_t, const A *, uint64_t, const std::vector<double *> &, double, uint64_t, uint64_t, const int *, int, uint64_t *); \
    template double sample_noalloc<A>(hipStream_t, int, const A *, int, uint64_t, uint64_t, const int *, double *, uint64_t *); \
    template void launch_gather_logical<A>(hipStream_t, const A *, c128 *, uint64_t, uint64_t, const int *, int);                                      \
    template void launch_scatter_logical<A>(hipStream_t, A *, const c128 *, uint64_t, uint64_t, const int *, int); \
    template int permute_tile_bits<A>();                                                                                                                \
    template void launch_permute<A>(hipStream_t, A *, int, const std::vector<int> &);

    QSV_INSTANTIATE(c128)
    QSV_INSTANTIATE(c64)
    template void overlap<c128, c128>(hipStream_t, const c128 *, const c128 *, uint64_t, double *, double *);
    template void overlap<c64, c128>(hipStream_t, const c64 *, const c128 *, uint64_t, double *, double *);
    template void overlap<c128, c64>(hipStream_t, const c128 *, const c64 *, uint64_t, double *, double *);
    template void overlap<c64, c64>(hipStream_t, const c64 *, const c64 *, uint64_t, double *, double *);
} // namespace qsv
