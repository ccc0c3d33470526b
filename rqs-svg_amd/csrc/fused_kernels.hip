// Precision form: dense k-qubit blocks on a complex64 state, FP32 or BF16x3 split-precision
// MFMA (gfx950).
//
// A block acts on KQ qubits (KQ = 4 or 5).  Its 2^KQ x 2^KQ complex matrix M is given as
// the real D x D matrix R (D = 2^(KQ+1)) on interleaved (re, im) vectors:
//   R[2r][2c] = Re M[r][c], R[2r][2c+1] = -Im M[r][c], R[2r+1][2c] = Im M[r][c],
//   R[2r+1][2c+1] = Re M[r][c],
// so that y = R x for the interleaved float vector x of the 2^KQ amplitudes of one
// "column" (the amplitudes that differ only in the block's qubits).
//
// Work split: one wave takes 32 columns at a time.  The product R (D x D) * X (D x 32) runs
// as D/32 output tiles of v_mfma_f32_32x32x2f32 with the K dimension split in two halves:
// MFMA step s consumes k = s (lanes 0-31) and k = s + D/2 (lanes 32-63).  Lane (h, c)
// therefore needs the contiguous half h of column c -- 2^KQ/2 amplitudes loaded straight
// from HBM -- and keeps its rows of R in registers for the whole launch.  The D/32
// accumulators hold, per lane, pairs of consecutive amplitudes of column c, written back
// in place.  Accumulation is FP32.  The BF16x3 form (BlockMatrix<KQ, true>) runs the same
// K split in steps of 8 values per lane on the bf16 MFMA.  FP32 5-qubit blocks
// (BlockMatrix<5, false>) take M as Re M, Im M and Re M + Im M instead and form y with
// three real 32 x 32 products (see there).
#include "fused_kernels.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace qsv
{
    namespace
    {
        typedef float v2f __attribute__((ext_vector_type(2)));
        typedef float v4f __attribute__((ext_vector_type(4)));
        typedef float v16f __attribute__((ext_vector_type(16)));

        constexpr int kWaves = 4;
        constexpr int kThreads = 64 * kWaves;
// waves per stage workgroup by tile bits and arithmetic (FP32 or BF16x3 MFMA), measured on
// 32-qubit brick circuits
#ifndef QSV_STAGE_WAVES_14
#define QSV_STAGE_WAVES_14 4
#endif
#ifndef QSV_STAGE_WAVES_14_X3
#define QSV_STAGE_WAVES_14_X3 4
#endif
#ifndef QSV_STAGE_WAVES_13
#define QSV_STAGE_WAVES_13 4
#endif
#ifndef QSV_STAGE_WAVES_13_X3
#define QSV_STAGE_WAVES_13_X3 2
#endif
#ifndef QSV_STAGE_WAVES_12
#define QSV_STAGE_WAVES_12 2
#endif
#ifndef QSV_STAGE_WAVES_12_X3
#define QSV_STAGE_WAVES_12_X3 1
#endif
#ifndef QSV_STAGE_WAVES_11
#define QSV_STAGE_WAVES_11 1
#endif
#ifndef QSV_STAGE_WAVES_11_X3
#define QSV_STAGE_WAVES_11_X3 1
#endif
// BF16x3 stage blocks issue the next column groups' LDS gathers before the current groups'
// MFMA work, and take two column groups per pass with 5-qubit blocks (measured on 32-qubit
// brick stages of ~5 blocks: 27.3 -> 25.9 ms; FP32 stages gain nothing from either)
// streaming stage kernel: parts of a tile exchange, and parts issued per block
#ifndef QSV_STREAM_PARTS
#define QSV_STREAM_PARTS 4
#endif
#ifndef QSV_STREAM_PER_BLOCK
#define QSV_STREAM_PER_BLOCK 1
#endif
#ifndef QSV_STAGE_GATHER_AHEAD
#define QSV_STAGE_GATHER_AHEAD 1
#endif
#ifndef QSV_STAGE_GROUPS_KQ5_X3
#define QSV_STAGE_GROUPS_KQ5_X3 2
#endif
#ifndef QSV_STAGE_GROUPS_KQ5
#define QSV_STAGE_GROUPS_KQ5 1
#endif
#ifndef QSV_STREAM_GROUPS
#define QSV_STREAM_GROUPS 1 // column groups per pass in the streaming stage kernel (FP32)
#endif
#ifndef QSV_STREAM_GROUPS_X3
#define QSV_STREAM_GROUPS_X3 1 // the same, BF16x3
#endif
#ifndef QSV_STREAM_AHEAD_FP32
#define QSV_STREAM_AHEAD_FP32 1 // the streaming kernel gathers the next FP32 pass during the products too
#endif
#ifndef QSV_STAGE_GROUPS_KQ4
#define QSV_STAGE_GROUPS_KQ4 1
#endif
#ifndef QSV_BLOCK_PREFETCH
#define QSV_BLOCK_PREFETCH 1
#endif

        typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
        typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

        // The block matrix R as MFMA A-operands, held in registers for the whole launch, and
        // its product with a wave's 32 columns.  Lane (h, c) supplies rows 32*mt + c of R and
        // the half h of column c: x[0..H) are the K values h*H .. h*H + H - 1.
        template <int KQ, bool X3>
        struct BlockMatrix;

        // FP32 on v_mfma_f32_32x32x2f32: step st pairs K value st of lanes 0-31 with K value
        // st of lanes 32-63.
        template <int KQ>
        struct BlockMatrix<KQ, false>
        {
            static constexpr int D = 2 << KQ, MT = D / 32, H = D / 2;
            // accumulators per column group; output amplitude offset of lane half h
            static constexpr int NACC = MT, OUT_H = 2;
            // LDS copy (streaming stage kernel): 16-B unit j * 64 + lane holds what lane
            // `lane` loads as its j-th v4f, so each of a wave's UNITS reads is one
            // contiguous, conflict-free 1 KiB
            static constexpr int UNITS = MT * (H / 4);
            // the products of K values 8p .. 8p + 7 need part p: units 2p, 2p + 1 of each row tile
            static constexpr int PARTS = H / 8;
            float a[MT][H];
            // the 16-B unit of R (from its start) that LDS unit j * 64 + lane holds
            static __device__ __forceinline__ int source_unit(int j, int lane)
            {
                const int mt = j / (H / 4), q = j % (H / 4);
                return ((32 * mt + (lane & 31)) * D + (lane >> 5) * H) / 4 + q;
            }
            __device__ __forceinline__ void load_lds_part(const v4f *m, int lane, int p)
            {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int q = 2 * p; q < 2 * p + 2; ++q)
                    {
                        const v4f v = m[(mt * (H / 4) + q) * 64 + lane];
                        a[mt][4 * q + 0] = v.x;
                        a[mt][4 * q + 1] = v.y;
                        a[mt][4 * q + 2] = v.z;
                        a[mt][4 * q + 3] = v.w;
                    }
            }
            __device__ __forceinline__ void load_lds(const v4f *m, int lane)
            {
#pragma unroll
                for (int p = 0; p < PARTS; ++p)
                    load_lds_part(m, lane, p);
            }
            __device__ __forceinline__ void load(const float *R, int c, int h)
            {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                {
                    const v4f *rp = reinterpret_cast<const v4f *>(R + (32 * mt + c) * D + h * H);
#pragma unroll
                    for (int q = 0; q < H / 4; ++q)
                    {
                        const v4f v = rp[q];
                        a[mt][4 * q + 0] = v.x;
                        a[mt][4 * q + 1] = v.y;
                        a[mt][4 * q + 2] = v.z;
                        a[mt][4 * q + 3] = v.w;
                    }
                }
            }
            __device__ __forceinline__ void apply(const float (&x)[H], v16f (&d)[MT]) const
            {
#pragma unroll
                for (int st = 0; st < H; ++st)
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                        d[mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[mt][st], x[st], d[mt], 0, 0, 0);
            }
            // G column groups at once: G * MT independent accumulation chains; part p = the
            // products of K values 8p .. 8p + 7
            template <int G>
            __device__ __forceinline__ void apply_part(const float (&x)[G][H], v16f (&d)[G][MT], int p) const
            {
#pragma unroll
                for (int st = 8 * p; st < 8 * p + 8; ++st)
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                        for (int g = 0; g < G; ++g)
                            d[g][mt] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[mt][st], x[g][st], d[g][mt], 0, 0, 0);
            }
            template <int G>
            __device__ __forceinline__ void apply_n(const float (&x)[G][H], v16f (&d)[G][MT]) const
            {
#pragma unroll
                for (int p = 0; p < PARTS; ++p)
                    apply_part<G>(x, d, p);
            }
            // accumulator register r of tile mt holds real row 32*mt + 8*(r/4) + 4*h + r%4:
            // amplitudes e = 16*mt + 4*(r/4) (+ OUT_H*h) and e + 1 as {re, im, re, im}
            template <class W>
            static __device__ __forceinline__ void outputs2(const v16f (&d)[NACC], W &&w)
            {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        w(16 * mt + 4 * q, v4f{d[mt][4 * q + 0], d[mt][4 * q + 1], d[mt][4 * q + 2], d[mt][4 * q + 3]});
            }
        };

        // FP32 5-qubit blocks (32 x 32 complex) with three real products in place of four
        // (Gauss): P1 = Ar xr, P2 = Ai xi, P3 = (Ar + Ai)(xr + xi), then Re = P1 - P2 and
        // Im = P3 - P1 - P2, on v_mfma_f32_32x32x2f32: 48 products per 32 columns instead of
        // 64 (the form is MFMA-bound).  Step st pairs amplitude st of lanes 0-31 with amplitude
        // 16 + st of lanes 32-63, which are the lane half's x values 2st (re) and 2st + 1 (im).
        // R holds Ar, Ai and Ar + Ai, each 32 x 32 row-major FP32 (the sum formed in FP64).
        template <>
        struct BlockMatrix<5, false>
        {
            static constexpr int H = 32, NACC = 3, OUT_H = 4;
            // LDS copy: unit (4q + p) * 64 + lane holds the lane's Ar / Ai / As (q = 0, 1, 2)
            // entries 4p .. 4p + 3; part p = the products of steps 4p .. 4p + 3
            static constexpr int UNITS = 12, PARTS = 4;
            float a[3][16]; // [Ar, Ai, Ar + Ai][step]: row c, column 16h + step
            static __device__ __forceinline__ int source_unit(int j, int lane)
            {
                return (j / 4) * 256 + (lane & 31) * 8 + 4 * (lane >> 5) + j % 4;
            }
            __device__ __forceinline__ void load_lds_part(const v4f *m, int lane, int p)
            {
#pragma unroll
                for (int q = 0; q < 3; ++q)
                {
                    const v4f v = m[(4 * q + p) * 64 + lane];
                    a[q][4 * p + 0] = v.x;
                    a[q][4 * p + 1] = v.y;
                    a[q][4 * p + 2] = v.z;
                    a[q][4 * p + 3] = v.w;
                }
            }
            __device__ __forceinline__ void load_lds(const v4f *m, int lane)
            {
#pragma unroll
                for (int p = 0; p < PARTS; ++p)
                    load_lds_part(m, lane, p);
            }
            __device__ __forceinline__ void load(const float *R, int c, int h)
            {
#pragma unroll
                for (int q = 0; q < 3; ++q)
                {
                    const v4f *rp = reinterpret_cast<const v4f *>(R + q * 1024 + c * 32 + 16 * h);
#pragma unroll
                    for (int p = 0; p < 4; ++p)
                    {
                        const v4f v = rp[p];
                        a[q][4 * p + 0] = v.x;
                        a[q][4 * p + 1] = v.y;
                        a[q][4 * p + 2] = v.z;
                        a[q][4 * p + 3] = v.w;
                    }
                }
            }
            template <int G>
            __device__ __forceinline__ void apply_part(const float (&x)[G][H], v16f (&d)[G][NACC], int p) const
            {
#pragma unroll
                for (int st = 4 * p; st < 4 * p + 4; ++st)
#pragma unroll
                    for (int g = 0; g < G; ++g)
                    {
                        const float xr = x[g][2 * st], xi = x[g][2 * st + 1];
                        d[g][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0][st], xr, d[g][0], 0, 0, 0);
                        d[g][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[1][st], xi, d[g][1], 0, 0, 0);
                        d[g][2] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[2][st], xr + xi, d[g][2], 0, 0, 0);
                    }
            }
            template <int G>
            __device__ __forceinline__ void apply_n(const float (&x)[G][H], v16f (&d)[G][NACC]) const
            {
#pragma unroll
                for (int p = 0; p < PARTS; ++p)
                    apply_part<G>(x, d, p);
            }
            __device__ __forceinline__ void apply(const float (&x)[H], v16f (&d)[NACC]) const
            {
#pragma unroll
                for (int st = 0; st < 16; ++st)
                {
                    const float xr = x[2 * st], xi = x[2 * st + 1];
                    d[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[0][st], xr, d[0], 0, 0, 0);
                    d[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[1][st], xi, d[1], 0, 0, 0);
                    d[2] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[2][st], xr + xi, d[2], 0, 0, 0);
                }
            }
            // accumulator register r holds complex row 8*(r/4) + 4*h + r%4 of P1, P2, P3:
            // amplitudes e = 8*(r/4) + r%4 (r%4 = 0 or 2, + OUT_H*h) and e + 1
            template <class W>
            static __device__ __forceinline__ void outputs2(const v16f (&d)[NACC], W &&w)
            {
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int i = 0; i < 4; i += 2)
                    {
                        const int r = 4 * q + i;
                        const float p10 = d[0][r], p20 = d[1][r], p30 = d[2][r];
                        const float p11 = d[0][r + 1], p21 = d[1][r + 1], p31 = d[2][r + 1];
                        w(8 * q + i, v4f{p10 - p20, (p30 - p10) - p20, p11 - p21, (p31 - p11) - p21});
                    }
            }
        };

        // BF16x3 split precision on v_mfma_f32_32x32x16_bf16: R = Rh + Rl and x = xh + xl
        // with each part a bfloat16 rounded to nearest even (the two parts carry 16
        // significand bits), and R x ~ Rl xh + Rh xl + Rh xh accumulated in FP32.  The
        // dropped Rl xl term is about 2^-16 of each product.  Step st takes the K values
        // h*H + 8*st .. h*H + 8*st + 7 from lane half h.  R points to Rh (D*D bfloat16,
        // row-major) followed by Rl.
        template <int KQ>
        struct BlockMatrix<KQ, true>
        {
            static constexpr int D = 2 << KQ, MT = D / 32, H = D / 2, ST = H / 8;
            static constexpr int NACC = MT, OUT_H = 2; // as in the FP32 form
            // LDS copy, as in the FP32 form: the hi units, then the lo units
            static constexpr int UNITS = 2 * MT * ST;
            // the products of K step st (K values 8 st .. 8 st + 7) need part st
            static constexpr int PARTS = ST;
            bf16x8 hi[MT][ST], lo[MT][ST];
            static __device__ __forceinline__ int source_unit(int j, int lane)
            {
                const int part = j / (MT * ST), mt = (j % (MT * ST)) / ST, st = j % ST;
                return part * (D * D / 8) + ((32 * mt + (lane & 31)) * D + (lane >> 5) * H) / 8 + st;
            }
            __device__ __forceinline__ void load_lds_part(const v4f *m, int lane, int st)
            {
                const bf16x8 *u = reinterpret_cast<const bf16x8 *>(m);
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
                {
                    hi[mt][st] = u[(mt * ST + st) * 64 + lane];
                    lo[mt][st] = u[((MT + mt) * ST + st) * 64 + lane];
                }
            }
            __device__ __forceinline__ void load_lds(const v4f *m, int lane)
            {
#pragma unroll
                for (int st = 0; st < ST; ++st)
                    load_lds_part(m, lane, st);
            }
            __device__ __forceinline__ void load(const float *R, int c, int h)
            {
                const bf16x8 *rh = reinterpret_cast<const bf16x8 *>(R);
                const bf16x8 *rl = rh + D * D / 8;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int st = 0; st < ST; ++st)
                    {
                        const int v = ((32 * mt + c) * D + h * H) / 8 + st;
                        hi[mt][st] = rh[v];
                        lo[mt][st] = rl[v];
                    }
            }
            // x[8*st .. 8*st + 7] as bfloat16 parts xh + xl
            static __device__ __forceinline__ void split(const float *x, bf16x8 &xh, bf16x8 &xl)
            {
#pragma unroll
                for (int j = 0; j < 8; j += 2)
                {
                    // two values per conversion instruction
                    const v2f v = {x[j], x[j + 1]};
                    const bf16x2 hb = __builtin_convertvector(v, bf16x2);
                    const v2f hf = __builtin_convertvector(hb, v2f);
                    const v2f r = {v.x - hf.x, v.y - hf.y};
                    const bf16x2 lb = __builtin_convertvector(r, bf16x2);
                    xh[j] = hb.x;
                    xh[j + 1] = hb.y;
                    xl[j] = lb.x;
                    xl[j + 1] = lb.y;
                }
            }
            __device__ __forceinline__ void apply(const float (&x)[H], v16f (&d)[MT]) const
            {
#pragma unroll
                for (int st = 0; st < ST; ++st)
                {
                    bf16x8 xh, xl;
                    split(x + 8 * st, xh, xl);
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
                    {
                        d[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo[mt][st], xh, d[mt], 0, 0, 0);
                        d[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi[mt][st], xl, d[mt], 0, 0, 0);
                        d[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi[mt][st], xh, d[mt], 0, 0, 0);
                    }
                }
            }
            // G column groups at once: G * MT independent accumulation chains; part st = the
            // products of step st
            template <int G>
            __device__ __forceinline__ void apply_part(const float (&x)[G][H], v16f (&d)[G][MT], int st) const
            {
                {
                    bf16x8 xh[G], xl[G];
#pragma unroll
                    for (int g = 0; g < G; ++g)
                        split(x[g] + 8 * st, xh[g], xl[g]);
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                        for (int g = 0; g < G; ++g)
                            d[g][mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo[mt][st], xh[g], d[g][mt], 0, 0, 0);
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                        for (int g = 0; g < G; ++g)
                            d[g][mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi[mt][st], xl[g], d[g][mt], 0, 0, 0);
#pragma unroll
                    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                        for (int g = 0; g < G; ++g)
                            d[g][mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi[mt][st], xh[g], d[g][mt], 0, 0, 0);
                }
            }
            template <int G>
            __device__ __forceinline__ void apply_n(const float (&x)[G][H], v16f (&d)[G][MT]) const
            {
#pragma unroll
                for (int st = 0; st < ST; ++st)
                    apply_part<G>(x, d, st);
            }
            template <class W>
            static __device__ __forceinline__ void outputs2(const v16f (&d)[NACC], W &&w)
            {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        w(16 * mt + 4 * q, v4f{d[mt][4 * q + 0], d[mt][4 * q + 1], d[mt][4 * q + 2], d[mt][4 * q + 3]});
            }
        };

        __device__ __forceinline__ uint64_t expand(uint64_t t, const Subspace &sub)
        {
            for (int k = 0; k < sub.nops; ++k)
            {
                uint64_t lo = t & ((uint64_t(1) << sub.ops[k]) - 1);
                t = lo | ((t ^ lo) << 1);
            }
            return t | sub.set_bits;
        }

        template <int KQ, bool X3>
        __global__ __launch_bounds__(kThreads) void k_block(c64 *__restrict__ s, uint64_t ncols, Subspace sub, BlockOp op)
        {
            constexpr int D = 2 << KQ;   // real dimension
            constexpr int H = D / 2;     // K values per lane
            constexpr int AMPS = H / 2;  // amplitudes per lane

            const int lane = threadIdx.x & 63;
            const int c = lane & 31;
            const int h = lane >> 5;

            // rows 32*mt + c of R, columns [h*H, h*H + H): the A operands of every step
            BlockMatrix<KQ, X3> A;
            A.load(op.R, c, h);

            const uint64_t ntiles = (ncols + 31) / 32;
            const uint64_t nwaves = uint64_t(gridDim.x) * kWaves;
            for (uint64_t tile = uint64_t(blockIdx.x) * kWaves + (threadIdx.x >> 6); tile < ntiles; tile += nwaves)
            {
                const uint64_t col = tile * 32 + c;
                const bool ok = col < ncols;
                const uint64_t base = ok ? expand(col, sub) : 0;
                float b[H];
#pragma unroll
                for (int j = 0; j < AMPS; ++j)
                {
                    const uint64_t off = h ? op.off[AMPS + j] : op.off[j];
                    v2f v = ok ? __builtin_nontemporal_load(reinterpret_cast<const v2f *>(s + (base | off))) : v2f{0.f, 0.f};
                    b[2 * j] = v.x;
                    b[2 * j + 1] = v.y;
                }
                typedef BlockMatrix<KQ, X3> Mat;
                v16f d[Mat::NACC];
#pragma unroll
                for (int i = 0; i < Mat::NACC; ++i)
                    d[i] = v16f{0.f};
                A.apply(b, d);
                if (!ok)
                    continue;
                Mat::outputs2(d, [&](int e, v4f v) {
                    const uint64_t o0 = h ? op.off[e + Mat::OUT_H] : op.off[e];
                    const uint64_t o1 = h ? op.off[e + Mat::OUT_H + 1] : op.off[e + 1];
                    __builtin_nontemporal_store(v2f{v.x, v.y}, reinterpret_cast<v2f *>(s + (base | o0)));
                    __builtin_nontemporal_store(v2f{v.z, v.w}, reinterpret_cast<v2f *>(s + (base | o1)));
                });
            }
        }
        // LDS slot of tile index i in the complex64 stage kernels (stage_lds_offset in slots).
        // The map is linear over GF(2) (it XORs bits 1..4 with bits of i it leaves unchanged),
        // so for indices built from disjoint bit sets, swz64(a | b) = swz64(a) ^ swz64(b): the
        // stage kernel combines per-lane and uniform parts of an LDS address with one XOR.
        __device__ __forceinline__ int swz64(int i) { return int(stage_lds_offset(uint32_t(i)) >> 3); }
        // complex128 (16-B slots): XOR bits 0..3 with bits 4..7 and 8..11
        __device__ __forceinline__ int swz128(int i) { return i ^ (((i >> 4) ^ (i >> 8)) & 15); }

        // Read-only data of a launch (block descriptors, complex128 matrices) read through the
        // constant address space: uniform loads from it become scalar loads, which neither
        // occupy vector registers nor wait behind the vector loads of the tile prefetch.
        template <class T>
        __device__ __forceinline__ const __attribute__((address_space(4))) T *as_constant(const T *p)
        {
            return (const __attribute__((address_space(4))) T *)(p);
        }

        // Workgroup barrier that orders LDS accesses only.  __syncthreads() also waits for
        // every outstanding global load and store of the wave, which would stall on the
        // next tile's prefetch; the stage kernels exchange data between waves only through
        // LDS, and their global accesses touch disjoint tiles.
        __device__ __forceinline__ void lds_barrier()
        {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        }

        // Orders this wave's LDS accesses against each other (lanes exchange data through LDS).
        __device__ __forceinline__ void wave_sync()
        {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }

        // LDS-staged variant.  A wave's tile is the 2^KQ * 32 amplitudes whose index bits
        // outside the "tile bits" (the block qubits plus the 5 lowest other qubits) are
        // fixed.  Position p of the tile (bit b of p <-> tile bit TB[b], ascending) is
        // moved in address order -- lanes take consecutive positions, so each load/store
        // instruction covers contiguous runs -- and transposed through LDS into
        // [column][local amplitude] rows for the MFMA operands and back.
        template <int KQ, bool X3>
        __global__ __launch_bounds__(kThreads) void k_block_lds(c64 *__restrict__ s, uint64_t ntiles, Subspace tsub, BlockLdsOp op)
        {
            constexpr int D = 2 << KQ;
            constexpr int H = D / 2;
            constexpr int AMPS = H / 2;
            constexpr int K = 1 << KQ;
            constexpr int ROW = K + 2;           // complex per LDS row (padding against bank conflicts)
            constexpr int NPOS = 32 * K;         // tile positions
            constexpr int NJ = NPOS / 64;        // load/store instructions per lane
            __shared__ __attribute__((aligned(16))) v2f lds[kWaves][32 * ROW];

            const int lane = threadIdx.x & 63;
            const int wave = threadIdx.x >> 6;
            const int c = lane & 31;
            const int h = lane >> 5;
            v2f *tl = lds[wave];

            BlockMatrix<KQ, X3> A;
            A.load(op.R, c, h);
            // lane part of the position -> global offset and LDS index (tile bits 0..5)
            uint64_t g_lane = 0;
            int l_lane = 0;
#pragma unroll
            for (int b = 0; b < 6; ++b)
                if ((lane >> b) & 1)
                {
                    g_lane += op.gw[b];
                    l_lane += op.lw[b];
                }

            const uint64_t nwaves = uint64_t(gridDim.x) * kWaves;
            // The next tile's loads are issued before this tile's MFMA work and stores, so
            // each wave keeps a tile of reads in flight while it computes.
            auto load_tile = [&](uint64_t tile, v2f (&v)[NJ]) {
                const uint64_t base = expand(tile, tsub) + g_lane;
#pragma unroll
                for (int j = 0; j < NJ; ++j)
                {
                    uint64_t gj = 0;
#pragma unroll
                    for (int b = 6; b < 6 + 4; ++b)
                        if ((j >> (b - 6)) & 1)
                            gj += op.gw[b];
                    v[j] = __builtin_nontemporal_load(reinterpret_cast<const v2f *>(s + (base + gj)));
                }
            };
            v2f v[NJ];
            uint64_t tile = uint64_t(blockIdx.x) * kWaves + wave;
            if (tile < ntiles)
                load_tile(tile, v);
            for (; tile < ntiles; tile += nwaves)
            {
                const uint64_t base = expand(tile, tsub) + g_lane;
#pragma unroll
                for (int j = 0; j < NJ; ++j)
                {
                    int lj = 0;
#pragma unroll
                    for (int b = 6; b < 6 + 4; ++b)
                        if ((j >> (b - 6)) & 1)
                            lj += op.lw[b];
                    tl[l_lane + lj] = v[j];
                }
                wave_sync();
#if QSV_BLOCK_PREFETCH
                if (tile + nwaves < ntiles)
                    load_tile(tile + nwaves, v);
#endif
                float b_[H];
#pragma unroll
                for (int r = 0; r < AMPS / 2; ++r)
                {
                    v4f w = *reinterpret_cast<const v4f *>(&tl[c * ROW + h * AMPS + 2 * r]);
                    b_[4 * r + 0] = w.x;
                    b_[4 * r + 1] = w.y;
                    b_[4 * r + 2] = w.z;
                    b_[4 * r + 3] = w.w;
                }
                typedef BlockMatrix<KQ, X3> Mat;
                v16f d[Mat::NACC];
#pragma unroll
                for (int i = 0; i < Mat::NACC; ++i)
                    d[i] = v16f{0.f};
                A.apply(b_, d);
                wave_sync();
                Mat::outputs2(d, [&](int e, v4f v) { *reinterpret_cast<v4f *>(&tl[c * ROW + e + Mat::OUT_H * h]) = v; });
                wave_sync();
#pragma unroll
                for (int j = 0; j < NJ; ++j)
                {
                    uint64_t gj = 0;
                    int lj = 0;
#pragma unroll
                    for (int b = 6; b < 6 + 4; ++b)
                        if ((j >> (b - 6)) & 1)
                        {
                            gj += op.gw[b];
                            lj += op.lw[b];
                        }
                    __builtin_nontemporal_store(tl[l_lane + lj], reinterpret_cast<v2f *>(s + (base + gj)));
                }
                wave_sync();
#if !QSV_BLOCK_PREFETCH
                if (tile + nwaves < ntiles)
                    load_tile(tile + nwaves, v);
#endif
            }
        }
        // complex128 block on 4 qubits.  Tile = 16 amplitudes x 64 columns (tile bits = the 4
        // block qubits + the 6 lowest other qubits): moved in address order through LDS,
        // then each lane applies the 16x16 complex matrix to its column in FP64 with the
        // matrix read through uniform (scalar) loads.
        constexpr int kRow128 = 17; // complex per LDS row (one of padding)

        __global__ __launch_bounds__(kThreads) void k_block128_lds(c128 *__restrict__ s, uint64_t ntiles, Subspace tsub, Block128Op op)
        {
            typedef double v2d __attribute__((ext_vector_type(2)));
            __shared__ __attribute__((aligned(16))) v2d lds[kWaves][64 * kRow128];
            const int lane = threadIdx.x & 63;
            const int wave = threadIdx.x >> 6;
            v2d *tl = lds[wave];
            const double *__restrict__ M = op.M;
            uint64_t g_lane = 0;
            int l_lane = 0;
#pragma unroll
            for (int b = 0; b < 6; ++b)
                if ((lane >> b) & 1)
                {
                    g_lane += op.gw[b];
                    l_lane += op.lw[b];
                }
            const uint64_t nwaves = uint64_t(gridDim.x) * kWaves;
            for (uint64_t tile = uint64_t(blockIdx.x) * kWaves + wave; tile < ntiles; tile += nwaves)
            {
                const uint64_t base = expand(tile, tsub) + g_lane;
                v2d v[16];
#pragma unroll
                for (int j = 0; j < 16; ++j)
                {
                    uint64_t gj = 0;
#pragma unroll
                    for (int b = 6; b < 10; ++b)
                        if ((j >> (b - 6)) & 1)
                            gj += op.gw[b];
                    v[j] = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(s + (base + gj)));
                }
#pragma unroll
                for (int j = 0; j < 16; ++j)
                {
                    int lj = 0;
#pragma unroll
                    for (int b = 6; b < 10; ++b)
                        if ((j >> (b - 6)) & 1)
                            lj += op.lw[b];
                    tl[l_lane + lj] = v[j];
                }
                wave_sync();
                v2d x[16];
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    x[k] = tl[lane * kRow128 + k];
#pragma unroll
                for (int r = 0; r < 16; ++r)
                {
                    double yr = 0.0, yi = 0.0;
#pragma unroll
                    for (int k = 0; k < 16; ++k)
                    {
                        const double mr = M[2 * (16 * r + k)], mi = M[2 * (16 * r + k) + 1];
                        yr = __builtin_fma(mr, x[k].x, yr);
                        yr = __builtin_fma(-mi, x[k].y, yr);
                        yi = __builtin_fma(mr, x[k].y, yi);
                        yi = __builtin_fma(mi, x[k].x, yi);
                    }
                    tl[lane * kRow128 + r] = v2d{yr, yi};
                }
                wave_sync();
#pragma unroll
                for (int j = 0; j < 16; ++j)
                {
                    uint64_t gj = 0;
                    int lj = 0;
#pragma unroll
                    for (int b = 6; b < 10; ++b)
                        if ((j >> (b - 6)) & 1)
                        {
                            gj += op.gw[b];
                            lj += op.lw[b];
                        }
                    __builtin_nontemporal_store(tl[l_lane + lj], reinterpret_cast<v2d *>(s + (base + gj)));
                }
                wave_sync();
            }
        }

        // complex128 block without LDS staging (fewer than 6 qubits outside the block): one
        // lane per column, gathered directly.
        __global__ __launch_bounds__(kThreads) void k_block128(c128 *__restrict__ s, uint64_t ncols, Subspace sub, Block128Op op)
        {
            typedef double v2d __attribute__((ext_vector_type(2)));
            const double *__restrict__ M = op.M;
            for (uint64_t col = uint64_t(blockIdx.x) * kThreads + threadIdx.x; col < ncols; col += uint64_t(gridDim.x) * kThreads)
            {
                const uint64_t base = expand(col, sub);
                v2d x[16];
#pragma unroll
                for (int k = 0; k < 16; ++k)
                    x[k] = *reinterpret_cast<const v2d *>(s + (base | op.off[k]));
#pragma unroll
                for (int r = 0; r < 16; ++r)
                {
                    double yr = 0.0, yi = 0.0;
#pragma unroll
                    for (int k = 0; k < 16; ++k)
                    {
                        const double mr = M[2 * (16 * r + k)], mi = M[2 * (16 * r + k) + 1];
                        yr = __builtin_fma(mr, x[k].x, yr);
                        yr = __builtin_fma(-mi, x[k].y, yr);
                        yi = __builtin_fma(mr, x[k].y, yi);
                        yi = __builtin_fma(mi, x[k].x, yi);
                    }
                    *reinterpret_cast<v2d *>(s + (base | op.off[r])) = v2d{yr, yi};
                }
            }
        }

        // Barrier among the W waves of a stage workgroup: a single wave only orders its own
        // LDS accesses.
        template <int W>
        __device__ __forceinline__ void stage_barrier()
        {
            if (W == 1)
                wave_sync();
            else
                lds_barrier();
        }

        // Swizzled LDS byte offsets of one block's tile bits: tb[j] of the block's j-th qubit,
        // cb[j] of its j-th column bit; uniform, so kept in scalar registers.
        template <int KQ, int M>
        struct BlockOffsets
        {
            uint32_t tb[KQ], cb[M - KQ];
            __device__ __forceinline__ void read(const __attribute__((address_space(4))) StageBlock &B)
            {
#pragma unroll
                for (int j = 0; j < KQ; ++j)
                    tb[j] = B.lds[j];
#pragma unroll
                for (int j = 0; j < M - KQ; ++j)
                    cb[j] = B.lds[KQ + j];
            }
        };

        // One block of a stage on the tile in LDS: CWAVES waves (index `wave`) share the
        // 2^(M-KQ) columns, 32 per MFMA column group, GCAP (0: the tuned default) groups per
        // pass at most.  The caller separates blocks with a workgroup barrier.
        //
        // PIPE: the caller's load_part(p) reads part p of the matrix A from LDS (what the
        // products of the gathered amplitudes 4p .. 4p + 3 need), and the first pass is
        // software-pipelined over the parts: the gathers and matrix part p + 1 (and a quarter
        // of the next pass's gathers) are issued before the products of part p, so the
        // products start once the first part has arrived instead of after every LDS read of
        // the block.  Same products in the same order either way.
        template <int KQ, int M, bool X3, int CWAVES, int GCAP = 0, bool PIPE = false, class LoadPart>
        __device__ __forceinline__ void stage_block_at(v2f *tile, const BlockMatrix<KQ, X3> &A, LoadPart &&load_part, const BlockOffsets<KQ, M> &O, int wave, int c, int h)
        {
            constexpr int D = 2 << KQ;
            constexpr int H = D / 2;
            constexpr int AMPS = H / 2;
            constexpr int NT = 1 << M;
            constexpr int COLS = NT >> KQ;
            constexpr int PER_WAVE = COLS / 32 / CWAVES;
            static_assert(PER_WAVE >= 1, "tile too small");
            // column groups per pass (each adds MT accumulators and H floats per lane)
            constexpr int GMAX = GCAP > 0 ? GCAP : KQ == 5 ? (X3 ? QSV_STAGE_GROUPS_KQ5_X3 : QSV_STAGE_GROUPS_KQ5) : QSV_STAGE_GROUPS_KQ4;
            constexpr int G = PER_WAVE < GMAX ? PER_WAVE : GMAX;
            static_assert(PER_WAVE % G == 0, "groups per pass must divide the groups per wave");
            constexpr bool AHEAD = (X3 || (PIPE && QSV_STREAM_AHEAD_FP32)) && QSV_STAGE_GATHER_AHEAD;
            constexpr int PARTS = BlockMatrix<KQ, X3>::PARTS;
            static_assert(AMPS == 4 * PARTS, "a matrix part per 4 gathered amplitudes");
            const uint32_t(&tb)[KQ] = O.tb;
            const uint32_t(&cb)[M - KQ] = O.cb;
            // amplitude a of a column: XOR of tb[j] over the set bits j of a
            auto amp_off = [&](int a) {
                uint32_t o = 0;
#pragma unroll
                for (int j = 0; j < KQ; ++j)
                    if ((a >> j) & 1)
                        o ^= tb[j];
                return o;
            };
            // the lane's column bits 0..4 (c) and its amplitude half h: gathers read
            // amplitudes h*AMPS + e, scatters write 16*mt + 4*q + 2*h (+1)
            uint32_t lanep = 0;
#pragma unroll
            for (int j = 0; j < 5; ++j)
                lanep ^= uint32_t(-int((c >> j) & 1)) & cb[j];
            const uint32_t hmask = uint32_t(-h);
            const uint32_t g_lane = lanep ^ (hmask & amp_off(AMPS));
            const uint32_t s_lane = lanep ^ (hmask & amp_off(BlockMatrix<KQ, X3>::OUT_H));
            const char *tbytes = reinterpret_cast<const char *>(tile);
            char *wbytes = reinterpret_cast<char *>(tile);
            // G column groups per pass (independent MFMA chains).  The next pass's gathers are
            // issued before this pass's products and scatters (its columns are disjoint from
            // this pass's), so their LDS latency hides under the MFMA work.
            auto group_base = [&](int gr0, uint32_t (&U)[G]) {
#pragma unroll
                for (int g = 0; g < G; ++g)
                {
                    // column bits 5.. of this group (uniform)
                    const int grp = wave * PER_WAVE + gr0 + g;
                    U[g] = 0;
#pragma unroll
                    for (int j = 5; j < M - KQ; ++j)
                        if ((grp >> (j - 5)) & 1)
                            U[g] ^= cb[j];
                }
            };
            // amplitudes 4p .. 4p + 3 of every group of a pass
            auto gather_part = [&](const uint32_t (&U)[G], float (&x)[G][H], int p) {
#pragma unroll
                for (int g = 0; g < G; ++g)
                {
                    const uint32_t gbase = g_lane ^ U[g];
#pragma unroll
                    for (int e = 4 * p; e < 4 * p + 4; ++e)
                    {
                        const v2f a = *reinterpret_cast<const v2f *>(tbytes + (gbase ^ amp_off(e)));
                        x[g][2 * e] = a.x;
                        x[g][2 * e + 1] = a.y;
                    }
                }
            };
            auto gather = [&](int gr0, float (&x)[G][H], uint32_t (&U)[G]) {
                group_base(gr0, U);
#pragma unroll
                for (int p = 0; p < PARTS; ++p)
                    gather_part(U, x, p);
            };
            uint32_t U[G], Un[G];
            float x[G][H], xn[G][H];
            if (!PIPE)
                gather(0, x, U);
#pragma unroll
            for (int gr0 = 0; gr0 < PER_WAVE; gr0 += G)
            {
                const bool more = AHEAD && gr0 + G < PER_WAVE;
                constexpr int NACC = BlockMatrix<KQ, X3>::NACC;
                v16f d[G][NACC];
#pragma unroll
                for (int g = 0; g < G; ++g)
#pragma unroll
                    for (int i = 0; i < NACC; ++i)
                        d[g][i] = v16f{0.f};
                if (PIPE && gr0 == 0)
                {
                    group_base(0, U);
                    if (more)
                        group_base(G, Un);
                    gather_part(U, x, 0);
                    load_part(0);
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int p = 0; p < PARTS; ++p)
                    {
                        if (p + 1 < PARTS)
                        {
                            gather_part(U, x, p + 1);
                            load_part(p + 1);
                        }
                        if (more)
                            gather_part(Un, xn, p);
                        A.template apply_part<G>(x, d, p);
                        __builtin_amdgcn_sched_barrier(0);
                    }
                }
                else
                {
                    if (more)
                        gather(gr0 + G, xn, Un);
                    A.template apply_n<G>(x, d);
                }
                // amplitudes e and e + 1 (e even, plus OUT_H*h in s_lane) of each column
#pragma unroll
                for (int g = 0; g < G; ++g)
                {
                    const uint32_t sbase = s_lane ^ U[g];
                    BlockMatrix<KQ, X3>::outputs2(d[g], [&](int e, v4f v) {
                        const uint32_t o0 = sbase ^ amp_off(e);
                        *reinterpret_cast<v2f *>(wbytes + o0) = v2f{v.x, v.y};
                        *reinterpret_cast<v2f *>(wbytes + (o0 ^ tb[0])) = v2f{v.z, v.w};
                    });
                }
                if (gr0 + G < PER_WAVE)
                {
                    if (!AHEAD)
                        gather(gr0 + G, xn, Un);
#pragma unroll
                    for (int g = 0; g < G; ++g)
                    {
                        U[g] = Un[g];
#pragma unroll
                        for (int i = 0; i < H; ++i)
                            x[g][i] = xn[g][i];
                    }
                }
            }
        }

        // The same with the matrix already in registers.
        template <int KQ, int M, bool X3, int CWAVES, int GCAP = 0>
        __device__ __forceinline__ void stage_block_with(v2f *tile, const BlockMatrix<KQ, X3> &A, const __attribute__((address_space(4))) StageBlock &B, int wave, int c, int h)
        {
            BlockOffsets<KQ, M> O;
            O.read(B);
            stage_block_at<KQ, M, X3, CWAVES, GCAP, false>(tile, A, [](int) {}, O, wave, c, h);
        }

        template <int KQ, int M, bool X3, int CWAVES>
        __device__ __forceinline__ void stage_block(v2f *tile, const float *__restrict__ mats, const __attribute__((address_space(4))) StageBlock &B, int wave, int c, int h)
        {
            BlockMatrix<KQ, X3> A;
            A.load(mats + B.matrix, c, h);
            // all of the matrix loads are issued here, together (one L2 round trip),
            // not moved next to their first uses
            __builtin_amdgcn_sched_barrier(0);
            stage_block_with<KQ, M, X3, CWAVES>(tile, A, B, wave, c, h);
        }

        // Stage kernel: one workgroup holds a tile of 2^M amplitudes in LDS -- the
        // amplitudes whose index bits outside the stage's tile qubits T are fixed -- and
        // applies every block of the stage to it before writing it back, so a stage costs
        // one HBM sweep however many blocks it holds.  T[0..3] are qubits 0..3, so each
        // thread moves 2 adjacent amplitudes (16 B) per load/store instruction and 8 lanes
        // cover one 128-B line.  LDS index bit b <-> tile qubit T[b].
        //
        // Per block, the 2^(M-KQ) columns (tile bits outside the block) are split into
        // groups of 32 over the 4 waves; lane (h, c) of a wave gathers column c's half h
        // from LDS (LDS weights of the block's bits and of the column bits come with the
        // block), multiplies on MFMA as the block kernels do, and writes its rows back in
        // place.  Workgroup barriers (LDS-only) separate the blocks.  The block matrices come
        // from global memory (L2) into registers per block; block descriptors through scalar
        // loads.  One launch may serve several states (tasks): the launch's tiles are numbered
        // across its tasks, and a workgroup finds the task of each tile by binary search over
        // the tasks' first tile numbers (batched circuits use one launch per stage round).
        typedef const __attribute__((address_space(4))) StageTask *TaskPtr;

        // expand() reading the subspace through scalar loads
        __device__ __forceinline__ uint64_t expand_task(uint64_t t, TaskPtr T)
        {
            const int nops = T->tsub.nops;
            for (int k = 0; k < nops; ++k)
            {
                const int op = T->tsub.ops[k];
                const uint64_t lo = t & ((uint64_t(1) << op) - 1);
                t = lo | ((t ^ lo) << 1);
            }
            return t | T->tsub.set_bits;
        }

        // Moves tiles between HBM and LDS: NTHM threads (index tm) move 16 B (2 amplitudes)
        // per v4f; the thread's i-th v4f is LDS index 2*tm + i*2*NTHM, whose tile bits 1.. are
        // set by tm (bits 1 .. TB-1) and i (bits TB ..).
        template <int M, int NTHM>
        struct TileIO
        {
            static constexpr int LOADS = (1 << M) / (2 * NTHM);
            static constexpr int TB = 1 + __builtin_ctz(NTHM);
            static_assert(LOADS >= 1, "tile too small");
            TaskPtr all;
            int ntasks;
            int tm;
            // the task owning global tile g (the last one with tile_begin <= g)
            __device__ __forceinline__ TaskPtr task(uint64_t g) const
            {
                int lo = 0, hi = ntasks - 1;
                while (lo < hi)
                {
                    const int mid = (lo + hi + 1) >> 1;
                    if (all[mid].tile_begin <= g)
                        lo = mid;
                    else
                        hi = mid - 1;
                }
                return all + lo;
            }
            // tile g: its task and this thread's first global amplitude
            __device__ __forceinline__ void locate(uint64_t g, TaskPtr &T, uint64_t &base) const
            {
                T = task(g);
                uint64_t g_t = 0;
#pragma unroll
                for (int b = 1; b < TB && b < M; ++b)
                    g_t += uint64_t(-int64_t(((2 * tm) >> b) & 1)) & T->gw[b];
                base = expand_task(g - T->tile_begin, T) + g_t;
            }
            __device__ __forceinline__ uint64_t offset(TaskPtr T, int i) const
            {
                uint64_t gi = 0;
#pragma unroll
                for (int b = TB; b < M; ++b)
                    if (((i * 2 * NTHM) >> b) & 1)
                        gi += T->gw[b];
                return gi;
            }
            __device__ __forceinline__ void fetch(TaskPtr T, uint64_t base, v4f *v) const
            {
                const c64 *__restrict__ s = T->state;
                if (T->flags & kStageZeroState)
                {
                    // |0...0>: amplitude 1 at index 0 (the first of this v4f's pair)
#pragma unroll
                    for (int i = 0; i < LOADS; ++i)
                        v[i] = v4f{base + offset(T, i) == 0 ? 1.f : 0.f, 0.f, 0.f, 0.f};
                    return;
                }
#pragma unroll
                for (int i = 0; i < LOADS; ++i)
                    v[i] = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(s + (base + offset(T, i))));
            }
            __device__ __forceinline__ void put(v2f *tile, const v4f *v) const
            {
#pragma unroll
                for (int i = 0; i < LOADS; ++i)
                    *reinterpret_cast<v4f *>(&tile[swz64(2 * tm + i * 2 * NTHM)]) = v[i];
            }
            __device__ __forceinline__ void store(const v2f *tile, TaskPtr T, uint64_t base, const float *__restrict__ mats) const
            {
                c64 *__restrict__ s = T->state;
                if (T->flags & kStageTrailTable) // uniform: the diagonal blocks after the others
                {
                    // unrolled by 4 only: fully unrolled, this rarely taken path raised the
                    // kernel's register count (k_stage<4, 11, true, 1>: 155 -> 178) and cost
                    // the dense stages a wave per SIMD
                    const v4f *tab = reinterpret_cast<const v4f *>(mats + T->trail);
#pragma unroll 4
                    for (int i = 0; i < LOADS; ++i)
                    {
                        // tile indices e, e + 1 with e = 2 * tm + i * 2 * NTHM
                        const v4f p = tab[tm + i * NTHM];
                        const v4f v = *reinterpret_cast<const v4f *>(&tile[swz64(2 * tm + i * 2 * NTHM)]);
                        const v4f y{__builtin_fmaf(p.x, v.x, -p.y * v.y), __builtin_fmaf(p.x, v.y, p.y * v.x),
                                    __builtin_fmaf(p.z, v.z, -p.w * v.w), __builtin_fmaf(p.z, v.w, p.w * v.z)};
                        __builtin_nontemporal_store(y, reinterpret_cast<v4f *>(s + (base + offset(T, i))));
                    }
                    return;
                }
#pragma unroll
                for (int i = 0; i < LOADS; ++i)
                {
                    const v4f v = *reinterpret_cast<const v4f *>(&tile[swz64(2 * tm + i * 2 * NTHM)]);
                    __builtin_nontemporal_store(v, reinterpret_cast<v4f *>(s + (base + offset(T, i))));
                }
            }
        };

        template <int KQ, int M, bool X3, int W>
        __global__ __launch_bounds__(64 * W) void k_stage(const float *__restrict__ mats, const StageTask *__restrict__ tasks, int ntasks, uint64_t total_tiles)
        {
            __shared__ __attribute__((aligned(16))) v2f tile[1 << M];
            const int lane = threadIdx.x & 63;
            const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            const int c = lane & 31;
            const int h = lane >> 5;
            const TileIO<M, 64 * W> io{as_constant(tasks), ntasks, int(threadIdx.x)};

            // one tile per workgroup (a loop only for launches beyond the grid limit): the
            // load -> blocks -> store phases of different workgroups overlap on a CU, and no
            // wave waits for its own stores before loading the next tile
            for (uint64_t g = blockIdx.x; g < total_tiles; g += gridDim.x)
            {
                TaskPtr T;
                uint64_t base;
                io.locate(g, T, base);
                {
                    v4f v[TileIO<M, 64 * W>::LOADS];
                    io.fetch(T, base, v);
                    io.put(tile, v);
                }
                stage_barrier<W>();
                const int nblocks = T->nblocks;
                const __attribute__((address_space(4))) StageBlock *blocks = as_constant(reinterpret_cast<const StageBlock *>(mats + T->blocks));
                for (int bi = 0; bi < nblocks; ++bi)
                {
                    stage_block<KQ, M, X3, W>(tile, mats, blocks[bi], wave, c, h);
                    stage_barrier<W>();
                }
                io.store(tile, T, base, mats);
                // the LDS tile is reused only when this workgroup takes another tile
                if (g + gridDim.x < total_tiles)
                    stage_barrier<W>();
            }
        }

        // Streaming stage kernel: a persistent workgroup per CU with two LDS tile buffers and
        // two LDS block-matrix buffers (160 KiB), and three kinds of wave:
        //   * 4 compute waves run the blocks of tile i on one tile buffer, exactly as
        //     k_stage<KQ, M, X3, 4> does (same columns per wave, same products: the same
        //     bits), reading each block's matrix from LDS;
        //   * 2 tile waves store tile i - 1 from the other tile buffer and refill it with
        //     tile i + 1 through direct-to-LDS loads (no registers hold the data in flight);
        //   * 2 matrix waves copy the next block's matrix into the other matrix buffer, also
        //     direct to LDS.
        // A wave's vector-memory counter completes in issue order and counts stores as well as
        // loads; giving each kind of traffic its own waves means that no wait for one kind
        // waits for the other.  The compute waves issue no global memory operations at all.
        // All waves pass the same barriers: one after the first tile's load and the first
        // matrix, then one per block.
        //
        // Tile wave tw owns the 1-KiB chunks CH*tw .. CH*tw + CH - 1 of both tile buffers (64
        // amplitude pairs each, CH = 2^(M-7)/2); lane l of chunk j holds LDS pair slot
        // 64 j + l, whose tile index is swz64(128 j + 2 l).  The per-lane global offsets of
        // those tile indices are the same for every tile, so they are computed once.
        typedef __attribute__((address_space(3))) void lds_void;
        typedef __attribute__((address_space(1))) void glb_void;
        constexpr int kStreamCompute = 4, kStreamTile = 2, kStreamMatrix = 2;
        constexpr int kStreamWaves = kStreamCompute + kStreamTile + kStreamMatrix;

        template <int M>
        struct StreamChunks
        {
            static constexpr int CH = (1 << M) / 128 / kStreamTile;
            uint32_t off[CH]; // global offset of the lane's pair in pairs (the state has at most 2^33 amplitudes), per owned chunk
            int first;        // first owned chunk
            int lane;
            __device__ __forceinline__ int tile_index(int k) const { return swz64(128 * (first + k) + 2 * lane); }
            __device__ __forceinline__ void init(TaskPtr T, int tw, int l)
            {
                first = CH * tw;
                lane = l;
#pragma unroll
                for (int k = 0; k < CH; ++k)
                {
                    const int e = tile_index(k);
                    uint64_t o = 0;
#pragma unroll
                    for (int b = 1; b < M; ++b)
                        o += T->gw[b] & (uint64_t(0) - uint64_t((e >> b) & 1));
                    off[k] = uint32_t(o >> 1);
                }
            }
            // owned chunks K0 .. K0 + P - 1
            template <int K0, int P>
            __device__ __forceinline__ void load(v2f *buf, const c64 *s, uint64_t base) const
            {
#pragma unroll
                for (int k = K0; k < K0 + P; ++k)
                    __builtin_amdgcn_global_load_lds((const glb_void *)(s + (base + 2 * uint64_t(off[k]))),
                                                     (lds_void *)(buf + 128 * (first + k)), 16, 0, 0);
            }
            // owned chunks of a finished tile from LDS into v, times the trailing table
            template <int K0, int P>
            __device__ __forceinline__ void read(const v2f *buf, TaskPtr T, const float *__restrict__ mats, v4f (&v)[P]) const
            {
#pragma unroll
                for (int k = 0; k < P; ++k)
                    v[k] = *reinterpret_cast<const v4f *>(buf + 128 * (first + K0 + k) + 2 * lane);
                if (T->flags & kStageTrailTable) // uniform: the diagonal blocks after the others
                {
                    const v4f *tab = reinterpret_cast<const v4f *>(mats + T->trail);
#pragma unroll
                    for (int k = 0; k < P; ++k)
                    {
                        const v4f p = tab[tile_index(K0 + k) >> 1];
                        const v4f x = v[k];
                        v[k] = v4f{__builtin_fmaf(p.x, x.x, -p.y * x.y), __builtin_fmaf(p.x, x.y, p.y * x.x),
                                   __builtin_fmaf(p.z, x.z, -p.w * x.w), __builtin_fmaf(p.z, x.w, p.w * x.z)};
                    }
                }
            }
            template <int K0, int P>
            __device__ __forceinline__ void write(c64 *s, uint64_t base, const v4f (&v)[P]) const
            {
#pragma unroll
                for (int k = 0; k < P; ++k)
                    __builtin_nontemporal_store(v[k], reinterpret_cast<v4f *>(s + (base + 2 * uint64_t(off[K0 + k]))));
            }
            // Tile `out` (if any) from buf to HBM and tile `in` (if any) from HBM into buf, in
            // PARTS parts: each part's LDS reads, then its loads, then its stores.  Waiting
            // for the loads (issue order) then waits for all but the last part's stores.
            // The parts are spread over the blocks of the tile (PER_BLOCK parts per block, the
            // rest with the last block), so that no barrier waits for the whole exchange.
            static constexpr int PARTS = QSV_STREAM_PARTS, PART = CH / PARTS, PER_BLOCK = QSV_STREAM_PER_BLOCK;
            static_assert(CH % PARTS == 0, "parts must divide the chunks");
            // the parts due at block bi of nblocks
            template <int Q = 0>
            __device__ __forceinline__ void exchange_at(int bi, int nblocks, v2f *buf, c64 *s, TaskPtr T, const float *__restrict__ mats, bool has_out, uint64_t out, bool has_in, uint64_t in) const
            {
                if constexpr (Q < PARTS)
                {
                    const int due = Q / PER_BLOCK < nblocks - 1 ? Q / PER_BLOCK : nblocks - 1;
                    if (due == bi)
                        part<Q * PART>(buf, s, T, mats, has_out, out, has_in, in);
                    exchange_at<Q + 1>(bi, nblocks, buf, s, T, mats, has_out, out, has_in, in);
                }
            }
            template <int K0>
            __device__ __forceinline__ void part(v2f *buf, c64 *s, TaskPtr T, const float *__restrict__ mats, bool has_out, uint64_t out, bool has_in, uint64_t in) const
            {
                v4f v[PART];
                if (has_out)
                    read<K0, PART>(buf, T, mats, v);
                if (has_in)
                {
                    // the LDS reads have returned (the table products, when there are any,
                    // waited for the table loads); nothing may move the loads above them.
                    // lgkmcnt(0) only: a vector-memory wait here would wait for the earlier
                    // parts' stores.
                    __builtin_amdgcn_s_waitcnt(0xc07f);
                    __builtin_amdgcn_sched_barrier(0);
                    load<K0, PART>(buf, s, in);
                    __builtin_amdgcn_sched_barrier(0);
                }
                if (has_out)
                    write<K0, PART>(s, out, v);
            }
        };

        // The barrier of the tile and matrix waves.  lds_barrier's release fence would make
        // the wave wait for its direct-to-LDS loads at every barrier (they write LDS); these
        // waves wait for them explicitly where needed, and their own LDS reads complete
        // before the stores that use them.  The acquire fence keeps the LDS reads of the next
        // tile's stores after the barrier.
        __device__ __forceinline__ void stream_barrier()
        {
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        }

        // s_waitcnt immediate that waits until at most n vector-memory operations are
        // outstanding (lgkmcnt and expcnt at their maxima: no wait on them)
        constexpr int vm_wait(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }

        // tile number -> its first amplitude (single task, tile_begin 0)
        __device__ __forceinline__ uint64_t stream_base(TaskPtr T, uint64_t g) { return expand_task(g, T); }

        template <int KQ, int M, bool X3>
        __global__ __launch_bounds__(64 * kStreamWaves) void k_stage_stream(const float *__restrict__ mats, const StageTask *__restrict__ task, uint64_t total_tiles)
        {
            typedef BlockMatrix<KQ, X3> Mat;
            __shared__ __attribute__((aligned(16))) v2f buf0[1 << M];
            __shared__ __attribute__((aligned(16))) v2f buf1[1 << M];
            __shared__ __attribute__((aligned(16))) v4f mat0[Mat::UNITS * 64];
            __shared__ __attribute__((aligned(16))) v4f mat1[Mat::UNITS * 64];
            const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            const int lane = threadIdx.x & 63;
            const TaskPtr T = as_constant(task);
            const int nblocks = T->nblocks;
            const __attribute__((address_space(4))) StageBlock *blocks = as_constant(reinterpret_cast<const StageBlock *>(mats + T->blocks));
            const uint64_t stride = gridDim.x;
            // this workgroup's tiles: blockIdx.x + i * gridDim.x, i < count (gridDim.x <= total_tiles)
            const uint64_t count = (total_tiles - blockIdx.x + stride - 1) / stride;
            // steps: tile k / nblocks, block k % nblocks; step k reads matrix buffer k & 1
            const uint64_t steps = count * uint64_t(nblocks);
            if (wave >= kStreamCompute + kStreamTile)
            {
                // matrix waves: wave mw copies units j = mw, mw + 2, ... of each matrix
                const int mw = wave - kStreamCompute - kStreamTile;
                auto copy = [&](v4f *dst, int bi) {
                    const char *src = reinterpret_cast<const char *>(mats + blocks[bi].matrix);
#pragma unroll
                    for (int j = 0; j < Mat::UNITS; j += kStreamMatrix)
                        __builtin_amdgcn_global_load_lds((const glb_void *)(src + 16 * Mat::source_unit(j + mw, lane)),
                                                         (lds_void *)(dst + 64 * (j + mw)), 16, 0, 0);
                };
                copy(mat0, 0);
                __builtin_amdgcn_s_waitcnt(0);
                stream_barrier();
                int bi = 0;
                for (uint64_t k = 0; k < steps; ++k)
                {
                    bi = bi + 1 == nblocks ? 0 : bi + 1;
                    // the buffer the compute waves read in step k - 1 (finished before the
                    // barrier that began step k)
                    if (k + 1 < steps)
                        copy((k & 1) ? mat0 : mat1, bi);
                    __builtin_amdgcn_s_waitcnt(0);
                    stream_barrier();
                }
                return;
            }
            if (wave >= kStreamCompute)
            {
                StreamChunks<M> io;
                io.init(T, wave - kStreamCompute, lane);
                c64 *__restrict__ s = T->state;
                io.template load<0, StreamChunks<M>::CH>(buf0, s, stream_base(T, blockIdx.x));
                __builtin_amdgcn_s_waitcnt(0);
                stream_barrier();
                for (uint64_t i = 0; i < count; ++i)
                {
                    // the other buffer: tile i - 1 out, tile i + 1 in, part by part during the
                    // blocks of tile i (the stores of each part after its loads, so that they
                    // drain during the next tile)
                    v2f *other = (i & 1) ? buf0 : buf1;
                    const bool has_out = i > 0, has_in = i + 1 < count;
                    const uint64_t out = has_out ? stream_base(T, blockIdx.x + (i - 1) * stride) : 0;
                    const uint64_t in = has_in ? stream_base(T, blockIdx.x + (i + 1) * stride) : 0;
                    for (int bi = 0; bi < nblocks; ++bi)
                    {
                        // opaque per block: keeps the compiler from computing every part's
                        // addresses ahead of the loop (they would not fit in registers)
                        uint64_t out_b = out, in_b = in;
                        asm volatile("" : "+s"(out_b), "+s"(in_b));
                        io.exchange_at(bi, nblocks, other, s, T, mats, has_out, out_b, has_in, in_b);
                        if (bi == nblocks - 1) // tile i + 1 in LDS; the last part's stores may be in flight
                        {
                            if (i > 0)
                                __builtin_amdgcn_s_waitcnt(vm_wait(StreamChunks<M>::PART));
                            else
                                __builtin_amdgcn_s_waitcnt(0);
                        }
                        stream_barrier();
                    }
                }
                io.exchange_at(0, 1, (count & 1) ? buf0 : buf1, s, T, mats, true, stream_base(T, blockIdx.x + (count - 1) * stride), false, 0);
                return;
            }
            const int c = lane & 31;
            const int h = lane >> 5;
            // The next block's LDS offsets are read at the end of each step, where the
            // barrier waits for the step's LDS accesses anyway: a scalar load outstanding
            // during the step would make each wait for its LDS reads a wait for all of them.
            BlockOffsets<KQ, M> O;
            O.read(blocks[0]);
            lds_barrier();
            uint64_t i = 0;
            int bi = 0;
            for (uint64_t k = 0; k < steps; ++k)
            {
                Mat A;
                const v4f *m = (k & 1) ? mat1 : mat0;
                // the matrix is read from LDS part by part, pipelined with the first products
                stage_block_at<KQ, M, X3, kStreamCompute, X3 ? QSV_STREAM_GROUPS_X3 : QSV_STREAM_GROUPS, true>((i & 1) ? buf1 : buf0, A, [&](int p) { A.load_lds_part(m, lane, p); }, O, wave, c, h);
                bi = bi + 1 == nblocks ? 0 : bi + 1;
                i += bi == 0;
                __builtin_amdgcn_sched_barrier(0);
                O.read(blocks[bi]);
                lds_barrier();
            }
        }

        // complex128 stage: a tile of 2^M amplitudes (M = 11, 12, 13: 32-128 KiB) in LDS, tile
        // bits as in k_stage; a block of KQ = 3 or 4 qubits has 2^(M-KQ) columns, one or two
        // per thread, applied in FP64 with the 2^KQ x 2^KQ matrix read through uniform loads;
        // the next tile is prefetched into registers while the blocks run.  Threads per
        // workgroup: 2^(M-4) (16 prefetch registers per thread), at most one per column.
        constexpr int kStage128Bits = 12; // the default tile
        template <int KQ, int M>
        constexpr int stage128_threads() { return (1 << (M - 4)) < (1 << (M - KQ)) ? (1 << (M - 4)) : (1 << (M - KQ)); }

        template <int KQ, int M>
        __global__ __launch_bounds__((stage128_threads<KQ, M>())) void k_stage128(c128 *__restrict__ s, uint64_t ntiles, Subspace tsub, Stage128Op op)
        {
            typedef double v2d __attribute__((ext_vector_type(2)));
            constexpr int kThreads = stage128_threads<KQ, M>();
            constexpr int TB = __builtin_ctz(kThreads); // tile bits set by the thread index
            constexpr int NT = 1 << M;
            constexpr int LOADS = NT / kThreads;
            constexpr int K = 1 << KQ;
            constexpr int PER_THREAD = (NT >> KQ) / kThreads; // columns per thread and block
            static_assert(PER_THREAD >= 1 && LOADS >= 1, "stage128 shape");
            __shared__ __attribute__((aligned(16))) v2d tile[NT];
            const int t = threadIdx.x;
            uint64_t g_t = 0;
#pragma unroll
            for (int b = 0; b < TB; ++b)
                if ((t >> b) & 1)
                    g_t += op.gw[b];
            auto tile_offset = [&](int i) {
                uint64_t gi = 0;
#pragma unroll
                for (int b = TB; b < M; ++b)
                    if (((i * kThreads) >> b) & 1)
                        gi += op.gw[b];
                return gi;
            };
            // the next tile's loads stay in flight in registers while the blocks run (the
            // matrices come through scalar loads, the tile through LDS: neither waits on them)
            v2d pf[LOADS];
            uint64_t tl = blockIdx.x;
            if (tl < ntiles)
            {
                const uint64_t base = expand(tl, tsub) + g_t;
#pragma unroll
                for (int i = 0; i < LOADS; ++i)
                    pf[i] = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(s + (base + tile_offset(i))));
            }
            for (; tl < ntiles; tl += gridDim.x)
            {
                const uint64_t base = expand(tl, tsub) + g_t;
#pragma unroll
                for (int i = 0; i < LOADS; ++i)
                    tile[swz128(t + i * kThreads)] = pf[i];
                lds_barrier();
                const uint64_t next = tl + gridDim.x;
                if (next < ntiles)
                {
                    const uint64_t nbase = expand(next, tsub) + g_t;
#pragma unroll
                    for (int i = 0; i < LOADS; ++i)
                        pf[i] = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(s + (nbase + tile_offset(i))));
                }
                for (int bi = 0; bi < op.nblocks; ++bi)
                {
                    const __attribute__((address_space(4))) StageBlock &B = as_constant(op.blocks)[bi];
                    const __attribute__((address_space(4))) double *Mx = as_constant(op.M) + B.matrix;
                    if (B.pad[0] & kStageBlockDiagonal) // uniform: K products per column, not K^2
                    {
#pragma unroll
                        for (int cc = 0; cc < PER_THREAD; ++cc)
                        {
                            const int col = t + cc * kThreads;
                            int cbase = 0;
#pragma unroll
                            for (int j = 0; j < M - KQ; ++j)
                                if ((col >> j) & 1)
                                    cbase += B.w[KQ + j];
#pragma unroll
                            for (int e = 0; e < K; ++e)
                            {
                                int off = cbase;
#pragma unroll
                                for (int j = 0; j < KQ; ++j)
                                    if ((e >> j) & 1)
                                        off += B.w[j];
                                const v2d x = tile[swz128(off)];
                                const double mr = Mx[2 * (K * e + e)], mi = Mx[2 * (K * e + e) + 1];
                                tile[swz128(off)] = v2d{__builtin_fma(mr, x.x, -mi * x.y), __builtin_fma(mr, x.y, mi * x.x)};
                            }
                        }
                        lds_barrier();
                        continue;
                    }
#pragma unroll
                    for (int cc = 0; cc < PER_THREAD; ++cc)
                    {
                        const int col = t + cc * kThreads;
                        int cbase = 0;
#pragma unroll
                        for (int j = 0; j < M - KQ; ++j)
                            if ((col >> j) & 1)
                                cbase += B.w[KQ + j];
                        int off[K];
#pragma unroll
                        for (int e = 0; e < K; ++e)
                        {
                            off[e] = cbase;
#pragma unroll
                            for (int j = 0; j < KQ; ++j)
                                if ((e >> j) & 1)
                                    off[e] += B.w[j];
                        }
                        v2d x[K];
#pragma unroll
                        for (int k = 0; k < K; ++k)
                            x[k] = tile[swz128(off[k])];
#pragma unroll
                        for (int r = 0; r < K; ++r)
                        {
                            double yr = 0.0, yi = 0.0;
#pragma unroll
                            for (int k = 0; k < K; ++k)
                            {
                                const double mr = Mx[2 * (K * r + k)], mi = Mx[2 * (K * r + k) + 1];
                                yr = __builtin_fma(mr, x[k].x, yr);
                                yr = __builtin_fma(-mi, x[k].y, yr);
                                yi = __builtin_fma(mr, x[k].y, yi);
                                yi = __builtin_fma(mi, x[k].x, yi);
                            }
                            tile[swz128(off[r])] = v2d{yr, yi};
                        }
                    }
                    lds_barrier();
                }
                if (op.trail != nullptr) // uniform: the trailing diagonal blocks' product
                {
                    v2d ph[LOADS];
#pragma unroll
                    for (int i = 0; i < LOADS; ++i)
                        ph[i] = *reinterpret_cast<const v2d *>(op.trail + 2 * (t + i * kThreads));
#pragma unroll
                    for (int i = 0; i < LOADS; ++i)
                    {
                        const v2d x = tile[swz128(t + i * kThreads)];
                        const v2d y{__builtin_fma(ph[i].x, x.x, -ph[i].y * x.y), __builtin_fma(ph[i].x, x.y, ph[i].y * x.x)};
                        __builtin_nontemporal_store(y, reinterpret_cast<v2d *>(s + (base + tile_offset(i))));
                    }
                }
                else
                {
#pragma unroll
                    for (int i = 0; i < LOADS; ++i)
                        __builtin_nontemporal_store(tile[swz128(t + i * kThreads)], reinterpret_cast<v2d *>(s + (base + tile_offset(i))));
                }
                lds_barrier();
            }
        }
        // complex128 stage whose blocks are all diagonal: their product is one table over the
        // 2^M tile indices, and each amplitude is multiplied by its entry.  The state is
        // streamed in order (no tiles): amplitude a's tile index is a's bits 0..3 (tile qubits
        // 0..3) and its bits hi[0..M-5] (the other tile qubits) as bits 4..M-1.  Each workgroup
        // copies the 32-128 KiB table into LDS once; the 8 XCDs each stream one contiguous
        // eighth.  1024-thread workgroups (two per CU with the 64 KiB table) keep as many loads
        // in flight as the pair kernels' 256-thread ones.
        constexpr int kTablePer = 4, kTableThreads = 1024;
        template <int M>
        __global__ __launch_bounds__(kTableThreads) void k_stage128_table(c128 *__restrict__ s, uint64_t count, Subspace tsub, const double *__restrict__ table)
        {
            typedef double v2d __attribute__((ext_vector_type(2)));
            constexpr int kThreads = kTableThreads;
            __shared__ __attribute__((aligned(16))) v2d tab[1 << M];
            for (int i = threadIdx.x; i < (1 << M); i += kThreads)
                tab[i] = *reinterpret_cast<const v2d *>(table + 2 * i);
            __syncthreads();
            int hi[M - 4];
#pragma unroll
            for (int b = 0; b < M - 4; ++b)
                hi[b] = tsub.ops[4 + b];
            const uint64_t per_block = uint64_t(kThreads) * kTablePer;
            const uint64_t nchunks = count / per_block;
            uint64_t nb = gridDim.x, b0 = blockIdx.x;
            if ((nb & 7) == 0)
                b0 = (b0 & 7) * (nb >> 3) + (b0 >> 3);
            for (uint64_t c = b0; c < nchunks; c += nb)
            {
                v2d a[kTablePer];
#pragma unroll
                for (int j = 0; j < kTablePer; ++j)
                    a[j] = __builtin_nontemporal_load(reinterpret_cast<const v2d *>(s + (c * per_block + uint64_t(j) * kThreads + threadIdx.x)));
#pragma unroll
                for (int j = 0; j < kTablePer; ++j)
                {
                    const uint64_t i = c * per_block + uint64_t(j) * kThreads + threadIdx.x;
                    uint32_t e = uint32_t(i & 15);
#pragma unroll
                    for (int b = 0; b < M - 4; ++b)
                        e |= uint32_t((i >> hi[b]) & 1) << (4 + b);
                    const v2d ph = tab[e];
                    const v2d y{__builtin_fma(ph.x, a[j].x, -ph.y * a[j].y), __builtin_fma(ph.x, a[j].y, ph.y * a[j].x)};
                    __builtin_nontemporal_store(y, reinterpret_cast<v2d *>(s + i));
                }
            }
        }
        // complex64 stage whose blocks are all diagonal: as k_stage128_table, with the table
        // of 2^M complex floats in LDS and two adjacent amplitudes (tile indices e, e | 1) per
        // 16-byte access.
        constexpr int kTable64Per = 4, kTable64Threads = 1024;
        template <int M>
        __global__ __launch_bounds__(kTable64Threads) void k_stage_table64(c64 *__restrict__ s, uint64_t count, Subspace tsub, const float *__restrict__ table)
        {
            __shared__ __attribute__((aligned(16))) v2f tab[1 << M];
            for (int i = threadIdx.x; i < (1 << M); i += kTable64Threads)
                tab[i] = *reinterpret_cast<const v2f *>(table + 2 * i);
            __syncthreads();
            constexpr int kChunkBits = __builtin_ctz(kTable64Threads * kTable64Per * 2);
            const uint64_t per_block = uint64_t(1) << kChunkBits;
            const uint64_t nchunks = count / per_block;
            // An amplitude's tile index takes its bits 0..3 and its bits at the other tile
            // qubits; the bits below kChunkBits come from the thread's offset in a chunk (the
            // same in every chunk), the others from the chunk's base (uniform).  Both parts
            // are computed once: per thread here, per chunk below.
            uint32_t eoff[kTable64Per];
#pragma unroll
            for (int j = 0; j < kTable64Per; ++j)
            {
                const uint32_t o = 2 * (uint32_t(j) * kTable64Threads + threadIdx.x);
                uint32_t e = o & 15;
#pragma unroll
                for (int b = 0; b < M - 4; ++b)
                    if (tsub.ops[4 + b] < kChunkBits)
                        e |= ((o >> tsub.ops[4 + b]) & 1) << (4 + b);
                eoff[j] = e;
            }
            uint64_t nb = gridDim.x, b0 = blockIdx.x;
            if ((nb & 7) == 0)
                b0 = (b0 & 7) * (nb >> 3) + (b0 >> 3);
            for (uint64_t c = b0; c < nchunks; c += nb)
            {
                v4f a[kTable64Per];
#pragma unroll
                for (int j = 0; j < kTable64Per; ++j)
                    a[j] = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(s + (c * per_block + 2 * (uint64_t(j) * kTable64Threads + threadIdx.x))));
                const uint64_t cbase = c * per_block;
                uint32_t ebase = 0;
#pragma unroll
                for (int b = 0; b < M - 4; ++b)
                    if (tsub.ops[4 + b] >= kChunkBits)
                        ebase |= uint32_t((cbase >> tsub.ops[4 + b]) & 1) << (4 + b);
#pragma unroll
                for (int j = 0; j < kTable64Per; ++j)
                {
                    const uint64_t i = cbase + 2 * (uint64_t(j) * kTable64Threads + threadIdx.x);
                    // e is even: entries e and e + 1 in one 16-B read (a wave's reads of
                    // consecutive pairs are then one contiguous, conflict-free KiB)
                    const uint32_t e = ebase | eoff[j];
                    const v4f p = *reinterpret_cast<const v4f *>(&tab[e]);
                    const v2f p0{p.x, p.y}, p1{p.z, p.w};
                    const v4f y{__builtin_fmaf(p0.x, a[j].x, -p0.y * a[j].y), __builtin_fmaf(p0.x, a[j].y, p0.y * a[j].x),
                                __builtin_fmaf(p1.x, a[j].z, -p1.y * a[j].w), __builtin_fmaf(p1.x, a[j].w, p1.y * a[j].z)};
                    __builtin_nontemporal_store(y, reinterpret_cast<v4f *>(s + i));
                }
            }
        }
    } // namespace

    void launch_stage_table64(hipStream_t st, c64 *state, int n, const Subspace &tsub, const float *table)
    {
        const int m = tsub.nops;
        if (m < 11 || m > 14 || n <= m)
            throw std::runtime_error("stage table: tile bits must be 11..14 and below the state size");
        for (int k = 0; k < m; ++k)
            if (tsub.ops[k] < 0 || tsub.ops[k] >= n || (k > 0 && tsub.ops[k] <= tsub.ops[k - 1]))
                throw std::runtime_error("stage table: tile qubits must be ascending and < n");
        for (int b = 0; b < 4; ++b)
            if (tsub.ops[b] != b)
                throw std::runtime_error("stage table: tile qubits must include qubits 0..3");
        if (state == nullptr || table == nullptr)
            throw std::runtime_error("stage table: null pointer");
        const uint64_t count = uint64_t(1) << n;
        const uint64_t per_block = uint64_t(kTable64Threads) * kTable64Per * 2;
        if (count % per_block)
            throw std::runtime_error("stage table: state smaller than one chunk");
        const uint64_t chunks = count / per_block;
        const uint64_t grid = chunks < 4096 ? chunks : 4096;
        if (m == 11)
            hipLaunchKernelGGL(k_stage_table64<11>, dim3(grid), dim3(kTable64Threads), 0, st, state, count, tsub, table);
        else if (m == 12)
            hipLaunchKernelGGL(k_stage_table64<12>, dim3(grid), dim3(kTable64Threads), 0, st, state, count, tsub, table);
        else if (m == 13)
            hipLaunchKernelGGL(k_stage_table64<13>, dim3(grid), dim3(kTable64Threads), 0, st, state, count, tsub, table);
        else
            hipLaunchKernelGGL(k_stage_table64<14>, dim3(grid), dim3(kTable64Threads), 0, st, state, count, tsub, table);
        check(hipGetLastError(), "stage table kernel");
    }

    void launch_block128(hipStream_t st, c128 *state, int n, const Subspace &sub, const Block128Op &op, bool staged)
    {
        if (sub.nops > n || n > kMaxQubits)
            throw std::runtime_error("block128: operand count out of range");
        for (int k = 0; k < sub.nops; ++k)
            if (sub.ops[k] < 0 || sub.ops[k] >= n || (k > 0 && sub.ops[k] <= sub.ops[k - 1]))
                throw std::runtime_error("block128: operands must be ascending and < n");
        if (op.M == nullptr)
            throw std::runtime_error("block128: matrix pointer is NULL");
        const uint64_t count = uint64_t(1) << (n - sub.nops);
        if (staged)
        {
            if (sub.nops < 10)
                throw std::runtime_error("block128: staged form needs 10 tile bits");
            for (int b = 0; b < 10; ++b)
                if (n < 64 && (op.gw[b] >> n) != 0)
                    throw std::runtime_error("block128: tile bit outside the state");
            uint64_t grid = (count + kWaves - 1) / kWaves;
            if (grid > 4096)
                grid = 4096;
            hipLaunchKernelGGL(k_block128_lds, dim3(grid), dim3(kThreads), 0, st, state, count, sub, op);
        }
        else
        {
            if (sub.nops < 4)
                throw std::runtime_error("block128: needs 4 block qubits");
            for (int e = 0; e < 16; ++e)
                if (n < 64 && (op.off[e] >> n) != 0)
                    throw std::runtime_error("block128: offset outside the state");
            uint64_t grid = (count + kThreads - 1) / kThreads;
            if (grid > 4096)
                grid = 4096;
            hipLaunchKernelGGL(k_block128, dim3(grid), dim3(kThreads), 0, st, state, count, sub, op);
        }
        check(hipGetLastError(), "block128 kernel");
    }


    void check_stage_task(const StageTask &t, int n, int kq, int m)
    {
        if (kq != 4 && kq != 5)
            throw std::runtime_error("stage: kq must be 4 or 5");
        if (m < 11 || m > 14)
            throw std::runtime_error("stage: tile bits must be 11, 12, 13 or 14");
        if (n <= m || t.tsub.nops != m || n > kMaxQubits)
            throw std::runtime_error("stage: the tile must be a proper part of the state");
        for (int k = 0; k < t.tsub.nops; ++k)
            if (t.tsub.ops[k] < 0 || t.tsub.ops[k] >= n || (k > 0 && t.tsub.ops[k] <= t.tsub.ops[k - 1]))
                throw std::runtime_error("stage: tile qubits must be ascending and < n");
        for (int b = 0; b < 4; ++b)
            if (t.tsub.ops[b] != b)
                throw std::runtime_error("stage: tile qubits must include qubits 0..3");
        for (int b = 0; b < m; ++b)
            if (t.gw[b] != (uint64_t(1) << t.tsub.ops[b]))
                throw std::runtime_error("stage: tile weights do not match the tile qubits");
        if (t.tsub.set_bits != 0)
            throw std::runtime_error("stage: tiles take no controls");
        if (t.nblocks < 1 || t.state == nullptr || t.blocks % 4 != 0)
            throw std::runtime_error("stage: task without blocks or state, or misaligned descriptors");
        if ((t.flags & kStageTrailTable) && t.trail % 4 != 0)
            throw std::runtime_error("stage: misaligned trailing table");
    }

    void launch_stage(hipStream_t st, int kq, int m, bool x3, const float *base, const StageTask *tasks, int ntasks, uint64_t total_tiles)
    {
        if (kq != 4 && kq != 5)
            throw std::runtime_error("stage: kq must be 4 or 5");
        if (m < 11 || m > 14)
            throw std::runtime_error("stage: tile bits must be 11, 12, 13 or 14");
        if (base == nullptr || tasks == nullptr || ntasks < 1 || total_tiles < 1)
            throw std::runtime_error("stage: no tasks");
        // one workgroup per tile: the load -> blocks -> store phases of the workgroups on a CU
        // overlap, and no wave waits for its own stores before loading another tile
        // (measured on 32-qubit brick stages, 2^13 tiles: 19.5 ms per stage against 23.4 ms
        // with 8192 workgroups looping over the tiles; a persistent form that loads the next
        // tile during the last block measured 21.1 ms)
        const uint64_t grid = total_tiles < 0x7fffffffu ? total_tiles : 0x7fffffffu;
        // waves per workgroup by tile size: smaller tiles, fewer waves, more workgroups per CU
        auto go = [&](auto kern, int waves) { hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * waves), 0, st, base, tasks, ntasks, total_tiles); };
        constexpr int W13 = QSV_STAGE_WAVES_13, W12 = QSV_STAGE_WAVES_12, W11 = QSV_STAGE_WAVES_11;
        constexpr int X13 = QSV_STAGE_WAVES_13_X3, X12 = QSV_STAGE_WAVES_12_X3;
        constexpr int W14 = QSV_STAGE_WAVES_14, X14 = QSV_STAGE_WAVES_14_X3, X11 = QSV_STAGE_WAVES_11_X3;
        if (kq == 5 && m == 14)
            x3 ? go(k_stage<5, 14, true, X14>, X14) : go(k_stage<5, 14, false, W14>, W14);
        else if (m == 14)
            x3 ? go(k_stage<4, 14, true, X14>, X14) : go(k_stage<4, 14, false, W14>, W14);
        else if (kq == 5 && m == 13)
            x3 ? go(k_stage<5, 13, true, X13>, X13) : go(k_stage<5, 13, false, W13>, W13);
        else if (kq == 5 && m == 12)
            x3 ? go(k_stage<5, 12, true, X12>, X12) : go(k_stage<5, 12, false, W12>, W12);
        else if (kq == 5)
            x3 ? go(k_stage<5, 11, true, X11>, X11) : go(k_stage<5, 11, false, W11>, W11);
        else if (m == 13)
            x3 ? go(k_stage<4, 13, true, X13>, X13) : go(k_stage<4, 13, false, W13>, W13);
        else if (m == 12)
            x3 ? go(k_stage<4, 12, true, X12>, X12) : go(k_stage<4, 12, false, W12>, W12);
        else
            x3 ? go(k_stage<4, 11, true, X11>, X11) : go(k_stage<4, 11, false, W11>, W11);
        check(hipGetLastError(), "stage kernel");
    }

    bool stage_stream_supported(int kq, int m, int n) { return (kq == 4 || kq == 5) && m == 13 && n >= m + 9 && n <= 33; }

    void launch_stage_stream(hipStream_t st, int kq, int m, bool x3, const float *base, const StageTask *task, uint64_t total_tiles)
    {
        if (base == nullptr || task == nullptr || total_tiles < 1 || (total_tiles & (total_tiles - 1)))
            throw std::runtime_error("stage stream: no task, or a tile count that is not a power of two");
        if (!stage_stream_supported(kq, m, m + 63 - __builtin_clzll(total_tiles)))
            throw std::runtime_error("stage stream: kq must be 4 or 5, the tile 2^13 amplitudes and the state 22 to 33 qubits");
        int dev = 0, cus = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        // one resident workgroup per CU (2 x 64 KiB of LDS); every workgroup gets a tile
        const uint64_t grid = std::min<uint64_t>(total_tiles, uint64_t(std::max(cus, 1)));
        auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kStreamWaves), 0, st, base, task, total_tiles); };
        if (kq == 5)
            x3 ? go(k_stage_stream<5, 13, true>) : go(k_stage_stream<5, 13, false>);
        else
            x3 ? go(k_stage_stream<4, 13, true>) : go(k_stage_stream<4, 13, false>);
        check(hipGetLastError(), "stage stream kernel");
    }


    void launch_stage128(hipStream_t st, c128 *state, int n, const Subspace &tsub, const Stage128Op &op)
    {
        if (op.m < 11 || op.m > 13)
            throw std::runtime_error("stage128: tile bits must be 11, 12 or 13");
        if (op.kq != 3 && op.kq != 4)
            throw std::runtime_error("stage128: block qubits must be 3 or 4");
        if (n <= op.m || tsub.nops != op.m)
            throw std::runtime_error("stage128: the tile must be a proper part of the state");
        for (int k = 0; k < tsub.nops; ++k)
            if (tsub.ops[k] < 0 || tsub.ops[k] >= n || (k > 0 && tsub.ops[k] <= tsub.ops[k - 1]))
                throw std::runtime_error("stage128: tile qubits must be ascending and < n");
        for (int b = 0; b < 4; ++b)
            if (tsub.ops[b] != b)
                throw std::runtime_error("stage128: tile qubits must include qubits 0..3");
        for (int b = 0; b < op.m; ++b)
            if (op.gw[b] != (uint64_t(1) << tsub.ops[b]))
                throw std::runtime_error("stage128: tile weights do not match the tile qubits");
        if (tsub.set_bits != 0)
            throw std::runtime_error("stage128: tiles take no controls");
        const uint64_t ntiles = uint64_t(1) << (n - op.m);
        if (op.table != nullptr)
        {
            // a streaming pass in amplitude order; 2^n is a multiple of the chunk (n > 13)
            const uint64_t count = uint64_t(1) << n;
            if (count % (uint64_t(kTableThreads) * kTablePer) != 0)
                throw std::runtime_error("stage128 table: state smaller than one chunk");
            const uint64_t chunks = count / (uint64_t(kTableThreads) * kTablePer);
            const uint64_t grid = chunks < 4096 ? chunks : 4096;
            if (op.m == 11)
                hipLaunchKernelGGL(k_stage128_table<11>, dim3(grid), dim3(kTableThreads), 0, st, state, count, tsub, op.table);
            else if (op.m == 12)
                hipLaunchKernelGGL(k_stage128_table<12>, dim3(grid), dim3(kTableThreads), 0, st, state, count, tsub, op.table);
            else
                hipLaunchKernelGGL(k_stage128_table<13>, dim3(grid), dim3(kTableThreads), 0, st, state, count, tsub, op.table);
            check(hipGetLastError(), "stage128 table kernel");
            return;
        }
        if (op.nblocks < 1 || op.blocks == nullptr || op.M == nullptr)
            throw std::runtime_error("stage128: no blocks");
        const uint64_t grid = ntiles < 8192 ? ntiles : 8192;
        auto go = [&](auto kern, int threads) { hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, st, state, ntiles, tsub, op); };
        if (op.m == 11)
            op.kq == 4 ? go(k_stage128<4, 11>, stage128_threads<4, 11>()) : go(k_stage128<3, 11>, stage128_threads<3, 11>());
        else if (op.m == 12)
            op.kq == 4 ? go(k_stage128<4, 12>, stage128_threads<4, 12>()) : go(k_stage128<3, 12>, stage128_threads<3, 12>());
        else
            op.kq == 4 ? go(k_stage128<4, 13>, stage128_threads<4, 13>()) : go(k_stage128<3, 13>, stage128_threads<3, 13>());
        check(hipGetLastError(), "stage128 kernel");
    }

    void launch_block_lds(hipStream_t st, c64 *state, int n, const Subspace &tsub, const BlockLdsOp &op)
    {
        if (op.kq != 4 && op.kq != 5)
            throw std::runtime_error("block: kq must be 4 or 5");
        const int tile_bits = op.kq + 5;
        if (tsub.nops < tile_bits || tsub.nops > n)
            throw std::runtime_error("block: operand count out of range");
        for (int k = 0; k < tsub.nops; ++k)
            if (tsub.ops[k] < 0 || tsub.ops[k] >= n || (k > 0 && tsub.ops[k] <= tsub.ops[k - 1]))
                throw std::runtime_error("block: operands must be ascending and < n");
        for (int b = 0; b < tile_bits; ++b)
            if (n < 64 && (op.gw[b] >> n) != 0)
                throw std::runtime_error("block: tile bit outside the state");
        for (int b = tile_bits; b < 10; ++b)
            if (op.gw[b] != 0 || op.lw[b] != 0)
                throw std::runtime_error("block: unused tile bits must have zero weight");
        if (op.R == nullptr)
            throw std::runtime_error("block: matrix pointer is NULL");
        const uint64_t ntiles = uint64_t(1) << (n - tsub.nops);
        uint64_t grid = (ntiles + kWaves - 1) / kWaves;
        if (grid > 4096)
            grid = 4096;
        if (op.kq == 5 && op.x3)
            hipLaunchKernelGGL((k_block_lds<5, true>), dim3(grid), dim3(kThreads), 0, st, state, ntiles, tsub, op);
        else if (op.kq == 5)
            hipLaunchKernelGGL((k_block_lds<5, false>), dim3(grid), dim3(kThreads), 0, st, state, ntiles, tsub, op);
        else if (op.x3)
            hipLaunchKernelGGL((k_block_lds<4, true>), dim3(grid), dim3(kThreads), 0, st, state, ntiles, tsub, op);
        else
            hipLaunchKernelGGL((k_block_lds<4, false>), dim3(grid), dim3(kThreads), 0, st, state, ntiles, tsub, op);
        check(hipGetLastError(), "block kernel");
    }

    void launch_block(hipStream_t st, c64 *state, int n, const Subspace &sub, const BlockOp &op)
    {
        if (op.kq != 4 && op.kq != 5)
            throw std::runtime_error("block: kq must be 4 or 5");
        if (sub.nops < op.kq || sub.nops > n)
            throw std::runtime_error("block: operand count out of range");
        for (int k = 0; k < sub.nops; ++k)
            if (sub.ops[k] < 0 || sub.ops[k] >= n || (k > 0 && sub.ops[k] <= sub.ops[k - 1]))
                throw std::runtime_error("block: operands must be ascending and < n");
        for (int e = 0; e < (1 << op.kq); ++e)
            if (n < 64 && (op.off[e] >> n) != 0)
                throw std::runtime_error("block: offset outside the state");
        if (op.R == nullptr)
            throw std::runtime_error("block: matrix pointer is NULL");
        const uint64_t ncols = uint64_t(1) << (n - sub.nops);
        const uint64_t tiles = (ncols + 31) / 32;
        uint64_t grid = (tiles + kWaves - 1) / kWaves;
        if (grid > 4096)
            grid = 4096;
        if (op.kq == 5 && op.x3)
            hipLaunchKernelGGL((k_block<5, true>), dim3(grid), dim3(kThreads), 0, st, state, ncols, sub, op);
        else if (op.kq == 5)
            hipLaunchKernelGGL((k_block<5, false>), dim3(grid), dim3(kThreads), 0, st, state, ncols, sub, op);
        else if (op.x3)
            hipLaunchKernelGGL((k_block<4, true>), dim3(grid), dim3(kThreads), 0, st, state, ncols, sub, op);
        else
            hipLaunchKernelGGL((k_block<4, false>), dim3(grid), dim3(kThreads), 0, st, state, ncols, sub, op);
        check(hipGetLastError(), "block kernel");
    }
} // namespace qsv
