// Batching form: many independent small circuits on one GPU.
//
//  - k_batch_lds: persistent kernel.  Each workgroup takes circuits from a global work
//    counter, keeps the whole complex64 state in LDS (up to 2^14 amplitudes = 128 KiB),
//    applies the circuit's ops with workgroup barriers between ops, then reduces <Z_q>
//    for every qubit in FP64 and writes it out.  The loop exit is uniform per workgroup
//    (the taken index is broadcast through LDS), so every barrier is reached by all threads.
//  - k_zsum_partial / k_zsum_finish: <Z_q> for all qubits of a state in HBM in one pass
//    (per-thread FP64 accumulators, fixed-order trees), for circuits run out of LDS.
//  - k_zero_counter: resets the work counter inside a graph (no memset nodes).
#include "batch_kernels.hpp"

#include <stdexcept>
#include <string>

namespace qsv
{
    namespace
    {
        typedef float v2f __attribute__((ext_vector_type(2)));

        __device__ __forceinline__ v2f cm(v2f m, v2f a) { return v2f{m.x * a.x - m.y * a.y, m.x * a.y + m.y * a.x}; }

        template <int BS>
        __global__ __launch_bounds__(BS) void k_batch_lds(const BatchCircuit *__restrict__ circs, int ncirc, const BatchOp *__restrict__ ops, double *__restrict__ out, unsigned *__restrict__ counter)
        {
            extern __shared__ __attribute__((aligned(16))) v2f st[];
            __shared__ int taken;
            __shared__ double red[BS];
            for (;;)
            {
                if (threadIdx.x == 0)
                    taken = (int)atomicAdd(counter, 1u);
                __syncthreads();
                const int ci = taken;
                __syncthreads();
                if (ci >= ncirc)
                    break;
                const BatchCircuit C = circs[ci];
                const int n = C.n;
                const int N = 1 << n;
                for (int i = threadIdx.x; i < N; i += BS)
                    st[i] = v2f{i == 0 ? 1.f : 0.f, 0.f};
                __syncthreads();
                for (int k = 0; k < C.nops; ++k)
                {
                    const BatchOp op = ops[C.op_offset + k];
                    const int t = op.target;
                    const unsigned low = (1u << t) - 1u;
                    const v2f m00{op.m[0], op.m[1]}, m01{op.m[2], op.m[3]}, m10{op.m[4], op.m[5]}, m11{op.m[6], op.m[7]};
                    for (int p = threadIdx.x; p < N / 2; p += BS)
                    {
                        const unsigned i0 = (unsigned(p) & low) | ((unsigned(p) & ~low) << 1);
                        if (op.control >= 0 && !((i0 >> op.control) & 1u))
                            continue;
                        const unsigned i1 = i0 | (1u << t);
                        const v2f a = st[i0], b = st[i1];
                        st[i0] = cm(m00, a) + cm(m01, b);
                        st[i1] = cm(m10, a) + cm(m11, b);
                    }
                    __syncthreads();
                }
                // <Z_q>: per-thread FP64 sums in index order, then a tree over threads
                double acc[kBatchMaxLdsQubits];
#pragma unroll
                for (int q = 0; q < kBatchMaxLdsQubits; ++q)
                    acc[q] = 0.0;
                for (int i = threadIdx.x; i < N; i += BS)
                {
                    const v2f a = st[i];
                    const double pr = __builtin_fma(double(a.x), double(a.x), double(a.y) * double(a.y));
#pragma unroll
                    for (int q = 0; q < kBatchMaxLdsQubits; ++q)
                        acc[q] += ((i >> q) & 1) ? -pr : pr;
                }
#pragma unroll
                for (int q = 0; q < kBatchMaxLdsQubits; ++q)
                {
                    if (q < n) // uniform across the workgroup
                    {
                        red[threadIdx.x] = acc[q];
                        __syncthreads();
                        for (int s = BS / 2; s >= 1; s /= 2)
                        {
                            if (threadIdx.x < s)
                                red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
                            __syncthreads();
                        }
                        if (threadIdx.x == 0)
                            out[C.out_offset + q] = red[0];
                        __syncthreads();
                    }
                }
            }
        }

        constexpr int kZBlock = 256;
        constexpr int kZMax = 32;

        // Partial sums of workgroup `blk` of `nblk` over a state of n qubits (one block of
        // the grid-stride loop; the same order whatever the launch shape around it).
        __device__ __forceinline__ void zsum_partial_block(const c64 *__restrict__ s, int n, uint64_t blk, uint64_t nblk, double *__restrict__ part, double *red)
        {
            double acc[kZMax];
#pragma unroll
            for (int q = 0; q < kZMax; ++q)
                acc[q] = 0.0;
            const uint64_t N = uint64_t(1) << n;
            for (uint64_t i = blk * kZBlock + threadIdx.x; i < N; i += nblk * kZBlock)
            {
                const v2f a = *reinterpret_cast<const v2f *>(s + i);
                const double pr = __builtin_fma(double(a.x), double(a.x), double(a.y) * double(a.y));
#pragma unroll
                for (int q = 0; q < kZMax; ++q)
                    acc[q] += ((i >> q) & 1) ? -pr : pr;
            }
#pragma unroll
            for (int q = 0; q < kZMax; ++q)
            {
                if (q < n) // uniform across the workgroup
                {
                    red[threadIdx.x] = acc[q];
                    __syncthreads();
                    for (int st = kZBlock / 2; st >= 1; st /= 2)
                    {
                        if (threadIdx.x < st)
                            red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + st];
                        __syncthreads();
                    }
                    if (threadIdx.x == 0)
                        part[blk * kZMax + q] = red[0];
                    __syncthreads();
                }
            }
        }

        __global__ __launch_bounds__(kZBlock) void k_zsum_partial(const c64 *__restrict__ s, int n, double *__restrict__ part)
        {
            __shared__ double red[kZBlock];
            zsum_partial_block(s, n, blockIdx.x, gridDim.x, part, red);
        }

        // The task of workgroup g: the last task whose begin (read by `begin`) is <= g.
        template <class F>
        __device__ __forceinline__ int find_task(int ntasks, uint32_t g, F begin)
        {
            int lo = 0, hi = ntasks - 1;
            while (lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if (begin(mid) <= g)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            return lo;
        }

        __global__ __launch_bounds__(kZBlock) void k_zsum_partial_tasks(const ZsumTask *__restrict__ tasks, int ntasks)
        {
            __shared__ double red[kZBlock];
            const int t = find_task(ntasks, blockIdx.x, [&](int i) { return tasks[i].block_begin; });
            const ZsumTask &T = tasks[t];
            zsum_partial_block(T.state, T.n, blockIdx.x - T.block_begin, T.nblocks, T.part, red);
        }

        // One workgroup per qubit: thread j sums partials j, j + 256, ... in order, then a
        // fixed tree over the threads (the order does not depend on scheduling).
        __device__ __forceinline__ void zsum_finish_qubit(const double *__restrict__ part, int nblocks, int q, double *__restrict__ out, double *red)
        {
            double acc = 0.0;
            for (int b = threadIdx.x; b < nblocks; b += kZBlock)
                acc += part[uint64_t(b) * kZMax + q];
            red[threadIdx.x] = acc;
            __syncthreads();
            for (int st = kZBlock / 2; st >= 1; st /= 2)
            {
                if (threadIdx.x < st)
                    red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + st];
                __syncthreads();
            }
            if (threadIdx.x == 0)
                out[q] = red[0];
        }

        __global__ __launch_bounds__(kZBlock) void k_zsum_finish(const double *__restrict__ part, int nblocks, int n, double *__restrict__ out)
        {
            __shared__ double red[kZBlock];
            zsum_finish_qubit(part, nblocks, blockIdx.x, out, red);
        }

        __global__ __launch_bounds__(kZBlock) void k_zsum_finish_tasks(const ZsumTask *__restrict__ tasks, int ntasks)
        {
            __shared__ double red[kZBlock];
            const int t = find_task(ntasks, blockIdx.x, [&](int i) { return tasks[i].qubit_begin; });
            const ZsumTask &T = tasks[t];
            zsum_finish_qubit(T.part, T.nblocks, blockIdx.x - T.qubit_begin, T.out, red);
        }

        __global__ void k_zero_counter(unsigned *c) { *c = 0u; }

        __global__ void k_fill_zero_state(c64 *s, uint64_t count)
        {
            uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
            for (; i < count; i += uint64_t(gridDim.x) * blockDim.x)
                *reinterpret_cast<v2f *>(s + i) = v2f{i == 0 ? 1.f : 0.f, 0.f};
        }
    } // namespace

    size_t batch_lds_bytes(int max_n) { return (size_t(1) << max_n) * sizeof(c64); }

    void batch_prepare(int block_size, int max_n)
    {
        const size_t lds = batch_lds_bytes(max_n);
        hipError_t e = hipSuccess;
        if (block_size == 256)
            e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_batch_lds<256>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        else if (block_size == 512)
            e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_batch_lds<512>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        else if (block_size == 1024)
            e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_batch_lds<1024>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        else
            throw std::runtime_error("batch: block size must be 256, 512 or 1024");
        check(e, "hipFuncSetAttribute");
    }

    int batch_lds_grid(int block_size, int max_n)
    {
        int dev = 0, cus = 0, per = 0;
        check(hipGetDevice(&dev), "hipGetDevice");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
        const size_t lds = batch_lds_bytes(max_n);
        hipError_t e;
        if (block_size == 256)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_lds<256>, block_size, lds);
        else if (block_size == 512)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_lds<512>, block_size, lds);
        else
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_batch_lds<1024>, block_size, lds);
        check(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
        if (per < 1)
            throw std::runtime_error("batch: LDS kernel does not fit on a CU with this state size");
        return cus * per;
    }

    void launch_batch_lds(hipStream_t s, int block_size, int grid, int max_n, const BatchCircuit *circs, int ncirc, const BatchOp *ops, double *out, unsigned *counter)
    {
        if (max_n > kBatchMaxLdsQubits || max_n < 1)
            throw std::runtime_error("batch: LDS circuits must have 1.." + std::to_string(kBatchMaxLdsQubits) + " qubits");
        const size_t lds = batch_lds_bytes(max_n);
        hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, counter);
        check(hipGetLastError(), "counter kernel");
        if (block_size == 256)
            hipLaunchKernelGGL(k_batch_lds<256>, dim3(grid), dim3(256), lds, s, circs, ncirc, ops, out, counter);
        else if (block_size == 512)
            hipLaunchKernelGGL(k_batch_lds<512>, dim3(grid), dim3(512), lds, s, circs, ncirc, ops, out, counter);
        else if (block_size == 1024)
            hipLaunchKernelGGL(k_batch_lds<1024>, dim3(grid), dim3(1024), lds, s, circs, ncirc, ops, out, counter);
        else
            throw std::runtime_error("batch: block size must be 256, 512 or 1024");
        check(hipGetLastError(), "batch kernel");
    }

    int zsum_blocks(int n)
    {
        const uint64_t N = uint64_t(1) << n;
        uint64_t g = (N + kZBlock * 16 - 1) / (kZBlock * 16);
        if (g > 1024)
            g = 1024;
        return (int)(g < 1 ? 1 : g);
    }

    void launch_zsum(hipStream_t s, const c64 *state, int n, double *partials, double *out)
    {
        if (n < 1 || n > kZMax)
            throw std::runtime_error("zsum: qubit count out of range");
        const int g = zsum_blocks(n);
        hipLaunchKernelGGL(k_zsum_partial, dim3(g), dim3(kZBlock), 0, s, state, n, partials);
        check(hipGetLastError(), "zsum kernel");
        hipLaunchKernelGGL(k_zsum_finish, dim3(n), dim3(kZBlock), 0, s, partials, g, n, out);
        check(hipGetLastError(), "zsum kernel");
    }

    void launch_fill_zero_c64(hipStream_t s, c64 *state, uint64_t count)
    {
        uint64_t g = (count + 255) / 256;
        if (g > 65536)
            g = 65536;
        hipLaunchKernelGGL(k_fill_zero_state, dim3(g), dim3(256), 0, s, state, count);
        check(hipGetLastError(), "fill kernel");
    }

    void launch_zsum_tasks(hipStream_t s, const ZsumTask *tasks, int ntasks, uint32_t total_blocks, uint32_t total_qubits)
    {
        if (tasks == nullptr || ntasks < 1 || total_blocks < 1 || total_qubits < 1)
            throw std::runtime_error("zsum: no tasks");
        hipLaunchKernelGGL(k_zsum_partial_tasks, dim3(total_blocks), dim3(kZBlock), 0, s, tasks, ntasks);
        check(hipGetLastError(), "zsum kernel");
        hipLaunchKernelGGL(k_zsum_finish_tasks, dim3(total_qubits), dim3(kZBlock), 0, s, tasks, ntasks);
        check(hipGetLastError(), "zsum kernel");
    }

} // namespace qsv
