// Work orders of the 2x2 pair update over target positions on a 30-qubit complex128 state
// (16 GiB): the same arithmetic as the library's pair kernel (one pair per thread, 256
// threads per workgroup, non-temporal 16-B loads and stores), with the tiles of items
// visited in different orders.  Prints ms per launch (mean of 4 after one warm-up launch)
// per target and order.  Standalone: it allocates its own state and needs no library.
//
//   xcd      one tile per workgroup, each XCD a contiguous eighth of the tiles (the library)
//   noxcd    one tile per workgroup in blockIdx order
//   bfirst   as xcd, odd waves load the partner amplitude first
//   rev4     as xcd, the XCD-local tile index with its low 4 bits reversed
//   back     as xcd, odd XCDs walk their eighth backwards
//   g2k..    persistent grids of 2^11, 2^13, 2^15 workgroups (XCD order, grid stride), with
//            "+skew": each tile's first amplitudes loaded one tile ahead of their partners
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

typedef double v2d __attribute__((ext_vector_type(2)));

#define CHECK(x)                                                                                   \
    do                                                                                             \
    {                                                                                              \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess)                                                                      \
        {                                                                                          \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);             \
            std::exit(1);                                                                          \
        }                                                                                          \
    } while (0)

namespace
{
    struct Mat2
    {
        double m[8]; // (re, im) of m00, m01, m10, m11
    };

    __device__ __forceinline__ v2d cmul(double r, double i, v2d a) { return v2d{r * a.x - i * a.y, r * a.y + i * a.x}; }
    __device__ __forceinline__ v2d load(const v2d *p) { return __builtin_nontemporal_load(p); }
    __device__ __forceinline__ void store(v2d *p, v2d v) { __builtin_nontemporal_store(v, p); }

    // item -> amplitude index with a zero inserted at bit t
    __device__ __forceinline__ uint64_t pair_index(uint64_t item, uint64_t bit)
    {
        const uint64_t lo = item & (bit - 1);
        return lo | ((item ^ lo) << 1);
    }

    __device__ __forceinline__ void update(v2d *s, uint64_t idx, uint64_t bit, v2d a, v2d b, const Mat2 &op)
    {
        store(s + idx, cmul(op.m[0], op.m[1], a) + cmul(op.m[2], op.m[3], b));
        store(s + (idx | bit), cmul(op.m[4], op.m[5], a) + cmul(op.m[6], op.m[7], b));
    }

    // one tile of 256 items per workgroup; ORDER: 0 blockIdx, 1 xcd, 3 bfirst, 4 rev4, 5 back
    template <int ORDER>
    __global__ __launch_bounds__(256) void k_order(v2d *__restrict__ s, uint64_t ntiles, int t, Mat2 op)
    {
        const uint64_t nb = gridDim.x, per_x = nb >> 3, x = blockIdx.x & 7;
        uint64_t l = blockIdx.x >> 3, tile = blockIdx.x;
        if (ORDER != 0)
        {
            if (ORDER == 4)
                l = (l & ~uint64_t(15)) | ((l & 1) << 3) | ((l & 2) << 1) | ((l & 4) >> 1) | ((l & 8) >> 3);
            if (ORDER == 5 && (x & 1))
                l = per_x - 1 - l;
            tile = x * per_x + l;
        }
        if (tile >= ntiles)
            return;
        const uint64_t bit = uint64_t(1) << t;
        const uint64_t idx = pair_index(tile * 256 + threadIdx.x, bit);
        v2d a, b;
        if (ORDER == 3 && ((threadIdx.x >> 6) & 1))
        {
            b = load(s + (idx | bit));
            a = load(s + idx);
        }
        else
        {
            a = load(s + idx);
            b = load(s + (idx | bit));
        }
        update(s, idx, bit, a, b, op);
    }

    // persistent grid (a multiple of 8), XCD order, grid stride; SKEW: a of the next tile
    // loaded before this tile's update
    template <bool SKEW>
    __global__ __launch_bounds__(256) void k_persistent(v2d *__restrict__ s, uint64_t ntiles, int t, Mat2 op)
    {
        const uint64_t nb = gridDim.x;
        const uint64_t first = (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
        const uint64_t bit = uint64_t(1) << t;
        v2d next = {0.0, 0.0};
        if (SKEW && first < ntiles)
            next = load(s + pair_index(first * 256 + threadIdx.x, bit));
        for (uint64_t tile = first; tile < ntiles; tile += nb)
        {
            const uint64_t idx = pair_index(tile * 256 + threadIdx.x, bit);
            const v2d a = SKEW ? next : load(s + idx);
            const v2d b = load(s + (idx | bit));
            if (SKEW && tile + nb < ntiles)
                next = load(s + pair_index((tile + nb) * 256 + threadIdx.x, bit));
            update(s, idx, bit, a, b, op);
        }
    }

    template <class Launch>
    float time_ms(Launch launch)
    {
        hipEvent_t e0, e1;
        CHECK(hipEventCreate(&e0));
        CHECK(hipEventCreate(&e1));
        launch();
        CHECK(hipGetLastError());
        CHECK(hipEventRecord(e0));
        for (int r = 0; r < 4; ++r)
            launch();
        CHECK(hipEventRecord(e1));
        CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        CHECK(hipEventDestroy(e0));
        CHECK(hipEventDestroy(e1));
        return ms / 4;
    }
} // namespace

int main()
{
    constexpr int n = 30;
    const uint64_t amps = uint64_t(1) << n;
    const uint64_t ntiles = (amps / 2) / 256; // 2^21 tiles of 256 pairs
    v2d *s = nullptr;
    CHECK(hipMalloc(&s, amps * sizeof(v2d)));
    CHECK(hipMemset(s, 0, amps * sizeof(v2d)));
    const Mat2 op{{0.6, 0.0, 0.8, 0.0, 0.8, 0.0, -0.6, 0.0}};
    const dim3 full(static_cast<unsigned>(ntiles)), block(256);
    std::printf("target    xcd  noxcd bfirst   rev4   back    g2k g2k+skew  g8k g8k+skew  g32k g32k+skew (ms)\n");
    for (int t = 0; t < n; ++t)
    {
        const float o1 = time_ms([&] { hipLaunchKernelGGL(k_order<1>, full, block, 0, 0, s, ntiles, t, op); });
        const float o0 = time_ms([&] { hipLaunchKernelGGL(k_order<0>, full, block, 0, 0, s, ntiles, t, op); });
        const float o3 = time_ms([&] { hipLaunchKernelGGL(k_order<3>, full, block, 0, 0, s, ntiles, t, op); });
        const float o4 = time_ms([&] { hipLaunchKernelGGL(k_order<4>, full, block, 0, 0, s, ntiles, t, op); });
        const float o5 = time_ms([&] { hipLaunchKernelGGL(k_order<5>, full, block, 0, 0, s, ntiles, t, op); });
        float p[6];
        const unsigned grids[3] = {2048, 8192, 32768};
        for (int g = 0; g < 3; ++g)
        {
            p[2 * g] = time_ms([&] { hipLaunchKernelGGL(k_persistent<false>, dim3(grids[g]), block, 0, 0, s, ntiles, t, op); });
            p[2 * g + 1] = time_ms([&] { hipLaunchKernelGGL(k_persistent<true>, dim3(grids[g]), block, 0, 0, s, ntiles, t, op); });
        }
        std::printf("%2d   %6.3f %6.3f %6.3f %6.3f %6.3f %6.3f %6.3f %6.3f %6.3f %6.3f %6.3f\n", t, o1, o0, o3, o4, o5, p[0], p[1], p[2], p[3], p[4], p[5]);
        std::fflush(stdout);
    }
    CHECK(hipFree(s));
    return 0;
}
