// Checks the permutation pass planner (qsv::plan_permute_passes) on the CPU: for seeded
// random permutations and hand-picked ones (full reversal, single swaps, cycles through
// qubits 0..2), applying the passes in order moves every qubit where the permutation
// sends it, each pass is itself a permutation, moves at most `budget` qubits at or above
// `free_below`, and so fits one tile of the permutation kernel (budget + free_below tile
// bits).  Prints "ok" at the end; exits 1 at the first violation.
// Usage: perm_check <cases> <seed>
#include "sv_kernels.hpp"

#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>
#include <vector>

namespace
{
    uint64_t mix(uint64_t &s)
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }

    [[noreturn]] void fail(const std::string &what, int c)
    {
        std::printf("FAIL case %d: %s\n", c, what.c_str());
        std::exit(1);
    }

    void check(const std::vector<int> &dst, int budget, int low, int c)
    {
        const int n = (int)dst.size();
        const auto passes = qsv::plan_permute_passes(dst, budget, low);
        std::vector<int> at(n); // where the data of original qubit q is now
        std::iota(at.begin(), at.end(), 0);
        for (const auto &p : passes)
        {
            if ((int)p.size() != n)
                fail("pass size", c);
            std::vector<int> seen(n, 0);
            int moved = 0, cost = 0;
            for (int q = 0; q < n; ++q)
            {
                if (p[q] < 0 || p[q] >= n || seen[p[q]]++)
                    fail("pass is not a permutation", c);
                if (p[q] != q)
                {
                    ++moved;
                    cost += q >= low;
                }
            }
            if (moved == 0)
                fail("empty pass", c);
            if (cost > budget || moved > budget + low)
                fail("pass moves too many qubits: " + std::to_string(moved) + " (" + std::to_string(cost) + " high)", c);
            for (int q = 0; q < n; ++q)
                at[q] = p[at[q]];
        }
        for (int q = 0; q < n; ++q)
            if (at[q] != dst[q])
                fail("passes do not compose to the permutation", c);
        std::printf("case %d: n=%d passes=%zu\n", c, n, passes.size());
    }
} // namespace

int main(int argc, char **argv)
{
    const int cases = argc > 1 ? std::atoi(argv[1]) : 100;
    uint64_t seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1;
    int c = 0;
    for (int n : {2, 7, 13, 20, 30, 34})
    {
        std::vector<int> rev(n);
        for (int q = 0; q < n; ++q)
            rev[q] = n - 1 - q;
        check(rev, 6, 6, c++);
        std::vector<int> cyc(n);
        std::iota(cyc.begin(), cyc.end(), 0);
        if (n >= 3)
        {
            cyc[0] = 1;
            cyc[1] = 2;
            cyc[2] = 0;
            check(cyc, 6, 6, c++);
        }
        std::vector<int> sw(n);
        std::iota(sw.begin(), sw.end(), 0);
        std::swap(sw[0], sw[n - 1]);
        check(sw, 2, 6, c++);
    }
    for (int i = 0; i < cases; ++i)
    {
        const int n = 2 + int(mix(seed) % 33);
        std::vector<int> dst(n);
        std::iota(dst.begin(), dst.end(), 0);
        // a random permutation, or one that moves only a few qubits
        const int moves = (mix(seed) & 1) ? n : 2 + int(mix(seed) % 6);
        for (int k = 0; k < moves; ++k)
            std::swap(dst[mix(seed) % n], dst[mix(seed) % n]);
        const int budget = 2 + int(mix(seed) % 7);
        const int low = int(mix(seed) % 8);
        check(dst, budget, low, c++);
    }
    std::printf("ok\n");
    return 0;
}
