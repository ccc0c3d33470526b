// Checks the stage planner (qcore::plan_stages) on seeded random gate lists, on the CPU:
// every gate lands in exactly one stage, stages keep the gates' order, a stage's gates lie
// inside its tile (m qubits, ascending, 0..3 included), a stage without a tile holds one
// gate wider than the block size, and any two gates whose order the plan swaps commute
// (disjoint qubits, or both diagonal without measured targets).  Then the phase tables of
// stages of diagonal gates (qcore::prepare_stage128 and prepare_stage with tables): entry
// idx against the product of the gates' own phases on the tile qubits' bits of idx, from
// the gate definitions (P(t) = diag(1, e^{it}), RZ(t) = diag(e^{-it/2}, e^{it/2}), Z, S, T;
// controls: the phase only where every control bit is 1); and the trailing tables of
// stages of dense gates followed by diagonal ones (the greedy blocks after the last dense
// block, applied as a table at the store): the LDS block count, the table's alignment and
// every entry against the trailing gates' phases.  Also the complex64 LDS tile layout
// (qsv::stage_lds_offset: a bijection onto 8-B slots that keeps amplitude pairs 2j, 2j + 1
// in one aligned 16-B slot and is XOR-linear over disjoint index bits), the swizzled
// offsets the complex64 stage descriptors carry, and the FP32 block matrices' layouts (the
// real form's [[Re, -Im], [Im, Re]] pattern for 4-qubit blocks; Re, Im and Re + Im for
// 5-qubit blocks).  Prints one line per case and "ok" at the end; exits 1 at the first
// violation.
// Usage: plan_check <cases> <seed>
#include "fused.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace qcore;

namespace
{
    uint64_t mix(uint64_t &s)
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }

    bool diagonal(const ResolvedGate &g)
    {
        const Kind k = g.spec.kind;
        return g.mmask == 0 && (k == Kind::GlobalPhase || k == Kind::PhaseOne || k == Kind::Diag1 || k == Kind::Diag2 || k == Kind::Identity);
    }

    std::vector<int> qubits_of(const ResolvedGate &g)
    {
        std::vector<int> q = g.tphys;
        q.insert(q.end(), g.pphys.begin(), g.pphys.end());
        q.insert(q.end(), g.nphys.begin(), g.nphys.end());
        return q;
    }

    [[noreturn]] void fail(const std::string &what, int c)
    {
        std::printf("FAIL case %d: %s\n", c, what.c_str());
        std::exit(1);
    }
} // namespace

// the complex64 stage kernels' LDS slot map on tiles of 2^m amplitudes
static void check_lds_offsets(uint64_t seed)
{
    for (int m = 11; m <= 14; ++m)
    {
        const uint32_t nt = 1u << m;
        std::vector<char> seen(nt, 0);
        for (uint32_t i = 0; i < nt; ++i)
        {
            const uint32_t off = qsv::stage_lds_offset(i);
            if (off % 8 != 0 || off / 8 >= nt || seen[off / 8]++)
                fail("stage_lds_offset is not a bijection onto the tile's 8-B slots", m);
            if ((i & 1) == 0 && (off % 16 != 0 || qsv::stage_lds_offset(i | 1) != off + 8))
                fail("stage_lds_offset splits an amplitude pair", m);
        }
        for (int r = 0; r < 1000; ++r)
        {
            const uint32_t a = uint32_t(mix(seed)) & (nt - 1), mask = uint32_t(mix(seed)) & (nt - 1);
            const uint32_t x = a & mask, y = a & ~mask;
            if (qsv::stage_lds_offset(x | y) != (qsv::stage_lds_offset(x) ^ qsv::stage_lds_offset(y)))
                fail("stage_lds_offset is not XOR-linear over disjoint bits", m);
        }
    }
    std::printf("lds offsets: m 11..14 ok\n");
}

// complex64 stage descriptors and FP32 block matrices of a prepared dense stage
static void check_c64_stage(const PreparedStage &p, const std::vector<float> &mats, int kb, int m, int c)
{
    const auto *desc = reinterpret_cast<const qsv::StageBlock *>(mats.data() + p.desc_offset);
    for (size_t b = 0; b < p.nblocks; ++b)
    {
        const qsv::StageBlock &sb = desc[b];
        for (int j = 0; j < m; ++j)
            if (sb.lds[j] != qsv::stage_lds_offset(uint32_t(sb.w[j])))
                fail("descriptor LDS offset differs from the swizzled weight", c);
        const float *R = mats.data() + sb.matrix;
        if (kb == 5)
        { // Re M, Im M, Re M + Im M (the sum formed in FP64, then rounded)
            double norm = 0;
            for (int i = 0; i < 1024; ++i)
            {
                const double re = R[i], im = R[1024 + i], sum = R[2048 + i];
                if (std::abs(sum - (re + im)) > 2.5e-7 * (std::abs(re) + std::abs(im)) + 1e-30)
                    fail("5-qubit FP32 matrix: third part is not Re + Im", c);
                norm += re * re + im * im;
            }
            if (std::abs(norm - 32.0) > 1e-4) // a unitary 32 x 32 matrix: sum of |entries|^2 = 32
                fail("5-qubit FP32 matrix: Re and Im parts are not those of a unitary", c);
        }
        else
        { // real form R[2r][2c] = R[2r+1][2c+1] = Re, R[2r+1][2c] = -R[2r][2c+1] = Im
            const int D = 2 << kb;
            for (int r = 0; r < D / 2; ++r)
                for (int q = 0; q < D / 2; ++q)
                    if (R[(2 * r) * D + 2 * q] != R[(2 * r + 1) * D + 2 * q + 1] ||
                        R[(2 * r + 1) * D + 2 * q] != -R[(2 * r) * D + 2 * q + 1])
                        fail("FP32 real-form block matrix pattern broken", c);
        }
    }
}

int main(int argc, char **argv)
{
    const int cases = argc > 1 ? std::atoi(argv[1]) : 50;
    uint64_t seed = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1;
    check_lds_offsets(seed);
    // gate names by target count; p, rz, rzz, z, s, t are diagonal, the rest not
    const char *one[] = {"h", "x", "z", "s", "t", "p", "rz", "rx", "u3", "sx"};
    const char *two[] = {"swap", "rzz", "rxx", "ecr", "iswap"};
    for (int c = 0; c < cases; ++c)
    {
        const int n = 14 + int(mix(seed) % 10);
        const int m = 11 + int(mix(seed) % 3);
        const int k = 3 + int(mix(seed) % 3);
        const int ngates = 20 + int(mix(seed) % 200);
        std::vector<ResolvedGate> gates;
        for (int i = 0; i < ngates; ++i)
        {
            ResolvedGate r;
            const int nt = (mix(seed) % 5 == 0) ? 2 : 1;
            const char *name = nt == 1 ? one[mix(seed) % 10] : two[mix(seed) % 5];
            double p[4] = {0.3, 0.7, 1.1, 0.2};
            r.spec = make_gate(name, p);
            std::vector<int> pool(n);
            for (int q = 0; q < n; ++q)
                pool[q] = q;
            for (int q = n - 1; q > 0; --q)
                std::swap(pool[q], pool[mix(seed) % (q + 1)]);
            int used = 0;
            for (int t = 0; t < nt; ++t)
                r.tphys.push_back(pool[used++]);
            const int nc = int(mix(seed) % 4); // 0..3 controls: some gates are wider than k
            for (int j = 0; j < nc; ++j)
                (mix(seed) & 1 ? r.pphys : r.nphys).push_back(pool[used++]);
            if (mix(seed) % 17 == 0)
                r.mmask = 1; // a measured target: stays in order
            gates.push_back(r);
        }
        for (int tiles : {kTilesGrow, kTilesWindow, kTilesAuto})
        {
            const auto plan = plan_stages(gates, n, m, k, tiles);
            std::vector<int> stage_of(gates.size(), -1);
            for (int s = 0; s < (int)plan.size(); ++s)
            {
                const auto &st = plan[s];
                if (st.gates.empty())
                    fail("empty stage", c);
                for (size_t i = 0; i < st.gates.size(); ++i)
                {
                    const int gi = st.gates[i];
                    if (gi < 0 || gi >= (int)gates.size() || stage_of[gi] != -1)
                        fail("gate missing, repeated or out of range", c);
                    stage_of[gi] = s;
                    if (i > 0 && st.gates[i] <= st.gates[i - 1])
                        fail("gates out of order inside a stage", c);
                }
                if (st.tile.empty())
                {
                    if (st.gates.size() != 1 || (int)qubits_of(gates[st.gates[0]]).size() <= k)
                        fail("a stage without a tile must hold one gate wider than k", c);
                    continue;
                }
                if ((int)st.tile.size() != m || !std::is_sorted(st.tile.begin(), st.tile.end()) ||
                    std::adjacent_find(st.tile.begin(), st.tile.end()) != st.tile.end() || st.tile.back() >= n)
                    fail("tile not m ascending qubits of the state", c);
                for (int q = 0; q < 4; ++q)
                    if (st.tile[q] != q)
                        fail("tile without qubits 0..3", c);
                for (int gi : st.gates)
                {
                    const auto qs = qubits_of(gates[gi]);
                    if ((int)qs.size() > k)
                        fail("gate wider than k inside a tiled stage", c);
                    for (int q : qs)
                        if (!std::binary_search(st.tile.begin(), st.tile.end(), q))
                            fail("gate qubit outside its stage's tile", c);
                }
            }
            for (size_t gi = 0; gi < gates.size(); ++gi)
                if (stage_of[gi] < 0)
                    fail("gate in no stage", c);
            // order swaps: a before b in the list but b in an earlier stage
            for (size_t a = 0; a < gates.size(); ++a)
                for (size_t b = a + 1; b < gates.size(); ++b)
                {
                    if (stage_of[b] >= stage_of[a])
                        continue;
                    const auto qa = qubits_of(gates[a]), qb = qubits_of(gates[b]);
                    bool share = false;
                    for (int q : qa)
                        share = share || std::find(qb.begin(), qb.end(), q) != qb.end();
                    if (share && !(diagonal(gates[a]) && diagonal(gates[b])))
                        fail("two gates that do not commute were reordered", c);
                }
            std::printf("case %d n %d m %d k %d gates %d tiles %d: %zu stages\n", c, n, m, k, ngates, tiles, plan.size());
        }
    }
    // phase tables of all-diagonal stages, on tiles of 2^11, 2^12 and 2^13
    for (int c = 0; c < cases; ++c)
    {
        const int n = 17 + int(mix(seed) % 6);
        const int m = 11 + c % 3;
        std::vector<int> tile = {0, 1, 2, 3};
        const int w0 = 4 + int(mix(seed) % (n - m + 1)); // a window of m - 4 qubits above qubit 3
        for (int q = w0; q < w0 + m - 4; ++q)
            tile.push_back(q);
        const char *names[] = {"p", "rz", "z", "s", "t"};
        std::vector<ResolvedGate> gates;
        std::vector<double> angle;
        const int ng = 1 + int(mix(seed) % 40);
        for (int i = 0; i < ng; ++i)
        {
            ResolvedGate r;
            const int kind = int(mix(seed) % 5);
            double prm[4] = {6.283 * double(mix(seed) >> 11) * 0x1.0p-53, 0, 0, 0};
            r.spec = make_gate(names[kind], prm);
            std::vector<int> pool = tile;
            for (int q = m - 1; q > 0; --q)
                std::swap(pool[q], pool[mix(seed) % (q + 1)]);
            r.tphys = {pool[0]};
            const int nc = kind == 1 ? 0 : int(mix(seed) % 3); // rz without controls
            for (int j = 0; j < nc; ++j)
                r.pphys.push_back(pool[1 + j]);
            gates.push_back(r);
            angle.push_back(prm[0]);
        }
        PlannedStage st;
        st.tile = tile;
        for (int i = 0; i < ng; ++i)
            st.gates.push_back(i);
        // the phase of gate i on the basis state whose tile bits are those of idx
        auto expect = [&](int idx) {
            std::complex<double> v(1.0, 0.0);
            auto bit = [&](int q) { return (idx >> (std::find(tile.begin(), tile.end(), q) - tile.begin())) & 1; };
            for (int i = 0; i < ng; ++i)
            {
                const ResolvedGate &g = gates[i];
                bool on = true;
                for (int q : g.pphys)
                    on = on && bit(q);
                if (!on)
                    continue;
                const int t = bit(g.tphys[0]);
                const double a = angle[i];
                const std::string nm = g.spec.name;
                if (nm == "rz")
                    v *= std::polar(1.0, t ? a / 2 : -a / 2);
                else if (t)
                    v *= std::polar(1.0, nm == "p" ? a : nm == "z" ? M_PI : nm == "s" ? M_PI / 2 : M_PI / 4);
            }
            return v;
        };
        for (int kq : {3, 4})
        {
            std::vector<double> mats64;
            const auto p128 = prepare_stage128(gates, st, n, mats64, kq);
            if (p128.table_offset == SIZE_MAX || p128.nblocks != 0)
                fail("complex128 stage of diagonal gates without a table", c);
            std::vector<float> mats;
            const auto p64 = prepare_stage(gates, st, kq + 1, m, n, mats, false, true);
            if (p64.table_offset == SIZE_MAX)
                fail("complex64 stage of diagonal gates without a table", c);
            double e128 = 0, e64 = 0;
            for (int idx = 0; idx < (1 << m); ++idx)
            {
                const std::complex<double> want = expect(idx);
                const std::complex<double> got(mats64[p128.table_offset + 2 * idx], mats64[p128.table_offset + 2 * idx + 1]);
                const std::complex<double> got64(mats[p64.table_offset + 2 * idx], mats[p64.table_offset + 2 * idx + 1]);
                e128 = std::max(e128, std::abs(got - want));
                e64 = std::max(e64, std::abs(got64 - want));
            }
            if (e128 > 1e-12 || e64 > 1e-6)
                fail("phase table differs from the gates' phases", c);
            std::printf("table case %d n %d m %d window %d gates %d kq %d: max|d| %.1e (c128) %.1e (c64)\n", c, n, m, w0, ng, kq, e128, e64);
        }
    }
    // trailing tables: dense gates, then a run of diagonal ones; the greedy blocks up to the
    // last one with a dense gate run through LDS (nblocks), the diagonal blocks after it
    // become one table applied at the store (trail_offset, 4-float aligned for complex64)
    for (int c = 0; c < cases; ++c)
    {
        const int n = 17 + int(mix(seed) % 6);
        const int m = 11 + c % 3;
        std::vector<int> tile = {0, 1, 2, 3};
        const int w0 = 4 + int(mix(seed) % (n - m + 1));
        for (int q = w0; q < w0 + m - 4; ++q)
            tile.push_back(q);
        const char *names[] = {"p", "rz", "z", "s", "t"};
        std::vector<ResolvedGate> gates;
        std::vector<double> angle;
        auto pick = [&]() {
            std::vector<int> pool = tile;
            for (int q = m - 1; q > 0; --q)
                std::swap(pool[q], pool[mix(seed) % (q + 1)]);
            return pool;
        };
        const int ndense = 1 + int(mix(seed) % 6);
        for (int i = 0; i < ndense; ++i)
        {
            ResolvedGate r;
            double prm[4] = {0.4 + 0.1 * i, 1.3, -0.7, 0};
            const auto pool = pick();
            r.spec = make_gate(i % 2 ? "x" : "u3", prm);
            r.tphys = {pool[0]};
            if (i % 2)
                r.pphys = {pool[1]};
            gates.push_back(r);
            angle.push_back(0.0);
        }
        const int ndiag = 1 + int(mix(seed) % 30);
        for (int i = 0; i < ndiag; ++i)
        {
            ResolvedGate r;
            const int kind = int(mix(seed) % 5);
            double prm[4] = {6.283 * double(mix(seed) >> 11) * 0x1.0p-53, 0, 0, 0};
            r.spec = make_gate(names[kind], prm);
            const auto pool = pick();
            r.tphys = {pool[0]};
            const int nc = kind == 1 ? 0 : int(mix(seed) % 3);
            for (int j = 0; j < nc; ++j)
                r.pphys.push_back(pool[1 + j]);
            gates.push_back(r);
            angle.push_back(prm[0]);
        }
        PlannedStage st;
        st.tile = tile;
        for (int i = 0; i < (int)gates.size(); ++i)
            st.gates.push_back(i);
        // product of the phases of the listed (diagonal) gates on tile index idx
        auto expect = [&](int idx, const std::vector<int> &which) {
            std::complex<double> v(1.0, 0.0);
            auto bit = [&](int q) { return (idx >> (std::find(tile.begin(), tile.end(), q) - tile.begin())) & 1; };
            for (int i : which)
            {
                const ResolvedGate &g = gates[i];
                if (!diagonal(g))
                    fail("a dense gate in the trailing table", c);
                bool on = true;
                for (int q : g.pphys)
                    on = on && bit(q);
                if (!on)
                    continue;
                const int t = bit(g.tphys[0]);
                const double a = angle[i];
                const std::string nm = g.spec.name;
                if (nm == "rz")
                    v *= std::polar(1.0, t ? a / 2 : -a / 2);
                else if (t)
                    v *= std::polar(1.0, nm == "p" ? a : nm == "z" ? M_PI : nm == "s" ? M_PI / 2 : M_PI / 4);
            }
            return v;
        };
        for (int kq : {3, 4})
            for (bool f64 : {true, false})
            {
                const int kb = f64 ? kq : kq + 1; // complex64 stages take 4- or 5-qubit blocks
                const auto plan = plan_greedy(gates, kb);
                const int lds = f64 ? stage128_lds_blocks(gates, kb) : stage64_lds_blocks(gates, kb);
                std::vector<int> trail;
                for (int b = lds; b < (int)plan.size(); ++b)
                    trail.insert(trail.end(), plan[b].gates.begin(), plan[b].gates.end());
                if (lds < 1)
                    fail("a stage with dense gates has no LDS block", c);
                size_t toff, nblocks;
                std::vector<double> mats64;
                std::vector<float> mats;
                if (f64)
                {
                    const auto p = prepare_stage128(gates, st, n, mats64, kb);
                    toff = p.trail_offset;
                    nblocks = p.nblocks;
                    if (p.table_offset != SIZE_MAX)
                        fail("a stage with dense gates ran as a table", c);
                }
                else
                {
                    const auto p = prepare_stage(gates, st, kb, m, n, mats, false, true);
                    toff = p.trail_offset;
                    nblocks = p.nblocks;
                    if (p.table_offset != SIZE_MAX)
                        fail("a stage with dense gates ran as a table", c);
                    if (toff != SIZE_MAX && toff % 4 != 0)
                        fail("complex64 trailing table not 4-float aligned", c);
                    check_c64_stage(p, mats, kb, m, c);
                }
                if ((int)nblocks != lds)
                    fail("LDS block count differs from the blocks up to the last dense one", c);
                if (trail.empty() != (toff == SIZE_MAX))
                    fail("trailing table present without trailing diagonal blocks, or missing", c);
                double err = 0;
                if (!trail.empty())
                    for (int idx = 0; idx < (1 << m); ++idx)
                    {
                        const std::complex<double> want = expect(idx, trail);
                        const std::complex<double> got = f64 ? std::complex<double>(mats64[toff + 2 * idx], mats64[toff + 2 * idx + 1])
                                                             : std::complex<double>(mats[toff + 2 * idx], mats[toff + 2 * idx + 1]);
                        err = std::max(err, std::abs(got - want));
                    }
                if (err > (f64 ? 1e-12 : 1e-6))
                    fail("trailing table differs from the trailing gates' phases", c);
                std::printf("trail case %d n %d dense %d diagonal %d k %d %s: %d LDS blocks, %zu trailing gates, max|d| %.1e\n",
                            c, n, ndense, ndiag, kb, f64 ? "c128" : "c64", lds, trail.size(), err);
            }
    }
    std::printf("ok\n");
    return 0;
}
