/*
 * qcs_oracle.cpp -- CPU restatement of the RQS-SVG plain method.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it (as oracle/_build/libqcs_oracle.so).  The product, libqcs.so, never
 * links or calls it.  It exports the same C ABI (include/qcs.h, include/qcs_ext.h) so that
 * one driver can run the same circuit on both.
 *
 * It restates, in plain C++17 with OpenMP and std::complex<double>, what the reference
 * computes (file:line into RQS-SVG):
 *   - amplitude index of a work item: insert a 0 bit at every sorted operand position,
 *     OR in the positive-control bits, target j = bit j of the local index
 *     (src/qcs.cu:156-376)                                          -> subspace_base()
 *   - the gate formulas (src/qcs.cu:378-1104)                       -> make_gate()
 *   - one in-place pass per gate over the controlled subspace; an input on the branch of
 *     a measured target that was not observed reads as zero (src/qcs.cu:1139-1174 for one
 *     target; for t >= 2 the intended per-bit rule is used, not the reference's
 *     `(bits & 2^j) == 1` test that is always false for j >= 1, src/qcs.cu:1216-1609)
 *                                                                   -> apply_gate()
 *   - measured, non-targeted qubits become implicit controls; a contradicting explicit
 *     control skips the gate; targets leave the measured lists (src/qcs.cu:2161-2204,
 *     2480-2484, 2679-2703)                                         -> operate()
 *   - measurement: cached result, {p0,p1} over the subspace consistent with earlier
 *     measurements, r = uniform_real_distribution(0, p0+p1)(mt19937_64), result = p0 < r,
 *     collapse of the other branch without renormalisation (src/qcs.cu:2545-2677,
 *     1106-1123)                                                    -> measure()
 *   - global->local swap bookkeeping of the mapping (src/qcs.cu:2206-2332), used here for
 *     the "local qubits" window                                     -> ensure_local()
 *   - mapping reset (src/qcs.cu:1860-1885), zero/sequential/flat/entangled states
 *     (src/qcs.cu:71-82, 2025-2095), logical-order save (src/qcs.cu:2494-2543), events
 *     (src/qcs.cu:2847-2897), ABI error convention (src/qcs.cu:2903-3006, 3383-3759).
 *
 * The reduction order for {p0,p1}, the probability prefix for sampling and the uniforms
 * are this project's definitions (include/qcs_ext.h); the GPU library implements the same
 * definitions so that, on bit-identical states, probabilities and outcomes are bit-exact.
 */
#include <qcs.h>
#include <qcs_ext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace oracle
{
    using cplx = std::complex<double>;

    static std::string fmt(const char *f, ...) __attribute__((format(printf, 1, 2)));
    static std::string fmt(const char *f, ...)
    {
        char buf[512];
        va_list ap;
        va_start(ap, f);
        vsnprintf(buf, sizeof buf, f, ap);
        va_end(ap);
        return buf;
    }

    static inline cplx mul_i(cplx z) { return cplx(-z.imag(), z.real()); }    // i*z, src/qcs.cu:56-59
    static inline cplx mul_mi(cplx z) { return cplx(z.imag(), -z.real()); }   // -i*z, src/qcs.cu:61-64
    static inline cplx mul_i_real(cplx z, double m) { return cplx(-m * z.imag(), m * z.real()); } // i*m*z, src/qcs.cu:66-69
    static inline cplx expi(double a) { return cplx(std::cos(a), std::sin(a)); }
    static const double SQRT1_2 = 0.70710678118654752440;

    /* One gate: number of targets and the in-place map of 2^t amplitudes.  `in` holds the
     * (masked) inputs; `out` arrives holding the current memory and the function writes
     * only the rows the reference's apply() writes. */
    struct Gate
    {
        int ntargets;
        std::function<void(const cplx *in, cplx *out)> apply;
    };

    /* Gate formulas: src/qcs.cu:381-1102 (Qiskit conventions, target j = index bit j). */
    static Gate make_gate(const std::string &name, const double *p)
    {
        if (name == "global_phase")
        { // src/qcs.cu:474-487
            cplx e = expi(p[0]);
            return {0, [e](const cplx *in, cplx *out) { out[0] = e * in[0]; }};
        }
        if (name == "h") // src/qcs.cu:489-499
            return {1, [](const cplx *in, cplx *out) { out[0] = SQRT1_2 * (in[0] + in[1]); out[1] = SQRT1_2 * (in[0] - in[1]); }};
        if (name == "x") // src/qcs.cu:509-519
            return {1, [](const cplx *in, cplx *out) { cplx a = in[0]; out[0] = in[1]; out[1] = a; }};
        if (name == "y") // src/qcs.cu:520-532
            return {1, [](const cplx *in, cplx *out) { cplx a = in[0], b = in[1]; out[0] = cplx(b.imag(), -b.real()); out[1] = cplx(-a.imag(), a.real()); }};
        if (name == "z") // src/qcs.cu:535-545
            return {1, [](const cplx *in, cplx *out) { out[1] = -in[1]; }};
        if (name == "s") // src/qcs.cu:547-556
            return {1, [](const cplx *in, cplx *out) { out[1] = mul_i(in[1]); }};
        if (name == "sdg") // src/qcs.cu:558-567
            return {1, [](const cplx *in, cplx *out) { out[1] = mul_mi(in[1]); }};
        if (name == "t") // src/qcs.cu:569-577
            return {1, [](const cplx *in, cplx *out) { out[1] = cplx(SQRT1_2, SQRT1_2) * in[1]; }};
        if (name == "tdg") // src/qcs.cu:579-587
            return {1, [](const cplx *in, cplx *out) { out[1] = cplx(SQRT1_2, -SQRT1_2) * in[1]; }};
        if (name == "sx" || name == "sxdg")
        { // src/qcs.cu:589-620
            double sgn = name == "sx" ? 0.5 : -0.5;
            return {1, [sgn](const cplx *in, cplx *out) {
                        cplx a = 0.5 * (in[0] + in[1]);
                        cplx b = sgn * mul_i(in[0] - in[1]);
                        out[0] = a + b;
                        out[1] = a - b;
                    }};
        }
        if (name == "rx")
        { // src/qcs.cu:622-639
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            return {1, [c, s](const cplx *in, cplx *out) { cplx a = in[0], b = in[1]; out[0] = c * a + mul_i_real(b, -s); out[1] = mul_i_real(a, -s) + c * b; }};
        }
        if (name == "ry")
        { // src/qcs.cu:641-658
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            return {1, [c, s](const cplx *in, cplx *out) { cplx a = in[0], b = in[1]; out[0] = c * a - s * b; out[1] = s * a + c * b; }};
        }
        if (name == "rz")
        { // src/qcs.cu:660-675
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            return {1, [c, s](const cplx *in, cplx *out) { out[0] = cplx(c, -s) * in[0]; out[1] = cplx(c, s) * in[1]; }};
        }
        if (name == "u3" || name == "u")
        { // src/qcs.cu:407-432, 471
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            cplx ep = expi(p[1]), el = expi(p[2]), epl = expi(p[1] + p[2]);
            return {1, [=](const cplx *in, cplx *out) { cplx a = in[0], b = in[1]; out[0] = c * a - s * el * b; out[1] = ep * s * a + epl * c * b; }};
        }
        if (name == "u4")
        { // src/qcs.cu:381-405
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            cplx ep = expi(p[1]), el = expi(p[2]), eg = expi(p[3]);
            return {1, [=](const cplx *in, cplx *out) { cplx a = in[0], b = in[1]; out[0] = eg * (c * a - s * el * b); out[1] = eg * (ep * (s * a + el * c * b)); }};
        }
        if (name == "u2")
        { // src/qcs.cu:434-454
            cplx ep = expi(p[0]), el = expi(p[1]), epl = expi(p[0] + p[1]);
            return {1, [=](const cplx *in, cplx *out) { cplx a = in[0], b = in[1]; out[0] = SQRT1_2 * (a - el * b); out[1] = SQRT1_2 * (ep * a + epl * b); }};
        }
        if (name == "u1" || name == "p")
        { // src/qcs.cu:456-472
            cplx el = expi(p[0]);
            return {1, [el](const cplx *in, cplx *out) { out[1] = el * in[1]; }};
        }
        if (name == "id") // src/qcs.cu:707-719
            return {1, [](const cplx *, cplx *) {}};
        if (name == "r")
        { // src/qcs.cu:721-751
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            cplx ep = expi(p[1]), em = std::conj(ep);
            return {1, [=](const cplx *in, cplx *out) { cplx a = in[0], b = in[1]; out[0] = c * a + mul_i_real(em * b, -s); out[1] = mul_i_real(ep * a, -s) + c * b; }};
        }
        if (name == "swap") // src/qcs.cu:677-689
            return {2, [](const cplx *in, cplx *out) { cplx a = in[1]; out[1] = in[2]; out[2] = a; }};
        if (name == "iswap") // src/qcs.cu:691-705
            return {2, [](const cplx *in, cplx *out) { cplx a = in[1], b = in[2]; out[1] = mul_i(b); out[2] = mul_i(a); }};
        if (name == "rzz")
        { // src/qcs.cu:753-781
            cplx m = cplx(std::cos(0.5 * p[0]), -std::sin(0.5 * p[0])), pl = cplx(std::cos(0.5 * p[0]), std::sin(0.5 * p[0]));
            return {2, [m, pl](const cplx *in, cplx *out) { out[0] = m * in[0]; out[1] = pl * in[1]; out[2] = pl * in[2]; out[3] = m * in[3]; }};
        }
        if (name == "rxx")
        { // src/qcs.cu:783-815
            double c = std::cos(0.5 * p[0]), ms = -std::sin(0.5 * p[0]);
            return {2, [c, ms](const cplx *in, cplx *out) {
                        cplx a0 = in[0], a1 = in[1], a2 = in[2], a3 = in[3];
                        out[0] = c * a0 + mul_i_real(a3, ms);
                        out[1] = c * a1 + mul_i_real(a2, ms);
                        out[2] = c * a2 + mul_i_real(a1, ms);
                        out[3] = c * a3 + mul_i_real(a0, ms);
                    }};
        }
        if (name == "ryy")
        { // src/qcs.cu:817-851
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            return {2, [c, s](const cplx *in, cplx *out) {
                        cplx a0 = in[0], a1 = in[1], a2 = in[2], a3 = in[3];
                        out[0] = c * a0 + mul_i_real(a3, s);
                        out[1] = c * a1 + mul_i_real(a2, -s);
                        out[2] = c * a2 + mul_i_real(a1, -s);
                        out[3] = c * a3 + mul_i_real(a0, s);
                    }};
        }
        if (name == "rzx")
        { // src/qcs.cu:853-886
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            return {2, [c, s](const cplx *in, cplx *out) {
                        cplx a0 = in[0], a1 = in[1], a2 = in[2], a3 = in[3];
                        out[0] = c * a0 + mul_i_real(a2, -s);
                        out[1] = c * a1 + mul_i(a3) * s;
                        out[2] = c * a2 + mul_i_real(a0, -s);
                        out[3] = c * a3 + mul_i(a1) * s;
                    }};
        }
        if (name == "xx_plus_yy" || name == "xx_minus_yy")
        { // src/qcs.cu:888-960
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            cplx eb = expi(p[1]), emb = cplx(std::cos(p[1]), -std::sin(p[1]));
            int lo = name == "xx_plus_yy" ? 1 : 0, hi = name == "xx_plus_yy" ? 2 : 3;
            return {2, [=](const cplx *in, cplx *out) {
                        cplx a = in[lo], b = in[hi];
                        out[lo] = c * a + mul_i_real(emb * b, -s);
                        out[hi] = mul_i_real(eb * a, -s) + c * b;
                    }};
        }
        if (name == "dcx") // src/qcs.cu:962-983
            return {2, [](const cplx *in, cplx *out) { cplx a1 = in[1], a2 = in[2], a3 = in[3]; out[1] = a3; out[2] = a1; out[3] = a2; }};
        if (name == "ecr") // src/qcs.cu:985-1007
            return {2, [](const cplx *in, cplx *out) {
                        cplx a0 = in[0], a1 = in[1], a2 = in[2], a3 = in[3];
                        out[0] = SQRT1_2 * (a1 + mul_i(a3));
                        out[1] = SQRT1_2 * (a0 + mul_mi(a2));
                        out[2] = SQRT1_2 * (mul_i(a1) + a3);
                        out[3] = SQRT1_2 * (mul_mi(a0) + a2);
                    }};
        if (name == "rccx") // src/qcs.cu:1009-1042
            return {3, [](const cplx *in, cplx *out) { cplx a3 = in[3], a5 = in[5], a7 = in[7]; out[3] = mul_mi(a7); out[5] = -a5; out[7] = mul_i(a3); }};
        if (name == "rcccx") // src/qcs.cu:1044-1102
            return {4, [](const cplx *in, cplx *out) { cplx a3 = in[3], a7 = in[7], a11 = in[11], a15 = in[15]; out[3] = mul_i(a3); out[7] = a15; out[11] = mul_mi(a11); out[15] = -a7; }};
        throw std::runtime_error("unknown gate " + name);
    }

    /* Insert a 0 bit at each position of `sorted_ops` (ascending) into `tid`
     * (src/qcs.cu:168-175), then OR in `set_bits` (src/qcs.cu:179-183). */
    static inline uint64_t subspace_base(uint64_t tid, const int *sorted_ops, int nops, uint64_t set_bits)
    {
        uint64_t idx = tid;
        for (int k = 0; k < nops; ++k)
        {
            int q = sorted_ops[k];
            uint64_t low = idx & ((UINT64_C(1) << q) - 1);
            idx = ((idx >> q) << (q + 1)) | low;
        }
        return idx | set_bits;
    }

    static inline double prob(cplx z) { return std::fma(z.real(), z.real(), z.imag() * z.imag()); }

    /* Canonical sum (include/qcs_ext.h): tiles of 4096, stride-halving tree, recursive. */
    static double canon_sum(std::vector<double> x)
    {
        const size_t W = 4096;
        if (x.empty())
            return 0.0;
        while (true)
        {
            size_t ntiles = (x.size() + W - 1) / W;
            std::vector<double> sums(ntiles);
#pragma omp parallel for schedule(static) if (ntiles > 16)
            for (size_t t = 0; t < ntiles; ++t)
            {
                double buf[4096];
                for (size_t i = 0; i < W; ++i)
                {
                    size_t j = t * W + i;
                    buf[i] = j < x.size() ? x[j] : 0.0;
                }
                for (size_t s = W / 2; s >= 1; s /= 2)
                    for (size_t i = 0; i < s; ++i)
                        buf[i] = buf[i] + buf[i + s];
                sums[t] = buf[0];
            }
            if (ntiles == 1)
                return sums[0];
            x.swap(sums);
        }
    }

    static inline uint64_t splitmix64(uint64_t z)
    {
        z = (z ^ (z >> 30)) * UINT64_C(0xBF58476D1CE4E5B9);
        z = (z ^ (z >> 27)) * UINT64_C(0x94D049BB133111EB);
        return z ^ (z >> 31);
    }
    static inline double uniform(uint64_t seed, uint64_t j)
    {
        return (double)(splitmix64(seed + (j + 1) * UINT64_C(0x9E3779B97F4A7C15)) >> 11) * 0x1.0p-53;
    }

    struct Sim
    {
        int num_qubits_cfg = 0;
        int num_clbits = 0;
        std::vector<bool> clbits;
        std::vector<int> initial_p2l;

        bool allocated = false;
        int n = 0;
        int n_local = 0;
        std::vector<cplx> state;
        std::vector<int> p2l, l2p;
        std::vector<int> m0, m1; // measured logical qubits with value 0 / 1
        uint64_t seed = 0;
        std::mt19937_64 engine;
        bool renormalize = false;
        std::vector<std::chrono::steady_clock::time_point> events;
        double prof_seconds[QCS_EXT_K_COUNT] = {};
        uint64_t prof_launches[QCS_EXT_K_COUNT] = {};
        bool profile = false;

        Sim()
        {
            std::random_device rd; // src/qcs.cu:1845-1852
            seed = rd();
            engine.seed(seed);
        }

        void need_alloc() const
        {
            if (!allocated)
                throw std::runtime_error("allocate_memory has not been called");
        }

        void reinitialize_mapping()
        { // src/qcs.cu:1860-1885
            p2l.assign(n, 0);
            l2p.assign(n, 0);
            if (initial_p2l.empty())
            {
                for (int q = 0; q < n; ++q)
                    p2l[q] = l2p[q] = q;
            }
            else
            {
                if ((int)initial_p2l.size() != n)
                    throw std::runtime_error(fmt("mapping size %d does not match num_qubits %d", (int)initial_p2l.size(), n));
                p2l = initial_p2l;
                for (int q = 0; q < n; ++q)
                    l2p[p2l[q]] = q;
            }
        }

        void allocate(int nq)
        { // src/qcs.cu:1887-1998 (without the swap buffer)
            if (allocated)
                throw std::runtime_error(fmt("num_qubits is already set %d > 0", n));
            if (nq < 1 || nq > 40)
                throw std::runtime_error(fmt("num_qubits %d out of range [1, 40]", nq));
            n = nq;
            n_local = nq;
            reinitialize_mapping();
            state.assign(size_t(1) << n, cplx(0.0, 0.0));
            state[0] = 1.0;
            allocated = true;
        }

        void free_memory()
        {
            state.clear();
            state.shrink_to_fit();
            allocated = false;
            n = 0;
        }

        uint64_t logical_index(uint64_t phys) const
        {
            uint64_t l = 0;
            for (int q = 0; q < n; ++q)
                l |= ((phys >> q) & 1) << p2l[q];
            return l;
        }

        /* Swap the data of physical qubits a and b (the in-HBM replacement of the
         * reference's NCCL exchange). */
        void swap_physical(int a, int b)
        {
            if (a == b)
                return;
            int lo = std::min(a, b), hi = std::max(a, b);
            uint64_t nitems = uint64_t(1) << (n - 2);
            int ops[2] = {lo, hi};
#pragma omp parallel for schedule(static)
            for (uint64_t t = 0; t < nitems; ++t)
            {
                uint64_t base = subspace_base(t, ops, 2, 0);
                std::swap(state[base | (uint64_t(1) << lo)], state[base | (uint64_t(1) << hi)]);
            }
        }

        /* src/qcs.cu:2206-2332: a target on a global physical qubit is swapped with local
         * slot n_local-1-i; the maps are updated as at src/qcs.cu:2306-2328. */
        void ensure_local(std::vector<int> &tphys)
        {
            // Slot for the i-th global target: n_local-1-i as in the reference, except that
            // slots holding a local target of the same gate are skipped (the reference would move
            // that target away and apply the gate to the wrong qubit).
            std::vector<int> glob, loc;
            int next_slot = n_local - 1;
            for (auto &t : tphys)
            {
                if (t >= n_local)
                {
                    while (std::find(tphys.begin(), tphys.end(), next_slot) != tphys.end())
                        --next_slot;
                    glob.push_back(t);
                    loc.push_back(next_slot);
                    t = next_slot--;
                }
            }
            for (size_t i = 0; i < glob.size(); ++i)
                swap_physical(glob[i], loc[i]);
            std::vector<int> loc_log(glob.size()), glob_log(glob.size());
            for (size_t i = 0; i < glob.size(); ++i)
            {
                loc_log[i] = p2l[loc[i]];
                glob_log[i] = p2l[glob[i]];
            }
            for (size_t i = 0; i < glob.size(); ++i)
            {
                p2l[glob[i]] = loc_log[i];
                p2l[loc[i]] = glob_log[i];
                l2p[glob_log[i]] = loc[i];
                l2p[loc_log[i]] = glob[i];
            }
        }

        void check_qubits(const std::vector<int> &t, const std::vector<int> &ng, const std::vector<int> &ps) const
        {
            std::vector<int> all;
            all.insert(all.end(), t.begin(), t.end());
            all.insert(all.end(), ng.begin(), ng.end());
            all.insert(all.end(), ps.begin(), ps.end());
            for (int q : all)
                if (q < 0 || q >= n)
                    throw std::runtime_error(fmt("qubit %d out of range [0, %d)", q, n));
            std::sort(all.begin(), all.end());
            if (std::adjacent_find(all.begin(), all.end()) != all.end())
                throw std::runtime_error("a qubit appears more than once among targets and controls");
        }

        /* The per-gate pass (src/qcs.cu:1125-1640): tphys are physical targets, pos/neg
         * physical controls, meas_mask/meas_val the measured targets (bit j = target j). */
        void apply_gate(const Gate &g, const std::vector<int> &tphys, const std::vector<int> &pos, const std::vector<int> &neg, unsigned meas_mask, unsigned meas_val)
        {
            auto t0 = std::chrono::steady_clock::now();
            std::vector<int> ops;
            ops.insert(ops.end(), pos.begin(), pos.end());
            ops.insert(ops.end(), neg.begin(), neg.end());
            ops.insert(ops.end(), tphys.begin(), tphys.end());
            std::sort(ops.begin(), ops.end());
            uint64_t setb = 0;
            for (int q : pos)
                setb |= uint64_t(1) << q;
            const int T = g.ntargets, K = 1 << T;
            uint64_t offs[16];
            for (int k = 0; k < K; ++k)
            {
                offs[k] = 0;
                for (int j = 0; j < T; ++j)
                    if ((k >> j) & 1)
                        offs[k] |= uint64_t(1) << tphys[j];
            }
            bool readable[16];
            for (int k = 0; k < K; ++k)
                readable[k] = ((unsigned(k) ^ meas_val) & meas_mask) == 0;
            uint64_t nitems = uint64_t(1) << (n - (int)ops.size());
            const int *opsp = ops.data();
            const int nops = (int)ops.size();
#pragma omp parallel for schedule(static)
            for (uint64_t tid = 0; tid < nitems; ++tid)
            {
                uint64_t base = subspace_base(tid, opsp, nops, setb);
                cplx in[16], out[16];
                for (int k = 0; k < K; ++k)
                {
                    out[k] = state[base | offs[k]];
                    in[k] = readable[k] ? out[k] : cplx(0.0, 0.0);
                }
                g.apply(in, out);
                for (int k = 0; k < K; ++k)
                    state[base | offs[k]] = out[k];
            }
            note(QCS_EXT_K_PAIR, t0);
        }

        void note(int cls, std::chrono::steady_clock::time_point t0)
        {
            if (!profile)
                return;
            prof_seconds[cls] += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            prof_launches[cls] += 1;
        }

        static bool contains(const std::vector<int> &v, int x) { return std::find(v.begin(), v.end(), x) != v.end(); }

        /* src/qcs.cu:2679-2703 with src/qcs.cu:2161-2204 and 2373-2492. */
        void operate(const Gate &g, std::vector<int> targets, std::vector<int> neg, std::vector<int> pos)
        {
            need_alloc();
            if ((int)targets.size() != g.ntargets)
                throw std::runtime_error(fmt("gate needs %d target qubit(s), %d given", g.ntargets, (int)targets.size()));
            check_qubits(targets, neg, pos);
            // lazy measurement: measured, non-targeted qubits become controls of their value
            std::vector<int> *ctrl[2] = {&neg, &pos};
            std::vector<int> *meas[2] = {&m0, &m1};
            std::vector<int> keep[2];
            bool condition = true;
            for (int v = 0; v < 2; ++v)
            {
                for (int q : *meas[v])
                {
                    if (contains(targets, q))
                        continue;
                    keep[v].push_back(q);
                    bool same = contains(*ctrl[v], q), other = contains(*ctrl[1 - v], q);
                    if (!same && !other)
                        ctrl[v]->push_back(q);
                    else if (other)
                        condition = false;
                }
            }
            if (!condition)
                return; // contradicting control: the gate does nothing (src/qcs.cu:2690-2691)
            std::vector<int> tphys(targets.size()), pphys, nphys;
            for (size_t j = 0; j < targets.size(); ++j)
                tphys[j] = l2p[targets[j]];
            ensure_local(tphys);
            for (int q : pos)
                pphys.push_back(l2p[q]);
            for (int q : neg)
                nphys.push_back(l2p[q]);
            unsigned mmask = 0, mval = 0;
            for (size_t j = 0; j < targets.size(); ++j)
            {
                if (contains(m0, targets[j]))
                    mmask |= 1u << j;
                else if (contains(m1, targets[j]))
                {
                    mmask |= 1u << j;
                    mval |= 1u << j;
                }
            }
            apply_gate(g, tphys, pphys, nphys, mmask, mval);
            m0 = keep[0];
            m1 = keep[1];
        }

        /* src/qcs.cu:2545-2677 */
        int measure(int q)
        {
            need_alloc();
            if (q < 0 || q >= n)
                throw std::runtime_error(fmt("qubit %d out of range [0, %d)", q, n));
            if (contains(m0, q))
                return 0;
            if (contains(m1, q))
                return 1;
            auto t0 = std::chrono::steady_clock::now();
            std::vector<int> tphys = {l2p[q]};
            ensure_local(tphys);
            int tq = tphys[0];
            std::vector<int> ops = {tq};
            uint64_t setb = 0;
            for (int x : m1)
            {
                ops.push_back(l2p[x]);
                setb |= uint64_t(1) << l2p[x];
            }
            for (int x : m0)
                ops.push_back(l2p[x]);
            std::sort(ops.begin(), ops.end());
            uint64_t nitems = uint64_t(1) << (n - (int)ops.size());
            std::vector<double> v0(nitems), v1(nitems);
#pragma omp parallel for schedule(static)
            for (uint64_t t = 0; t < nitems; ++t)
            {
                uint64_t b = subspace_base(t, ops.data(), (int)ops.size(), setb);
                v0[t] = prob(state[b]);
                v1[t] = prob(state[b | (uint64_t(1) << tq)]);
            }
            double p0 = canon_sum(std::move(v0)), p1 = canon_sum(std::move(v1));
            note(QCS_EXT_K_REDUCE, t0);
            std::uniform_real_distribution<double> dist(0, p0 + p1);
            double r = dist(engine);
            int result = p0 < r;
            (result ? m1 : m0).push_back(q);
            t0 = std::chrono::steady_clock::now();
            collapse(tq, result, result ? p1 : p0);
            note(QCS_EXT_K_COLLAPSE, t0);
            return result;
        }

        /* src/qcs.cu:1106-1123 (+ optional FP64 renormalisation of the kept branch). */
        void collapse(int tq, int result, double pkept)
        {
            uint64_t half = uint64_t(1) << (n - 1);
            int ops[1] = {tq};
            double scale = (renormalize && pkept > 0) ? 1.0 / std::sqrt(pkept) : 1.0;
#pragma omp parallel for schedule(static)
            for (uint64_t t = 0; t < half; ++t)
            {
                uint64_t b = subspace_base(t, ops, 1, 0);
                uint64_t kill = result ? b : (b | (uint64_t(1) << tq));
                uint64_t keep = result ? (b | (uint64_t(1) << tq)) : b;
                state[kill] = 0.0;
                if (renormalize)
                    state[keep] = cplx(state[keep].real() * scale, state[keep].imag() * scale);
            }
        }

        std::vector<double> probs_physical() const
        {
            std::vector<double> p(state.size());
#pragma omp parallel for schedule(static)
            for (size_t i = 0; i < state.size(); ++i)
                p[i] = prob(state[i]);
            return p;
        }

        /* Level totals of the 64-ary right-nested prefix (include/qcs_ext.h). */
        static std::vector<std::vector<double>> fold_levels(std::vector<double> p)
        {
            std::vector<std::vector<double>> lv;
            lv.push_back(std::move(p));
            while (lv.back().size() > 1)
            {
                const auto &cur = lv.back();
                size_t nb = (cur.size() + 63) / 64;
                std::vector<double> nxt(nb);
#pragma omp parallel for schedule(static) if (nb > 1024)
                for (size_t b = 0; b < nb; ++b)
                {
                    double acc = 0.0;
                    size_t e = std::min(cur.size(), b * 64 + 64);
                    for (size_t i = b * 64; i < e; ++i)
                        acc = (i == b * 64) ? cur[i] : acc + cur[i];
                    nxt[b] = acc;
                }
                lv.push_back(std::move(nxt));
            }
            return lv;
        }

        /* Smallest index whose right-nested prefix is >= r. */
        static uint64_t search(const std::vector<std::vector<double>> &lv, double r)
        {
            int L = (int)lv.size() - 1; // lv[L] has one entry
            double offs[64];            // E values on the path, outermost first
            int depth = 0;
            uint64_t block = 0;         // index of the current block at level `lev`
            for (int lev = L - 1; lev >= 0; --lev)
            {
                const auto &cur = lv[lev];
                uint64_t b0 = block * 64, b1 = std::min<uint64_t>(cur.size(), b0 + 64);
                double E = 0.0;
                uint64_t pick = b1 - 1;
                for (uint64_t j = b0; j < b1; ++j)
                {
                    double incl = (j == b0) ? cur[j] : E + cur[j];
                    // value of the boundary at the end of child j: offs nested around incl
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
                // E now holds the sum of the children before `pick` (0 if none)
                double Epick = 0.0;
                for (uint64_t j = b0; j < pick; ++j)
                    Epick = (j == b0) ? cur[j] : Epick + cur[j];
                offs[depth++] = Epick;
                block = pick;
            }
            return block;
        }

        double probability_total() const
        {
            auto lv = fold_levels(probs_physical());
            return lv.back()[0];
        }

        void sample(uint64_t sd, uint64_t shots, uint64_t *out) const
        {
            auto lv = fold_levels(probs_physical());
            double total = lv.back()[0];
#pragma omp parallel for schedule(dynamic, 16)
            for (uint64_t j = 0; j < shots; ++j)
            {
                double r = uniform(sd, j) * total;
                out[j] = logical_index(search(lv, r));
            }
        }
    };
} // namespace oracle

struct qcs_simulator
{
    oracle::Sim *core;
};

/* Cost-model MLP on the CPU (F -> 64 -> 64 -> 1, ReLU, FP32): the same functions the GPU
 * library computes, written as plain loops; used to check the GPU forward and gradients. */
struct qcs_ext_mlp
{
    static constexpr int H = 64;
    int F;
    std::vector<float> p, m, v;
    int step = 0;
    int w1() const { return 0; }
    int b1() const { return H * F; }
    int w2() const { return H * F + H; }
    int b2() const { return H * F + H + H * H; }
    int w3() const { return H * F + 2 * H + H * H; }
    int b3() const { return H * F + 3 * H + H * H; }
    int P() const { return b3() + 1; }
    float forward(const float *x, float *h1, float *h2) const
    {
        for (int i = 0; i < H; ++i)
        {
            float a = p[b1() + i];
            for (int k = 0; k < F; ++k)
                a = std::fma(p[w1() + i * F + k], x[k], a);
            h1[i] = a > 0.f ? a : 0.f;
        }
        for (int i = 0; i < H; ++i)
        {
            float a = p[b2() + i];
            for (int j = 0; j < H; ++j)
                a = std::fma(p[w2() + i * H + j], h1[j], a);
            h2[i] = a > 0.f ? a : 0.f;
        }
        float y = p[b3()];
        for (int j = 0; j < H; ++j)
            y = std::fma(p[w3() + j], h2[j], y);
        return y;
    }
    float gradients(const float *X, const float *T, int B, float *G) const
    {
        std::vector<float> g(P(), 0.f);
        double loss = 0;
        for (int b = 0; b < B; ++b)
        {
            float h1[H], h2[H];
            const float y = forward(X + size_t(b) * F, h1, h2);
            loss += double(y - T[b]) * double(y - T[b]);
            const float dy = 2.f * (y - T[b]) / float(B);
            float d2[H], d1[H];
            for (int i = 0; i < H; ++i)
                d2[i] = h2[i] > 0.f ? dy * p[w3() + i] : 0.f;
            for (int j = 0; j < H; ++j)
            {
                float a = 0.f;
                for (int i = 0; i < H; ++i)
                    a = std::fma(p[w2() + i * H + j], d2[i], a);
                d1[j] = h1[j] > 0.f ? a : 0.f;
            }
            for (int i = 0; i < H; ++i)
            {
                for (int k = 0; k < F; ++k)
                    g[w1() + i * F + k] = std::fma(d1[i], X[size_t(b) * F + k], g[w1() + i * F + k]);
                g[b1() + i] += d1[i];
                for (int j = 0; j < H; ++j)
                    g[w2() + i * H + j] = std::fma(d2[i], h1[j], g[w2() + i * H + j]);
                g[b2() + i] += d2[i];
                g[w3() + i] = std::fma(dy, h2[i], g[w3() + i]);
            }
            g[b3()] += dy;
        }
        if (G)
            std::copy(g.begin(), g.end(), G);
        return float(loss / B);
    }
};

/* Batching form in the oracle: the plain method, circuit after circuit. */
struct qcs_ext_batch
{
    struct Gate
    {
        std::string name;
        int target, control;
        double params[4];
    };
    struct Circuit
    {
        int n;
        std::vector<Gate> gates;
    };
    std::vector<Circuit> circuits;
    int total_qubits = 0;
};

namespace
{
    qcs_exception_callback g_callback = nullptr;

    void report(const char *m)
    {
        if (g_callback)
            g_callback(m, std::strlen(m));
    }

    template <typename F>
    int guard(F &&f)
    {
        try
        {
            f();
            return 1;
        }
        catch (const std::exception &e)
        {
            report(e.what());
        }
        catch (...)
        {
            report("Unknown C++ exception");
        }
        return 0;
    }

    template <typename R, typename V>
    void put(R *r, V v)
    {
        if (r == nullptr)
            throw std::runtime_error("result pointer must not be NULL");
        *r = static_cast<R>(v);
    }

    std::vector<int> vec(const bit_num_t *a, bit_num_t c)
    { // src/qcs.cu:2914-2925
        if (c < 0)
            throw std::runtime_error("array element count must be non-negative");
        if (c > 0 && a == nullptr)
            throw std::runtime_error("array pointer must not be NULL when element count is positive");
        return std::vector<int>(a, a + c);
    }

    oracle::Sim &core(const qcs_simulator *s)
    {
        if (s == nullptr || s->core == nullptr)
            throw std::runtime_error("simulator is NULL or disposed");
        return *s->core;
    }
} // namespace

/* ---- C++ entry points ---- */
qcs_simulator *qcs_simulator_create_cxx(void) { return new qcs_simulator{new oracle::Sim()}; }
void qcs_simulator_destroy_cxx(qcs_simulator *sim)
{
    if (!sim)
        return;
    delete sim->core;
    delete sim;
}
void qcs_simulator_allocate_memory_cxx(qcs_simulator *sim) { core(sim).allocate(core(sim).num_qubits_cfg); }
void qcs_simulator_dispose_cxx(qcs_simulator *sim)
{
    if (sim && sim->core)
    {
        delete sim->core;
        sim->core = nullptr;
    }
}
int qcs_simulator_get_num_procs_cxx(qcs_simulator *sim) { (void)core(sim); return 1; }
int qcs_simulator_get_proc_num_cxx(qcs_simulator *sim) { (void)core(sim); return 0; }
int qcs_simulator_get_num_qubits_cxx(const qcs_simulator *sim) { return core(sim).num_qubits_cfg; }
int qcs_simulator_get_num_clbits_cxx(const qcs_simulator *sim) { return core(sim).num_clbits; }
void qcs_simulator_set_num_qubits_cxx(qcs_simulator *sim, bit_num_t n) { core(sim).num_qubits_cfg = n; }
void qcs_simulator_set_mapping_cxx(qcs_simulator *sim, const bit_num_t *perm, bit_num_t count)
{ // src/qcs.cu:3045-3062
    auto p = vec(perm, count);
    auto &c = core(sim);
    if (c.num_qubits_cfg <= 0)
        throw std::runtime_error("set_num_qubits must be called before set_mapping");
    if ((int)p.size() != c.num_qubits_cfg)
        throw std::runtime_error(oracle::fmt("mapping size %d does not match num_qubits %d", (int)p.size(), c.num_qubits_cfg));
    std::vector<bool> used(c.num_qubits_cfg, false);
    for (int l : p)
    {
        if (l < 0 || l >= c.num_qubits_cfg)
            throw std::runtime_error(oracle::fmt("mapping value %d is out of range [0, %d)", l, c.num_qubits_cfg));
        if (used[l])
            throw std::runtime_error(oracle::fmt("mapping value %d appears multiple times", l));
        used[l] = true;
    }
    c.initial_p2l = p;
}
void qcs_simulator_set_num_clbits_cxx(qcs_simulator *sim, bit_num_t nc)
{
    if (nc < 0)
        throw std::runtime_error("num_clbits must be non-negative");
    core(sim).num_clbits = nc;
    core(sim).clbits.resize(nc);
}
void qcs_simulator_get_clbits_cxx(const qcs_simulator *sim, bit_t *clbits)
{
    auto &c = core(sim);
    if (c.num_clbits > 0 && clbits == nullptr)
        throw std::runtime_error("clbits pointer must not be NULL");
    for (int i = 0; i < c.num_clbits; ++i)
        clbits[i] = c.clbits[i] ? 1 : 0;
}
void qcs_simulator_get_clbits_string_cxx(const qcs_simulator *sim, char *s)
{ // src/qcs.cu:3371-3377: clbit n-1 first
    auto &c = core(sim);
    if (s == nullptr)
        throw std::runtime_error("clbits_string pointer must not be NULL");
    int o = 0;
    for (int i = c.num_clbits - 1; i >= 0; --i)
        s[o++] = c.clbits[i] ? '1' : '0';
    s[o] = '\0';
}
void qcs_simulator_reset_cxx(qcs_simulator *sim, bit_num_t q)
{ // src/qcs.cu:3077-3081
    if (qcs_simulator_measure_cxx(sim, q))
        qcs_simulator_gate_x_cxx(sim, &q, 1, nullptr, 0, nullptr, 0);
}
void qcs_simulator_set_zero_state_cxx(qcs_simulator *sim)
{ // src/qcs.cu:2065-2079
    auto &c = core(sim);
    c.need_alloc();
    c.m0.clear();
    c.m1.clear();
    std::fill(c.state.begin(), c.state.end(), oracle::cplx(0.0, 0.0));
    c.state[0] = 1.0;
}
void qcs_simulator_set_sequential_state_cxx(qcs_simulator *sim)
{ // src/qcs.cu:71-76, 2025-2043
    auto &c = core(sim);
    c.need_alloc();
    c.m0.clear();
    c.m1.clear();
    for (size_t i = 0; i < c.state.size(); ++i)
        c.state[i] = oracle::cplx((double)i, 0.0);
}
void qcs_simulator_set_flat_state_cxx(qcs_simulator *sim)
{ // src/qcs.cu:78-82, 2045-2063
    auto &c = core(sim);
    c.need_alloc();
    c.m0.clear();
    c.m1.clear();
    std::fill(c.state.begin(), c.state.end(), oracle::cplx(1.0, 0.0));
}
void qcs_simulator_set_entangled_state_cxx(qcs_simulator *sim)
{ // src/qcs.cu:2081-2095 with one rank: amp[0] = 1, nothing cleared
    auto &c = core(sim);
    c.need_alloc();
    c.m0.clear();
    c.m1.clear();
    c.state[0] = 1.0;
}
void qcs_simulator_set_random_state_cxx(qcs_simulator *sim)
{ // src/qcs.cu:2097-2119 uses curand normals; this build uses its own generator
  // (SplitMix64 + Box-Muller, seeded by the measurement seed).  Measured lists are kept,
  // as in the reference.
    auto &c = core(sim);
    c.need_alloc();
    uint64_t sd = c.seed;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < c.state.size(); ++i)
    {
        double u1 = oracle::uniform(sd, 2 * i), u2 = oracle::uniform(sd, 2 * i + 1);
        double rad = std::sqrt(-2.0 * std::log(1.0 - u1));
        c.state[i] = oracle::cplx(rad * std::cos(2.0 * M_PI * u2), rad * std::sin(2.0 * M_PI * u2));
    }
}
void qcs_simulator_reset_clbits_cxx(qcs_simulator *sim)
{
    auto &c = core(sim);
    std::fill(c.clbits.begin(), c.clbits.end(), false);
}
void qcs_simulator_reset_measurement_state_cxx(qcs_simulator *sim)
{
    auto &c = core(sim);
    c.m0.clear();
    c.m1.clear();
}
void qcs_simulator_reinitialize_mapping_cxx(qcs_simulator *sim)
{
    auto &c = core(sim);
    c.need_alloc();
    c.reinitialize_mapping();
}

static void run_gate(qcs_simulator *sim, const char *name, const double *params, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps)
{
    auto tv = vec(t, nt), nv = vec(ng, nng), pv = vec(ps, nps);
    core(sim).operate(oracle::make_gate(name, params), tv, nv, pv);
}

void qcs_simulator_gate_global_phase_cxx(qcs_simulator *sim, double theta, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps)
{
    double p[1] = {theta};
    run_gate(sim, "global_phase", p, nullptr, 0, ng, nng, ps, nps);
}

#define ORACLE_GATE0(NAME)                                                                                                                               \
    void qcs_simulator_gate_##NAME##_cxx(qcs_simulator *sim, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps) \
    {                                                                                                                                                    \
        run_gate(sim, #NAME, nullptr, t, nt, ng, nng, ps, nps);                                                                                          \
    }
#define ORACLE_GATE1(NAME, A)                                                                                                                                      \
    void qcs_simulator_gate_##NAME##_cxx(qcs_simulator *sim, double A, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps) \
    {                                                                                                                                                              \
        double p[1] = {A};                                                                                                                                         \
        run_gate(sim, #NAME, p, t, nt, ng, nng, ps, nps);                                                                                                          \
    }
#define ORACLE_GATE2(NAME, A, B)                                                                                                                                             \
    void qcs_simulator_gate_##NAME##_cxx(qcs_simulator *sim, double A, double B, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps) \
    {                                                                                                                                                                        \
        double p[2] = {A, B};                                                                                                                                                \
        run_gate(sim, #NAME, p, t, nt, ng, nng, ps, nps);                                                                                                                    \
    }
#define ORACLE_GATE3(NAME, A, B, C)                                                                                                                                                    \
    void qcs_simulator_gate_##NAME##_cxx(qcs_simulator *sim, double A, double B, double C, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps) \
    {                                                                                                                                                                                  \
        double p[3] = {A, B, C};                                                                                                                                                       \
        run_gate(sim, #NAME, p, t, nt, ng, nng, ps, nps);                                                                                                                              \
    }

ORACLE_GATE0(h)
ORACLE_GATE0(x)
ORACLE_GATE0(y)
ORACLE_GATE0(z)
ORACLE_GATE0(s)
ORACLE_GATE0(sdg)
ORACLE_GATE0(t)
ORACLE_GATE0(tdg)
ORACLE_GATE0(sx)
ORACLE_GATE0(sxdg)
ORACLE_GATE0(id)
ORACLE_GATE0(swap)
ORACLE_GATE0(iswap)
ORACLE_GATE0(dcx)
ORACLE_GATE0(ecr)
ORACLE_GATE0(rccx)
ORACLE_GATE0(rcccx)
ORACLE_GATE1(rx, theta)
ORACLE_GATE1(ry, theta)
ORACLE_GATE1(rz, theta)
ORACLE_GATE1(u1, lambda)
ORACLE_GATE1(p, theta)
ORACLE_GATE1(rxx, theta)
ORACLE_GATE1(ryy, theta)
ORACLE_GATE1(rzz, theta)
ORACLE_GATE1(rzx, theta)
ORACLE_GATE2(u2, phi, lambda)
ORACLE_GATE2(r, theta, phi)
ORACLE_GATE2(xx_plus_yy, theta, beta)
ORACLE_GATE2(xx_minus_yy, theta, beta)
ORACLE_GATE3(u3, theta, phi, lambda)
ORACLE_GATE3(u, theta, phi, lambda)

void qcs_simulator_gate_u4_cxx(qcs_simulator *sim, double theta, double phi, double lambda, double gamma, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps)
{
    double p[4] = {theta, phi, lambda, gamma};
    run_gate(sim, "u4", p, t, nt, ng, nng, ps, nps);
}

int qcs_simulator_measure_cxx(qcs_simulator *sim, bit_num_t q) { return core(sim).measure(q); }
int qcs_simulator_measure_to_clbit_cxx(qcs_simulator *sim, bit_num_t q, bit_num_t cb)
{
    auto &c = core(sim);
    if (cb < 0 || cb >= c.num_clbits)
        throw std::runtime_error(oracle::fmt("clbit %d out of range [0, %d)", cb, c.num_clbits));
    int r = c.measure(q);
    c.clbits[cb] = r;
    return r;
}
int qcs_simulator_read_cxx(qcs_simulator *sim, bit_num_t cb)
{
    auto &c = core(sim);
    if (cb < 0 || cb >= c.num_clbits)
        throw std::runtime_error(oracle::fmt("clbit %d out of range [0, %d)", cb, c.num_clbits));
    return c.clbits[cb];
}
void qcs_simulator_save_statevector_cxx(qcs_simulator *sim, const char *fn)
{ // src/qcs.cu:2494-2543: raw little-endian complex<double>, logical order, no header
    auto &c = core(sim);
    c.need_alloc();
    if (fn == nullptr)
        throw std::runtime_error("output file name must not be NULL");
    std::vector<oracle::cplx> out(c.state.size());
    for (uint64_t i = 0; i < c.state.size(); ++i)
        out[c.logical_index(i)] = c.state[i];
    FILE *fp = std::fopen(fn, "wb");
    if (!fp)
        throw std::runtime_error("open failed");
    size_t w = std::fwrite(out.data(), sizeof(oracle::cplx), out.size(), fp);
    std::fclose(fp);
    if (w != out.size())
        throw std::runtime_error("fwrite failed");
}
int qcs_simulator_event_create_cxx(qcs_simulator *sim)
{
    auto &c = core(sim);
    c.events.push_back(std::chrono::steady_clock::now());
    return (int)c.events.size() - 1;
}
void qcs_simulator_event_record_cxx(qcs_simulator *sim, int e)
{
    auto &c = core(sim);
    if (e < 0 || e >= (int)c.events.size())
        throw std::runtime_error(oracle::fmt("event %d does not exist", e));
    c.events[e] = std::chrono::steady_clock::now();
}
double qcs_simulator_event_get_elapsed_time_cxx(qcs_simulator *sim, event_num_t a, event_num_t b)
{
    auto &c = core(sim);
    if (a < 0 || a >= (int)c.events.size() || b < 0 || b >= (int)c.events.size())
        throw std::runtime_error("event does not exist");
    return std::chrono::duration<double>(c.events[b] - c.events[a]).count();
}

/* ---- C entry points ---- */
extern "C"
{
    int qcs_set_exception_callback(qcs_exception_callback cb)
    {
        g_callback = cb;
        return 1;
    }
    int qcs_simulator_create(qcs_simulator **r)
    {
        return guard([&] {
            if (!r)
                throw std::runtime_error("result pointer must not be NULL");
            *r = qcs_simulator_create_cxx();
        });
    }
    int qcs_simulator_destroy(qcs_simulator *s) { return guard([&] { qcs_simulator_destroy_cxx(s); }); }
    int qcs_simulator_allocate_memory(qcs_simulator *s) { return guard([&] { qcs_simulator_allocate_memory_cxx(s); }); }
    int qcs_simulator_dispose(qcs_simulator *s) { return guard([&] { qcs_simulator_dispose_cxx(s); }); }
    int qcs_simulator_get_num_procs(qcs_simulator *s, bit_t *r) { return guard([&] { put(r, qcs_simulator_get_num_procs_cxx(s)); }); }
    int qcs_simulator_get_proc_num(qcs_simulator *s, bit_t *r) { return guard([&] { put(r, qcs_simulator_get_proc_num_cxx(s)); }); }
    int qcs_simulator_get_num_qubits(const qcs_simulator *s, bit_t *r) { return guard([&] { put(r, qcs_simulator_get_num_qubits_cxx(s)); }); }
    int qcs_simulator_get_num_clbits(const qcs_simulator *s, bit_t *r) { return guard([&] { put(r, qcs_simulator_get_num_clbits_cxx(s)); }); }
    int qcs_simulator_set_num_qubits(qcs_simulator *s, bit_num_t n) { return guard([&] { qcs_simulator_set_num_qubits_cxx(s, n); }); }
    int qcs_simulator_set_mapping(qcs_simulator *s, const bit_num_t *p, bit_num_t c) { return guard([&] { qcs_simulator_set_mapping_cxx(s, p, c); }); }
    int qcs_simulator_set_num_clbits(qcs_simulator *s, bit_num_t n) { return guard([&] { qcs_simulator_set_num_clbits_cxx(s, n); }); }
    int qcs_simulator_get_clbits(const qcs_simulator *s, bit_t *c) { return guard([&] { qcs_simulator_get_clbits_cxx(s, c); }); }
    int qcs_simulator_get_clbits_string(const qcs_simulator *s, char *c) { return guard([&] { qcs_simulator_get_clbits_string_cxx(s, c); }); }
    int qcs_simulator_reset(qcs_simulator *s, bit_num_t q) { return guard([&] { qcs_simulator_reset_cxx(s, q); }); }
    int qcs_simulator_set_zero_state(qcs_simulator *s) { return guard([&] { qcs_simulator_set_zero_state_cxx(s); }); }
    int qcs_simulator_set_sequential_state(qcs_simulator *s) { return guard([&] { qcs_simulator_set_sequential_state_cxx(s); }); }
    int qcs_simulator_set_flat_state(qcs_simulator *s) { return guard([&] { qcs_simulator_set_flat_state_cxx(s); }); }
    int qcs_simulator_set_entangled_state(qcs_simulator *s) { return guard([&] { qcs_simulator_set_entangled_state_cxx(s); }); }
    int qcs_simulator_set_random_state(qcs_simulator *s) { return guard([&] { qcs_simulator_set_random_state_cxx(s); }); }
    int qcs_simulator_reset_clbits(qcs_simulator *s) { return guard([&] { qcs_simulator_reset_clbits_cxx(s); }); }
    int qcs_simulator_reset_measurement_state(qcs_simulator *s) { return guard([&] { qcs_simulator_reset_measurement_state_cxx(s); }); }
    int qcs_simulator_reinitialize_mapping(qcs_simulator *s) { return guard([&] { qcs_simulator_reinitialize_mapping_cxx(s); }); }
    int qcs_simulator_gate_global_phase(qcs_simulator *s, double th, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps)
    {
        return guard([&] { qcs_simulator_gate_global_phase_cxx(s, th, ng, nng, ps, nps); });
    }
#define ORACLE_C0(NAME)                                                                                                                       \
    int qcs_simulator_gate_##NAME(qcs_simulator *s, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps) \
    {                                                                                                                                         \
        return guard([&] { qcs_simulator_gate_##NAME##_cxx(s, t, nt, ng, nng, ps, nps); });                                                   \
    }
#define ORACLE_C1(NAME)                                                                                                                                 \
    int qcs_simulator_gate_##NAME(qcs_simulator *s, double a, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps) \
    {                                                                                                                                                   \
        return guard([&] { qcs_simulator_gate_##NAME##_cxx(s, a, t, nt, ng, nng, ps, nps); });                                                          \
    }
#define ORACLE_C2(NAME)                                                                                                                                           \
    int qcs_simulator_gate_##NAME(qcs_simulator *s, double a, double b, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps) \
    {                                                                                                                                                             \
        return guard([&] { qcs_simulator_gate_##NAME##_cxx(s, a, b, t, nt, ng, nng, ps, nps); });                                                                 \
    }
#define ORACLE_C3(NAME)                                                                                                                                                     \
    int qcs_simulator_gate_##NAME(qcs_simulator *s, double a, double b, double c, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps) \
    {                                                                                                                                                                       \
        return guard([&] { qcs_simulator_gate_##NAME##_cxx(s, a, b, c, t, nt, ng, nng, ps, nps); });                                                                        \
    }
    ORACLE_C0(h)
    ORACLE_C0(x)
    ORACLE_C0(y)
    ORACLE_C0(z)
    ORACLE_C0(s)
    ORACLE_C0(sdg)
    ORACLE_C0(t)
    ORACLE_C0(tdg)
    ORACLE_C0(sx)
    ORACLE_C0(sxdg)
    ORACLE_C0(id)
    ORACLE_C0(swap)
    ORACLE_C0(iswap)
    ORACLE_C0(dcx)
    ORACLE_C0(ecr)
    ORACLE_C0(rccx)
    ORACLE_C0(rcccx)
    ORACLE_C1(rx)
    ORACLE_C1(ry)
    ORACLE_C1(rz)
    ORACLE_C1(u1)
    ORACLE_C1(p)
    ORACLE_C1(rxx)
    ORACLE_C1(ryy)
    ORACLE_C1(rzz)
    ORACLE_C1(rzx)
    ORACLE_C2(u2)
    ORACLE_C2(r)
    ORACLE_C2(xx_plus_yy)
    ORACLE_C2(xx_minus_yy)
    ORACLE_C3(u3)
    ORACLE_C3(u)
    int qcs_simulator_gate_u4(qcs_simulator *s, double a, double b, double c, double d, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps)
    {
        return guard([&] { qcs_simulator_gate_u4_cxx(s, a, b, c, d, t, nt, ng, nng, ps, nps); });
    }
    int qcs_simulator_measure(qcs_simulator *s, bit_num_t q, bit_t *r) { return guard([&] { put(r, qcs_simulator_measure_cxx(s, q)); }); }
    int qcs_simulator_measure_to_clbit(qcs_simulator *s, bit_num_t q, bit_num_t c, bit_t *r) { return guard([&] { put(r, qcs_simulator_measure_to_clbit_cxx(s, q, c)); }); }
    int qcs_simulator_read(qcs_simulator *s, bit_num_t c, bit_t *r) { return guard([&] { put(r, qcs_simulator_read_cxx(s, c)); }); }
    int qcs_simulator_save_statevector(qcs_simulator *s, const char *fn) { return guard([&] { qcs_simulator_save_statevector_cxx(s, fn); }); }
    int qcs_simulator_event_create(qcs_simulator *s, bit_t *r) { return guard([&] { put(r, qcs_simulator_event_create_cxx(s)); }); }
    int qcs_simulator_event_record(qcs_simulator *s, int e) { return guard([&] { qcs_simulator_event_record_cxx(s, e); }); }
    int qcs_simulator_event_get_elapsed_time(qcs_simulator *s, event_num_t a, event_num_t b, double *r)
    {
        return guard([&] { put(r, qcs_simulator_event_get_elapsed_time_cxx(s, a, b)); });
    }

    /* ---- extensions (include/qcs_ext.h) ---- */
    int qcs_ext_set_seed(qcs_simulator *s, uint64_t sd)
    {
        return guard([&] {
            auto &c = core(s);
            c.seed = sd;
            c.engine.seed(sd);
        });
    }
    int qcs_ext_get_statevector(qcs_simulator *s, double *out, uint64_t count)
    {
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            if (count != c.state.size() || out == nullptr)
                throw std::runtime_error("statevector buffer size mismatch");
            for (uint64_t i = 0; i < c.state.size(); ++i)
            {
                uint64_t l = c.logical_index(i);
                out[2 * l] = c.state[i].real();
                out[2 * l + 1] = c.state[i].imag();
            }
        });
    }
    int qcs_ext_set_statevector(qcs_simulator *s, const double *in, uint64_t count)
    {
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            if (count != c.state.size() || in == nullptr)
                throw std::runtime_error("statevector buffer size mismatch");
            for (uint64_t i = 0; i < c.state.size(); ++i)
            {
                uint64_t l = c.logical_index(i);
                c.state[i] = oracle::cplx(in[2 * l], in[2 * l + 1]);
            }
            // a new state: earlier measurement results no longer describe it
            c.m0.clear();
            c.m1.clear();
        });
    }
    int qcs_ext_load_statevector(qcs_simulator *s, const char *path)
    { // src/qcs.cu:2121-2159 with the file format of save_statevector (logical order)
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            if (path == nullptr)
                throw std::runtime_error("input file name must not be NULL");
            FILE *fp = std::fopen(path, "rb");
            if (!fp)
                throw std::runtime_error("open failed");
            std::vector<oracle::cplx> in(c.state.size() + 1);
            const size_t got = std::fread(in.data(), sizeof(oracle::cplx), in.size(), fp);
            std::fclose(fp);
            if (got != c.state.size())
                throw std::runtime_error(oracle::fmt("statevector file holds %zu amplitudes, expected %zu", got, c.state.size()));
            for (uint64_t i = 0; i < c.state.size(); ++i)
                c.state[i] = in[c.logical_index(i)];
            c.m0.clear();
            c.m1.clear();
        });
    }
    int qcs_ext_sample(qcs_simulator *s, uint64_t sd, uint64_t shots, uint64_t *out)
    {
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            if (shots > 0 && out == nullptr)
                throw std::runtime_error("outcome buffer must not be NULL");
            c.sample(sd, shots, out);
        });
    }
    int qcs_ext_sample_noalloc(qcs_simulator *s, uint64_t sd, uint64_t shots, uint64_t *out)
    {
        // the limits of libqcs.so's scratch array, then qcs_ext_sample
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            if (c.n > 30)
                throw std::runtime_error(oracle::fmt("sample_noalloc: %d qubits; the scratch array holds the prefix of at most 30", c.n));
            if (shots > 65536)
                throw std::runtime_error("sample_noalloc: more than 65536 shots");
            if (shots > 0 && out == nullptr)
                throw std::runtime_error("outcome buffer must not be NULL");
            c.sample(sd, shots, out);
        });
    }
    int qcs_ext_probability_total(qcs_simulator *s, double *total)
    {
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            put(total, c.probability_total());
        });
    }
    int qcs_ext_norm(qcs_simulator *s, double *norm)
    {
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            put(norm, oracle::canon_sum(c.probs_physical()));
        });
    }
    int qcs_ext_expect_z(qcs_simulator *s, double *out, int count)
    {
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            if (count < 0 || count > c.n || (count > 0 && out == nullptr))
                throw std::runtime_error("expect_z count out of range");
            auto p = c.probs_physical();
            for (int q = 0; q < count; ++q)
            {
                int pq = c.l2p[q];
                std::vector<double> v(p.size());
                for (size_t i = 0; i < p.size(); ++i)
                    v[i] = ((i >> pq) & 1) ? -p[i] : p[i];
                out[q] = oracle::canon_sum(std::move(v));
            }
        });
    }
    int qcs_ext_set_renormalize(qcs_simulator *s, int en)
    {
        return guard([&] { core(s).renormalize = en != 0; });
    }
    int qcs_ext_set_local_qubits(qcs_simulator *s, bit_num_t nl)
    {
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            if (nl < 1 || nl > c.n)
                throw std::runtime_error("num_local_qubits out of range");
            c.n_local = nl;
        });
    }
    int qcs_ext_get_mapping(qcs_simulator *s, bit_num_t *p, bit_num_t count)
    {
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            if (count != c.n || p == nullptr)
                throw std::runtime_error("mapping buffer size mismatch");
            for (int q = 0; q < c.n; ++q)
                p[q] = (bit_num_t)c.p2l[q];
        });
    }
    int qcs_ext_permute(qcs_simulator *s, const bit_num_t *np, bit_num_t count)
    {
        return guard([&] {
            auto &c = core(s);
            c.need_alloc();
            auto want = vec(np, count);
            if ((int)want.size() != c.n)
                throw std::runtime_error("mapping size mismatch");
            std::vector<bool> used(c.n, false);
            for (int l : want)
            {
                if (l < 0 || l >= c.n || used[l])
                    throw std::runtime_error("not a permutation");
                used[l] = true;
            }
            std::vector<oracle::cplx> nxt(c.state.size());
            std::vector<int> wl2p(c.n);
            for (int q = 0; q < c.n; ++q)
                wl2p[want[q]] = q;
            for (uint64_t i = 0; i < c.state.size(); ++i)
            {
                uint64_t j = 0;
                for (int q = 0; q < c.n; ++q)
                    j |= ((i >> q) & 1) << wl2p[c.p2l[q]];
                nxt[j] = c.state[i];
            }
            c.state.swap(nxt);
            c.p2l = want;
            c.l2p = wl2p;
        });
    }
    int qcs_ext_set_stage_bits(qcs_simulator *s, int m)
    { // the oracle has no fused path; the setting is validated and ignored
        return guard([&] {
            (void)core(s);
            if (m != 0 && (m < 11 || m > 14))
                throw std::runtime_error("stage tile bits must be 0, 11, 12, 13 or 14");
        });
    }
    int qcs_ext_set_stage_stream(qcs_simulator *s, int on)
    { // the oracle has no fused path; the setting is validated and ignored
        return guard([&] {
            (void)core(s);
            if (on != 0 && on != 1)
                throw std::runtime_error("stage stream setting must be 0 or 1");
        });
    }
    int qcs_ext_planner_set_check(qcs_simulator *s, int on)
    { // the oracle has no fused path; the setting is validated and ignored
        return guard([&] {
            (void)core(s);
            if (on != 0 && on != 1)
                throw std::runtime_error("planner check setting must be 0 or 1");
        });
    }
    int qcs_ext_set_plan(qcs_simulator *s, int, int, int, int, int, int)
    { // the oracle has no fused path (set_precision refuses every fused mode)
        return guard([&] {
            (void)core(s);
            throw std::runtime_error("the oracle implements only the complex128 plain method");
        });
    }
    int qcs_ext_set_precision(qcs_simulator *s, int mode, int k)
    {
        return guard([&] {
            (void)core(s);
            if (mode != 0)
                throw std::runtime_error("the oracle implements only the complex128 plain method");
            (void)k;
        });
    }
    int qcs_ext_flush(qcs_simulator *s) { return guard([&] { (void)core(s); }); }
    int qcs_ext_fusion_stats(qcs_simulator *s, uint64_t *gates, uint64_t *blocks)
    {
        return guard([&] {
            (void)core(s);
            put(gates, 0);
            put(blocks, 0);
        });
    }
    int qcs_ext_overlap(qcs_simulator *a, qcs_simulator *b, double *out)
    {
        return guard([&] {
            auto &x = core(a), &y = core(b);
            x.need_alloc();
            y.need_alloc();
            if (x.n != y.n || x.p2l != y.p2l || out == nullptr)
                throw std::runtime_error("overlap needs two states of the same size and qubit mapping");
            oracle::cplx acc = 0.0;
            double md = 0.0;
            for (size_t i = 0; i < x.state.size(); ++i)
            {
                acc += std::conj(x.state[i]) * y.state[i];
                md = std::max(md, std::abs(x.state[i] - y.state[i]));
            }
            out[0] = acc.real();
            out[1] = acc.imag();
            out[2] = md;
        });
    }
    int qcs_ext_profile(qcs_simulator *s, int en) { return guard([&] { core(s).profile = en != 0; }); }
    int qcs_ext_profile_get(qcs_simulator *s, int cls, uint64_t *launches, double *seconds, double *bytes)
    {
        return guard([&] {
            auto &c = core(s);
            if (cls < 0 || cls >= QCS_EXT_K_COUNT)
                throw std::runtime_error("kernel class out of range");
            put(launches, c.prof_launches[cls]);
            put(seconds, c.prof_seconds[cls]);
            put(bytes, 0.0);
        });
    }
    int qcs_ext_batch_create(qcs_ext_batch **out)
    {
        return guard([&] {
            if (!out)
                throw std::runtime_error("result pointer must not be NULL");
            *out = new qcs_ext_batch();
        });
    }
    int qcs_ext_batch_destroy(qcs_ext_batch *b) { return guard([&] { delete b; }); }
    int qcs_ext_batch_add_circuit(qcs_ext_batch *b, bit_num_t nq, int32_t ng, const char *const *names, const bit_num_t *targets, const bit_num_t *controls, const double *params)
    {
        return guard([&] {
            if (!b || nq < 1 || nq > 30 || ng < 0 || (ng > 0 && (!names || !targets || !controls || !params)))
                throw std::runtime_error("bad batch circuit");
            qcs_ext_batch::Circuit c{nq, {}};
            for (int i = 0; i < ng; ++i)
            {
                qcs_ext_batch::Gate g{names[i], targets[i], controls[i], {params[4 * i], params[4 * i + 1], params[4 * i + 2], params[4 * i + 3]}};
                if (oracle::make_gate(g.name, g.params).ntargets != 1 || g.target < 0 || g.target >= nq || g.control >= nq || g.control == g.target)
                    throw std::runtime_error("bad batch gate");
                c.gates.push_back(g);
            }
            b->circuits.push_back(c);
            b->total_qubits += nq;
        });
    }
    int qcs_ext_batch_set_tuning(qcs_ext_batch *b, int, int, int, int, int) { return guard([&] { if (!b) throw std::runtime_error("batch is NULL"); }); }
    int qcs_ext_batch_set_precision(qcs_ext_batch *b, int mode)
    {
        return guard([&] {
            if (!b)
                throw std::runtime_error("batch is NULL");
            if (mode != 1 && mode != 3)
                throw std::runtime_error("batch: precision must be 1 or 3");
        });
    }
    int qcs_ext_batch_autotune(qcs_ext_batch *b, int *tuning)
    {
        return guard([&] {
            if (!b || !tuning)
                throw std::runtime_error("batch is NULL");
            tuning[0] = tuning[1] = tuning[2] = tuning[3] = tuning[4] = 0;
        });
    }
    int qcs_ext_batch_run(qcs_ext_batch *b, double *expz, uint64_t count, double *seconds)
    {
        return guard([&] {
            if (!b || (expz && count != (uint64_t)b->total_qubits))
                throw std::runtime_error("bad batch run");
            auto t0 = std::chrono::steady_clock::now();
            size_t off = 0;
            for (const auto &c : b->circuits)
            {
                oracle::Sim sim;
                sim.num_qubits_cfg = c.n;
                sim.allocate(c.n);
                for (const auto &g : c.gates)
                {
                    std::vector<int> pos;
                    if (g.control >= 0)
                        pos.push_back(g.control);
                    sim.operate(oracle::make_gate(g.name, g.params), {g.target}, {}, pos);
                }
                auto p = sim.probs_physical();
                for (int q = 0; q < c.n; ++q)
                {
                    std::vector<double> v(p.size());
                    for (size_t i = 0; i < p.size(); ++i)
                        v[i] = ((i >> q) & 1) ? -p[i] : p[i];
                    if (expz)
                        expz[off + q] = oracle::canon_sum(std::move(v));
                }
                off += c.n;
            }
            if (seconds)
                *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        });
    }
    int qcs_ext_batch_run_direct(qcs_ext_batch *b, double *expz, uint64_t count, double *seconds)
    {
        return qcs_ext_batch_run(b, expz, count, seconds);
    }
    int qcs_ext_batch_run_plain(qcs_ext_batch *b, double *expz, uint64_t count, double *seconds)
    { // the oracle's batch run is already the plain method per circuit
        return qcs_ext_batch_run(b, expz, count, seconds);
    }
    int qcs_ext_batch_info(qcs_ext_batch *b, int *nc, int *tq, uint64_t *nodes)
    {
        return guard([&] {
            if (!b)
                throw std::runtime_error("batch is NULL");
            put(nc, (int)b->circuits.size());
            put(tq, b->total_qubits);
            put(nodes, 0);
        });
    }
    int qcs_ext_batch_traffic(qcs_ext_batch *b, double *bytes, uint64_t *launches)
    {
        return guard([&] {
            if (!b)
                throw std::runtime_error("batch is NULL");
            put(bytes, 0.0);
            put(launches, 0);
        });
    }
    int qcs_ext_mlp_create(qcs_ext_mlp **out, int F, uint64_t seed)
    {
        return guard([&] {
            if (!out || F < 1 || F > 32)
                throw std::runtime_error("bad model arguments");
            auto *m = new qcs_ext_mlp();
            m->F = F;
            m->p.assign(m->P(), 0.f);
            auto fill = [&](int off, int count, int fan_in) {
                const double a = std::sqrt(6.0 / fan_in);
                for (int i = 0; i < count; ++i)
                {
                    uint64_t z = (seed += UINT64_C(0x9E3779B97F4A7C15));
                    double u = double(oracle::splitmix64(z) >> 11) * 0x1.0p-53;
                    m->p[off + i] = float((2 * u - 1) * a);
                }
            };
            fill(m->w1(), 64 * F, F);
            fill(m->w2(), 64 * 64, 64);
            fill(m->w3(), 64, 64);
            m->m.assign(m->P(), 0.f);
            m->v.assign(m->P(), 0.f);
            *out = m;
        });
    }
    int qcs_ext_mlp_destroy(qcs_ext_mlp *m) { return guard([&] { delete m; }); }
    int qcs_ext_mlp_num_params(qcs_ext_mlp *m, int *count)
    {
        return guard([&] {
            if (!m)
                throw std::runtime_error("model is NULL");
            put(count, m->P());
        });
    }
    int qcs_ext_mlp_get_params(qcs_ext_mlp *m, float *p, int count)
    {
        return guard([&] {
            if (!m || !p || count != m->P())
                throw std::runtime_error("parameter buffer size mismatch");
            std::copy(m->p.begin(), m->p.end(), p);
        });
    }
    int qcs_ext_mlp_set_params(qcs_ext_mlp *m, const float *p, int count)
    {
        return guard([&] {
            if (!m || !p || count != m->P())
                throw std::runtime_error("parameter buffer size mismatch");
            m->p.assign(p, p + count);
            m->step = 0;
        });
    }
    int qcs_ext_mlp_predict(qcs_ext_mlp *m, const float *x, int rows, float *y)
    {
        return guard([&] {
            if (!m || rows < 0 || (rows > 0 && (!x || !y)))
                throw std::runtime_error("bad prediction buffers");
            float h1[64], h2[64];
            for (int b = 0; b < rows; ++b)
                y[b] = m->forward(x + size_t(b) * m->F, h1, h2);
        });
    }
    int qcs_ext_mlp_gradients(qcs_ext_mlp *m, const float *x, const float *t, int rows, float *grad, float *loss)
    {
        return guard([&] {
            if (!m || rows < 1 || !x || !t)
                throw std::runtime_error("bad gradient buffers");
            float l = m->gradients(x, t, rows, grad);
            if (loss)
                *loss = l;
        });
    }
    int qcs_ext_mlp_train(qcs_ext_mlp *m, const float *x, const float *t, int rows, int epochs, int batch, float lr, uint64_t seed, float *loss)
    {
        return guard([&] {
            if (!m || rows < 1 || !x || !t || batch < 1 || epochs < 0)
                throw std::runtime_error("bad training arguments");
            batch = std::min(batch, rows);
            std::vector<int> idx(rows);
            for (int i = 0; i < rows; ++i)
                idx[i] = i;
            std::vector<float> xs(size_t(batch) * m->F), ts(batch), g(m->P());
            for (int e = 0; e < epochs; ++e)
            {
                for (int i = rows - 1; i > 0; --i)
                {
                    uint64_t z = (seed += UINT64_C(0x9E3779B97F4A7C15));
                    std::swap(idx[i], idx[oracle::splitmix64(z) % uint64_t(i + 1)]);
                }
                for (int b0 = 0; b0 + batch <= rows; b0 += batch)
                {
                    for (int i = 0; i < batch; ++i)
                    {
                        std::copy(x + size_t(idx[b0 + i]) * m->F, x + size_t(idx[b0 + i] + 1) * m->F, xs.begin() + size_t(i) * m->F);
                        ts[i] = t[idx[b0 + i]];
                    }
                    m->gradients(xs.data(), ts.data(), batch, g.data());
                    ++m->step;
                    const float c1 = 1.f - std::pow(0.9f, float(m->step)), c2 = 1.f - std::pow(0.999f, float(m->step));
                    for (int i = 0; i < m->P(); ++i)
                    {
                        m->m[i] = 0.9f * m->m[i] + 0.1f * g[i];
                        m->v[i] = 0.999f * m->v[i] + 0.001f * g[i] * g[i];
                        m->p[i] -= lr * (m->m[i] / c1) / (std::sqrt(m->v[i] / c2) + 1e-8f);
                    }
                }
            }
            float l = m->gradients(x, t, rows, nullptr);
            if (loss)
                *loss = l;
        });
    }
    int qcs_ext_planner_num_features(int *count) { return guard([&] { put(count, 20); }); }
    int qcs_ext_planner_record(qcs_simulator *, int, uint64_t, float *, float *)
    {
        return guard([&] { throw std::runtime_error("the oracle has no kernels to time"); });
    }
    int qcs_ext_set_planner(qcs_simulator *s, qcs_ext_mlp *) { return guard([&] { (void)core(s); }); }
    int qcs_ext_planner_record_plans(qcs_simulator *, int, float *, float *, int *)
    {
        return guard([&] { throw std::runtime_error("the oracle has no kernels to time"); });
    }
    int qcs_ext_planner_last_choice(qcs_simulator *s, int *plan)
    {
        return guard([&] {
            (void)core(s);
            throw std::runtime_error("the oracle has no fusion planner");
        });
    }
    int qcs_ext_planner_stats(qcs_simulator *s, double *out)
    {
        return guard([&] {
            (void)core(s);
            if (!out)
                throw std::runtime_error("result pointer must not be NULL");
            for (int i = 0; i < 5; ++i)
                out[i] = 0;
        });
    }
    int qcs_ext_profile_reset(qcs_simulator *s)
    {
        return guard([&] {
            auto &c = core(s);
            for (int i = 0; i < QCS_EXT_K_COUNT; ++i)
            {
                c.prof_launches[i] = 0;
                c.prof_seconds[i] = 0;
            }
        });
    }
}
