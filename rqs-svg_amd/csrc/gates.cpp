// Gate table of the host core.  Matrices follow the reference's gate semantics
// (src/qcs.cu:381-1102: Qiskit conventions, target j = bit j of the local index); the
// coefficient formulas use libm cos/sin in double, as the reference's constructors do.
#include "gates.hpp"

#include <cmath>
#include <stdexcept>

namespace qcore
{
    namespace
    {
        cd expi(double a) { return cd(std::cos(a), std::sin(a)); }
        const double kS = 0.70710678118654752440; // 1/sqrt(2)

        GateSpec pair1(const std::string &name, cd a, cd b, cd c, cd d)
        {
            GateSpec g;
            g.name = name;
            g.ntargets = 1;
            g.kind = Kind::Pair1;
            g.m[0] = a;
            g.m[1] = b;
            g.m[2] = c;
            g.m[3] = d;
            return g;
        }

        GateSpec phase_one(const std::string &name, cd ph)
        {
            GateSpec g;
            g.name = name;
            g.ntargets = 1;
            g.kind = Kind::PhaseOne;
            g.phase[0] = 1.0;
            g.phase[1] = ph;
            return g;
        }

        GateSpec pair_sub(const std::string &name, int pa, int pb, cd a, cd b, cd c, cd d)
        {
            GateSpec g = pair1(name, a, b, c, d);
            g.ntargets = 2;
            g.kind = Kind::PairSub;
            g.pat[0] = pa;
            g.pat[1] = pb;
            return g;
        }

        GateSpec dense4(const std::string &name, const cd (&m)[16])
        {
            GateSpec g;
            g.name = name;
            g.ntargets = 2;
            g.kind = Kind::Dense4;
            for (int i = 0; i < 16; ++i)
                g.m[i] = m[i];
            return g;
        }

        GateSpec mono(const std::string &name, int t, std::initializer_list<std::tuple<int, int, cd>> rows)
        {
            GateSpec g;
            g.name = name;
            g.ntargets = t;
            g.kind = Kind::Mono;
            for (auto &r : rows)
            {
                g.row[g.nrows] = std::get<0>(r);
                g.src[g.nrows] = std::get<1>(r);
                g.coef[g.nrows] = std::get<2>(r);
                ++g.nrows;
            }
            return g;
        }
    } // namespace

    GateSpec make_gate(const std::string &name, const double *p)
    {
        const cd I(0.0, 1.0);
        if (name == "global_phase")
        {
            GateSpec g;
            g.name = name;
            g.ntargets = 0;
            g.kind = Kind::GlobalPhase;
            g.phase[0] = expi(p[0]);
            return g;
        }
        if (name == "h")
            return pair1(name, kS, kS, kS, -kS);
        if (name == "x")
            return pair1(name, 0.0, 1.0, 1.0, 0.0);
        if (name == "y")
            return pair1(name, 0.0, -I, I, 0.0);
        if (name == "sx")
            return pair1(name, cd(0.5, 0.5), cd(0.5, -0.5), cd(0.5, -0.5), cd(0.5, 0.5));
        if (name == "sxdg")
            return pair1(name, cd(0.5, -0.5), cd(0.5, 0.5), cd(0.5, 0.5), cd(0.5, -0.5));
        if (name == "rx")
        {
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            return pair1(name, c, cd(0, -s), cd(0, -s), c);
        }
        if (name == "ry")
        {
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            return pair1(name, c, -s, s, c);
        }
        if (name == "u3" || name == "u")
        {
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            return pair1(name, c, -s * expi(p[2]), s * expi(p[1]), c * expi(p[1] + p[2]));
        }
        if (name == "u4")
        {
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            cd eg = expi(p[3]);
            return pair1(name, eg * c, -eg * s * expi(p[2]), eg * s * expi(p[1]), eg * c * expi(p[1]) * expi(p[2]));
        }
        if (name == "u2")
            return pair1(name, kS, -kS * expi(p[1]), kS * expi(p[0]), kS * expi(p[0] + p[1]));
        if (name == "r")
        {
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            return pair1(name, c, -I * s * cd(std::cos(p[1]), -std::sin(p[1])), -I * s * expi(p[1]), c);
        }
        if (name == "z")
            return phase_one(name, -1.0);
        if (name == "s")
            return phase_one(name, I);
        if (name == "sdg")
            return phase_one(name, -I);
        if (name == "t")
            return phase_one(name, cd(kS, kS));
        if (name == "tdg")
            return phase_one(name, cd(kS, -kS));
        if (name == "p" || name == "u1")
            return phase_one(name, expi(p[0]));
        if (name == "rz")
        {
            GateSpec g;
            g.name = name;
            g.ntargets = 1;
            g.kind = Kind::Diag1;
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            g.phase[0] = cd(c, -s);
            g.phase[1] = cd(c, s);
            return g;
        }
        if (name == "id")
        {
            GateSpec g;
            g.name = name;
            g.ntargets = 1;
            g.kind = Kind::Identity;
            return g;
        }
        // two-target gates: local pattern = bit(t0) | bit(t1) << 1
        if (name == "swap")
            return pair_sub(name, 1, 2, 0.0, 1.0, 1.0, 0.0);
        if (name == "iswap")
            return pair_sub(name, 1, 2, 0.0, I, I, 0.0);
        if (name == "xx_plus_yy" || name == "xx_minus_yy")
        {
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            cd eb = expi(p[1]), emb = cd(std::cos(p[1]), -std::sin(p[1]));
            int pa = name == "xx_plus_yy" ? 1 : 0, pb = name == "xx_plus_yy" ? 2 : 3;
            return pair_sub(name, pa, pb, c, -I * s * emb, -I * s * eb, c);
        }
        if (name == "rzz")
        {
            GateSpec g;
            g.name = name;
            g.ntargets = 2;
            g.kind = Kind::Diag2;
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            g.phase[0] = g.phase[3] = cd(c, -s);
            g.phase[1] = g.phase[2] = cd(c, s);
            return g;
        }
        if (name == "rxx" || name == "ryy" || name == "rzx")
        {
            double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
            cd m[16] = {};
            for (int k = 0; k < 4; ++k)
                m[5 * k] = c;
            if (name == "rxx")
            { // c I - i s X(x)X
                m[3] = m[6] = m[9] = m[12] = cd(0, -s);
            }
            else if (name == "ryy")
            { // c I - i s Y(x)Y
                m[3] = m[12] = cd(0, s);
                m[6] = m[9] = cd(0, -s);
            }
            else
            { // exp(-i theta/2 Z_t0 X_t1)
                m[2] = m[8] = cd(0, -s);
                m[7] = m[13] = cd(0, s);
            }
            return dense4(name, m);
        }
        if (name == "ecr")
        {
            cd m[16] = {0.0, kS, 0.0, kS * I, kS, 0.0, -kS * I, 0.0, 0.0, kS * I, 0.0, kS, -kS * I, 0.0, kS, 0.0};
            return dense4(name, m);
        }
        if (name == "dcx")
            return mono(name, 2, {{1, 3, 1.0}, {2, 1, 1.0}, {3, 2, 1.0}});
        if (name == "rccx")
            return mono(name, 3, {{3, 7, -I}, {5, 5, -1.0}, {7, 3, I}});
        if (name == "rcccx")
            return mono(name, 4, {{3, 3, I}, {7, 15, 1.0}, {11, 11, -I}, {15, 7, -1.0}});
        throw std::runtime_error("unknown gate " + name);
    }

    std::vector<cd> dense_matrix(const GateSpec &g)
    {
        const int K = 1 << g.ntargets;
        std::vector<cd> u(K * K, 0.0);
        for (int i = 0; i < K; ++i)
            u[i * K + i] = 1.0;
        switch (g.kind)
        {
        case Kind::GlobalPhase:
            u[0] = g.phase[0];
            break;
        case Kind::PhaseOne:
            u[3] = g.phase[1];
            break;
        case Kind::Pair1:
            for (int i = 0; i < 4; ++i)
                u[i] = g.m[i];
            break;
        case Kind::Diag1:
        case Kind::Diag2:
            for (int i = 0; i < K; ++i)
                u[i * K + i] = g.phase[i];
            break;
        case Kind::Identity:
            break;
        case Kind::PairSub:
        {
            int a = g.pat[0], b = g.pat[1];
            u[a * K + a] = g.m[0];
            u[a * K + b] = g.m[1];
            u[b * K + a] = g.m[2];
            u[b * K + b] = g.m[3];
            break;
        }
        case Kind::Dense4:
            for (int i = 0; i < 16; ++i)
                u[i] = g.m[i];
            break;
        case Kind::Mono:
            for (int k = 0; k < g.nrows; ++k)
            {
                int r = g.row[k];
                for (int c = 0; c < K; ++c)
                    u[r * K + c] = 0.0;
            }
            for (int k = 0; k < g.nrows; ++k)
                u[g.row[k] * K + g.src[k]] = g.coef[k];
            break;
        }
        return u;
    }
} // namespace qcore
