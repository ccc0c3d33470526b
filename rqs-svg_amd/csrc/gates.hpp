// Gate specifications of the host core: coefficients computed in double on the host
// (the reference does the same in each gate's constructor, src/qcs.cu:381-1102) and the
// kernel family each gate runs on.
#pragma once
#include <complex>
#include <string>
#include <vector>

namespace qcore
{
    using cd = std::complex<double>;

    enum class Kind
    {
        GlobalPhase, // t = 0: e^{i theta} on the control subspace
        PhaseOne,    // t = 1, only the |1> branch changes by phase[1]: z s sdg t tdg p u1
        Pair1,       // t = 1, dense 2x2: h x y sx sxdg rx ry u u2 u3 u4 r
        Diag1,       // t = 1, diagonal: rz
        Identity,    // id: bookkeeping only
        PairSub,     // t = 2, 2x2 on two basis patterns: swap iswap xx_plus_yy xx_minus_yy
        Diag2,       // t = 2, diagonal: rzz
        Dense4,      // t = 2, dense 4x4: rxx ryy rzx ecr
        Mono,        // one nonzero per changed row: dcx rccx rcccx
    };

    struct GateSpec
    {
        std::string name;
        int ntargets = 0;
        Kind kind = Kind::Identity;
        cd m[16];         // Pair1/PairSub: m[0..3] = 2x2; Dense4: 4x4 row-major
        cd phase[4];      // GlobalPhase: phase[0]; PhaseOne: phase[1]; Diag1/Diag2: diagonal
        int pat[2] = {0, 0}; // PairSub: the two local basis patterns the 2x2 acts on
        int nrows = 0;    // Mono: rows out[row] = coef * in[src]
        int row[16] = {};
        int src[16] = {};
        cd coef[16];
    };

    // name as in the C API (qcs_simulator_gate_<name>), params in API order.
    GateSpec make_gate(const std::string &name, const double *params);

    // Full 2^t x 2^t matrix (row-major, local index bit j = target j).
    std::vector<cd> dense_matrix(const GateSpec &g);
} // namespace qcore
