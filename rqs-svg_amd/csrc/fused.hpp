// Precision form: complex64 state with gates fused into k-qubit blocks.
#pragma once
#include "simulator.hpp"

namespace qcore
{
    class FusedEngine
    {
    public:
        FusedEngine(Simulator &sim, int max_block_qubits);
        ~FusedEngine();
        void enqueue(const ResolvedGate &g);
        void flush();
        void fill(int kind);
        void swap_physical(int a, int b);
        int measure(int q);
        void get_statevector(double *out);
        void set_statevector(const double *in);
        double probability_total();
        void sample(uint64_t seed, uint64_t shots, uint64_t *out);
        double norm();
        void expect_z(double *out, int count);

    private:
        Simulator &sim_;
        int k_;
    };
} // namespace qcore
