// Precision form (complex64 fused blocks): not available in this build yet.
#include "fused.hpp"

#include <stdexcept>

namespace qcore
{
    FusedEngine::FusedEngine(Simulator &sim, int k) : sim_(sim), k_(k)
    {
        throw std::runtime_error("precision mode 1 is not available in this build");
    }
    FusedEngine::~FusedEngine() {}
    void FusedEngine::enqueue(const ResolvedGate &) {}
    void FusedEngine::flush() {}
    void FusedEngine::fill(int) {}
    void FusedEngine::swap_physical(int, int) {}
    int FusedEngine::measure(int) { return 0; }
    void FusedEngine::get_statevector(double *) {}
    void FusedEngine::set_statevector(const double *) {}
    double FusedEngine::probability_total() { return 0; }
    void FusedEngine::sample(uint64_t, uint64_t, uint64_t *) {}
    double FusedEngine::norm() { return 0; }
    void FusedEngine::expect_z(double *, int) {}
} // namespace qcore
