// C ABI of libqcs.so (include/qcs.h, include/qcs_ext.h).
//
// Convention of the reference shim (src/qcs.cu:2903-3006, 3383-3759): each extern "C"
// function runs its C++ body under try/catch, returns 1 on success and 0 on any
// exception, and passes the message to the callback installed with
// qcs_set_exception_callback.  Differences, all additive: qubit/clbit indices, operand
// duplicates and target counts are validated (the reference asserts or has undefined
// behaviour, src/qcs.cu:2683, 3073), and dispose followed by destroy is safe.
#include <qcs.h>
#include <qcs_ext.h>

#include "batch.hpp"
#include "gates.hpp"
#include "planner.hpp"
#include "simulator.hpp"

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

struct qcs_simulator
{
    qcore::Simulator *core;
};

struct qcs_ext_batch
{
    qcore::Batch *core;
};

struct qcs_ext_mlp
{
    qcore::CostModel *core;
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

    qcore::Simulator &core(const qcs_simulator *s)
    {
        if (s == nullptr || s->core == nullptr)
            throw std::runtime_error("simulator is NULL or disposed");
        return *s->core;
    }

    void run_gate(qcs_simulator *sim, const char *name, const double *params, const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps)
    {
        auto tv = vec(t, nt), nv = vec(ng, nng), pv = vec(ps, nps);
        core(sim).operate(qcore::make_gate(name, params), tv, nv, pv);
    }

    std::string range_msg(const char *what, int v, int hi)
    {
        return std::string(what) + " " + std::to_string(v) + " out of range [0, " + std::to_string(hi) + ")";
    }
} // namespace

/* ---------------- C++ entry points (include/qcs.h:18-80) ---------------- */
qcs_simulator *qcs_simulator_create_cxx(void) { return new qcs_simulator{new qcore::Simulator()}; }

void qcs_simulator_destroy_cxx(qcs_simulator *sim)
{ // src/qcs.cu:3015-3023
    if (sim == nullptr)
        return;
    delete sim->core;
    delete sim;
}

void qcs_simulator_allocate_memory_cxx(qcs_simulator *sim) { core(sim).allocate(); }

void qcs_simulator_dispose_cxx(qcs_simulator *sim)
{ // src/qcs.cu:3033-3038
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
{ // src/qcs.cu:3045-3062: validated now, applied at allocate / reinitialize_mapping
    auto p = vec(perm, count);
    auto &c = core(sim);
    if (c.num_qubits_cfg <= 0)
        throw std::runtime_error("set_num_qubits must be called before set_mapping");
    if ((int)p.size() != c.num_qubits_cfg)
        throw std::runtime_error("mapping size " + std::to_string(p.size()) + " does not match num_qubits " + std::to_string(c.num_qubits_cfg));
    std::vector<bool> used(c.num_qubits_cfg, false);
    for (int l : p)
    {
        if (l < 0 || l >= c.num_qubits_cfg)
            throw std::runtime_error(range_msg("mapping value", l, c.num_qubits_cfg));
        if (used[l])
            throw std::runtime_error("mapping value " + std::to_string(l) + " appears multiple times");
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
{ // src/qcs.cu:3077-3081: measure, then X if the result is 1
    if (qcs_simulator_measure_cxx(sim, q))
        qcs_simulator_gate_x_cxx(sim, &q, 1, nullptr, 0, nullptr, 0);
}

void qcs_simulator_set_zero_state_cxx(qcs_simulator *sim) { core(sim).set_zero_state(); }
void qcs_simulator_set_sequential_state_cxx(qcs_simulator *sim) { core(sim).set_sequential_state(); }
void qcs_simulator_set_flat_state_cxx(qcs_simulator *sim) { core(sim).set_flat_state(); }
void qcs_simulator_set_entangled_state_cxx(qcs_simulator *sim) { core(sim).set_entangled_state(); }
void qcs_simulator_set_random_state_cxx(qcs_simulator *sim) { core(sim).set_random_state(); }

void qcs_simulator_reset_clbits_cxx(qcs_simulator *sim)
{
    auto &c = core(sim);
    std::fill(c.clbits.begin(), c.clbits.end(), false);
}

void qcs_simulator_reset_measurement_state_cxx(qcs_simulator *sim) { core(sim).clear_measurement_state(); }

void qcs_simulator_reinitialize_mapping_cxx(qcs_simulator *sim)
{
    auto &c = core(sim);
    if (!c.allocated())
        throw std::runtime_error("allocate_memory has not been called");
    c.flush();
    c.reinitialize_mapping();
}

void qcs_simulator_gate_global_phase_cxx(qcs_simulator *sim, double theta, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps)
{
    double p[1] = {theta};
    run_gate(sim, "global_phase", p, nullptr, 0, ng, nng, ps, nps);
}

#define QCS_GATE_ARGS const bit_num_t *t, bit_num_t nt, const bit_num_t *ng, bit_num_t nng, const bit_num_t *ps, bit_num_t nps
#define QCS_GATE0(NAME)                                                          \
    void qcs_simulator_gate_##NAME##_cxx(qcs_simulator *sim, QCS_GATE_ARGS)      \
    {                                                                            \
        run_gate(sim, #NAME, nullptr, t, nt, ng, nng, ps, nps);                  \
    }
#define QCS_GATE1(NAME, A)                                                          \
    void qcs_simulator_gate_##NAME##_cxx(qcs_simulator *sim, double A, QCS_GATE_ARGS) \
    {                                                                               \
        double p[1] = {A};                                                          \
        run_gate(sim, #NAME, p, t, nt, ng, nng, ps, nps);                           \
    }
#define QCS_GATE2(NAME, A, B)                                                                 \
    void qcs_simulator_gate_##NAME##_cxx(qcs_simulator *sim, double A, double B, QCS_GATE_ARGS) \
    {                                                                                         \
        double p[2] = {A, B};                                                                 \
        run_gate(sim, #NAME, p, t, nt, ng, nng, ps, nps);                                     \
    }
#define QCS_GATE3(NAME, A, B, C)                                                                        \
    void qcs_simulator_gate_##NAME##_cxx(qcs_simulator *sim, double A, double B, double C, QCS_GATE_ARGS) \
    {                                                                                                   \
        double p[3] = {A, B, C};                                                                        \
        run_gate(sim, #NAME, p, t, nt, ng, nng, ps, nps);                                               \
    }

QCS_GATE0(h)
QCS_GATE0(x)
QCS_GATE0(y)
QCS_GATE0(z)
QCS_GATE0(s)
QCS_GATE0(sdg)
QCS_GATE0(t)
QCS_GATE0(tdg)
QCS_GATE0(sx)
QCS_GATE0(sxdg)
QCS_GATE0(id)
QCS_GATE0(swap)
QCS_GATE0(iswap)
QCS_GATE0(dcx)
QCS_GATE0(ecr)
QCS_GATE0(rccx)
QCS_GATE0(rcccx)
QCS_GATE1(rx, theta)
QCS_GATE1(ry, theta)
QCS_GATE1(rz, theta)
QCS_GATE1(u1, lambda)
QCS_GATE1(p, theta)
QCS_GATE1(rxx, theta)
QCS_GATE1(ryy, theta)
QCS_GATE1(rzz, theta)
QCS_GATE1(rzx, theta)
QCS_GATE2(u2, phi, lambda)
QCS_GATE2(r, theta, phi)
QCS_GATE2(xx_plus_yy, theta, beta)
QCS_GATE2(xx_minus_yy, theta, beta)
QCS_GATE3(u3, theta, phi, lambda)
QCS_GATE3(u, theta, phi, lambda)

void qcs_simulator_gate_u4_cxx(qcs_simulator *sim, double theta, double phi, double lambda, double gamma, QCS_GATE_ARGS)
{
    double p[4] = {theta, phi, lambda, gamma};
    run_gate(sim, "u4", p, t, nt, ng, nng, ps, nps);
}

int qcs_simulator_measure_cxx(qcs_simulator *sim, bit_num_t q) { return core(sim).measure(q); }

int qcs_simulator_measure_to_clbit_cxx(qcs_simulator *sim, bit_num_t q, bit_num_t cb)
{
    auto &c = core(sim);
    if (cb < 0 || cb >= c.num_clbits)
        throw std::runtime_error(range_msg("clbit", cb, c.num_clbits));
    int r = c.measure(q);
    c.clbits[cb] = r;
    return r;
}

int qcs_simulator_read_cxx(qcs_simulator *sim, bit_num_t cb)
{
    auto &c = core(sim);
    if (cb < 0 || cb >= c.num_clbits)
        throw std::runtime_error(range_msg("clbit", cb, c.num_clbits));
    return c.clbits[cb];
}

void qcs_simulator_save_statevector_cxx(qcs_simulator *sim, const char *fn) { core(sim).save_statevector(fn); }
int qcs_simulator_event_create_cxx(qcs_simulator *sim) { return core(sim).event_create(); }
void qcs_simulator_event_record_cxx(qcs_simulator *sim, int e) { core(sim).event_record(e); }
double qcs_simulator_event_get_elapsed_time_cxx(qcs_simulator *sim, event_num_t a, event_num_t b) { return core(sim).event_elapsed(a, b); }

/* ---------------- C entry points (include/qcs.h:85-148) ---------------- */
extern "C"
{
    int qcs_set_exception_callback(qcs_exception_callback cb)
    { // always succeeds, src/qcs.cu:3002-3006
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
#define QCS_C0(NAME)                                                                                        \
    int qcs_simulator_gate_##NAME(qcs_simulator *s, QCS_GATE_ARGS)                                          \
    {                                                                                                       \
        return guard([&] { qcs_simulator_gate_##NAME##_cxx(s, t, nt, ng, nng, ps, nps); });                 \
    }
#define QCS_C1(NAME)                                                                                        \
    int qcs_simulator_gate_##NAME(qcs_simulator *s, double a, QCS_GATE_ARGS)                                \
    {                                                                                                       \
        return guard([&] { qcs_simulator_gate_##NAME##_cxx(s, a, t, nt, ng, nng, ps, nps); });              \
    }
#define QCS_C2(NAME)                                                                                        \
    int qcs_simulator_gate_##NAME(qcs_simulator *s, double a, double b, QCS_GATE_ARGS)                      \
    {                                                                                                       \
        return guard([&] { qcs_simulator_gate_##NAME##_cxx(s, a, b, t, nt, ng, nng, ps, nps); });           \
    }
#define QCS_C3(NAME)                                                                                        \
    int qcs_simulator_gate_##NAME(qcs_simulator *s, double a, double b, double c, QCS_GATE_ARGS)            \
    {                                                                                                       \
        return guard([&] { qcs_simulator_gate_##NAME##_cxx(s, a, b, c, t, nt, ng, nng, ps, nps); });        \
    }
    QCS_C0(h)
    QCS_C0(x)
    QCS_C0(y)
    QCS_C0(z)
    QCS_C0(s)
    QCS_C0(sdg)
    QCS_C0(t)
    QCS_C0(tdg)
    QCS_C0(sx)
    QCS_C0(sxdg)
    QCS_C0(id)
    QCS_C0(swap)
    QCS_C0(iswap)
    QCS_C0(dcx)
    QCS_C0(ecr)
    QCS_C0(rccx)
    QCS_C0(rcccx)
    QCS_C1(rx)
    QCS_C1(ry)
    QCS_C1(rz)
    QCS_C1(u1)
    QCS_C1(p)
    QCS_C1(rxx)
    QCS_C1(ryy)
    QCS_C1(rzz)
    QCS_C1(rzx)
    QCS_C2(u2)
    QCS_C2(r)
    QCS_C2(xx_plus_yy)
    QCS_C2(xx_minus_yy)
    QCS_C3(u3)
    QCS_C3(u)
    int qcs_simulator_gate_u4(qcs_simulator *s, double a, double b, double c, double d, QCS_GATE_ARGS)
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

    /* ---------------- extensions (include/qcs_ext.h) ---------------- */
    int qcs_ext_set_seed(qcs_simulator *s, uint64_t sd) { return guard([&] { core(s).set_seed(sd); }); }
    int qcs_ext_get_statevector(qcs_simulator *s, double *out, uint64_t n) { return guard([&] { core(s).get_statevector(out, n); }); }
    int qcs_ext_set_statevector(qcs_simulator *s, const double *in, uint64_t n) { return guard([&] { core(s).set_statevector(in, n); }); }
    int qcs_ext_load_statevector(qcs_simulator *s, const char *path) { return guard([&] { core(s).load_statevector(path); }); }
    int qcs_ext_sample(qcs_simulator *s, uint64_t sd, uint64_t shots, uint64_t *out) { return guard([&] { core(s).sample(sd, shots, out); }); }
    int qcs_ext_sample_noalloc(qcs_simulator *s, uint64_t sd, uint64_t shots, uint64_t *out) { return guard([&] { core(s).sample_noalloc(sd, shots, out); }); }
    int qcs_ext_probability_total(qcs_simulator *s, double *total) { return guard([&] { put(total, core(s).probability_total()); }); }
    int qcs_ext_norm(qcs_simulator *s, double *norm) { return guard([&] { put(norm, core(s).norm()); }); }
    int qcs_ext_expect_z(qcs_simulator *s, double *out, int count) { return guard([&] { core(s).expect_z(out, count); }); }
    int qcs_ext_set_renormalize(qcs_simulator *s, int en) { return guard([&] { core(s).set_renormalize(en != 0); }); }
    int qcs_ext_set_local_qubits(qcs_simulator *s, bit_num_t nl) { return guard([&] { core(s).set_local_qubits(nl); }); }
    int qcs_ext_get_mapping(qcs_simulator *s, bit_num_t *p, bit_num_t count) { return guard([&] { core(s).get_mapping(p, count); }); }
    int qcs_ext_permute(qcs_simulator *s, const bit_num_t *p, bit_num_t count) { return guard([&] { core(s).permute(vec(p, count)); }); }
    int qcs_ext_set_stage_bits(qcs_simulator *s, int m) { return guard([&] { core(s).set_stage_bits(m); }); }
    int qcs_ext_set_stage_stream(qcs_simulator *s, int on) { return guard([&] { core(s).set_stage_stream(on); }); }
    int qcs_ext_planner_set_check(qcs_simulator *s, int on) { return guard([&] { core(s).set_planner_check(on); }); }
    int qcs_ext_set_precision(qcs_simulator *s, int mode, int k) { return guard([&] { core(s).set_precision(mode, k); }); }
    int qcs_ext_flush(qcs_simulator *s) { return guard([&] { core(s).flush(); }); }
    int qcs_ext_set_plan(qcs_simulator *s, int en, int bits, int k, int tiles, int window, int rule)
    {
        return guard([&] { core(s).set_plan(en != 0, bits, k, tiles, window, rule); });
    }
    int qcs_ext_fusion_stats(qcs_simulator *s, uint64_t *gates, uint64_t *blocks)
    {
        return guard([&] {
            if (!gates || !blocks)
                throw std::runtime_error("result pointer must not be NULL");
            core(s).fusion_stats(gates, blocks);
        });
    }
    int qcs_ext_overlap(qcs_simulator *a, qcs_simulator *b, double *out)
    {
        return guard([&] {
            if (!out)
                throw std::runtime_error("result pointer must not be NULL");
            core(a).overlap(core(b), out);
        });
    }
    int qcs_ext_profile(qcs_simulator *s, int en) { return guard([&] { core(s).profile_enable(en != 0); }); }
    int qcs_ext_profile_get(qcs_simulator *s, int cls, uint64_t *launches, double *seconds, double *bytes)
    {
        return guard([&] { core(s).profile_get(cls, launches, seconds, bytes); });
    }
    int qcs_ext_profile_reset(qcs_simulator *s) { return guard([&] { core(s).profile_reset(); }); }

    /* ---------------- learned planner ---------------- */
    int qcs_ext_mlp_create(qcs_ext_mlp **out, int F, uint64_t seed)
    {
        return guard([&] {
            if (!out)
                throw std::runtime_error("result pointer must not be NULL");
            *out = new qcs_ext_mlp{new qcore::CostModel(F, seed)};
        });
    }
    int qcs_ext_mlp_destroy(qcs_ext_mlp *m)
    {
        return guard([&] {
            if (!m)
                return;
            delete m->core;
            delete m;
        });
    }
    static qcore::CostModel &model(qcs_ext_mlp *m)
    {
        if (!m || !m->core)
            throw std::runtime_error("model is NULL");
        return *m->core;
    }
    int qcs_ext_mlp_num_params(qcs_ext_mlp *m, int *count) { return guard([&] { put(count, model(m).num_params()); }); }
    int qcs_ext_mlp_get_params(qcs_ext_mlp *m, float *p, int count)
    {
        return guard([&] {
            if (!p || count != model(m).num_params())
                throw std::runtime_error("parameter buffer size mismatch");
            model(m).get_params(p);
        });
    }
    int qcs_ext_mlp_set_params(qcs_ext_mlp *m, const float *p, int count)
    {
        return guard([&] {
            if (!p || count != model(m).num_params())
                throw std::runtime_error("parameter buffer size mismatch");
            model(m).set_params(p);
        });
    }
    int qcs_ext_mlp_predict(qcs_ext_mlp *m, const float *x, int rows, float *y)
    {
        return guard([&] {
            if (rows < 0 || (rows > 0 && (!x || !y)))
                throw std::runtime_error("bad prediction buffers");
            model(m).predict(x, rows, y);
        });
    }
    int qcs_ext_mlp_gradients(qcs_ext_mlp *m, const float *x, const float *t, int rows, float *grad, float *loss)
    {
        return guard([&] {
            if (rows < 1 || !x || !t)
                throw std::runtime_error("bad gradient buffers");
            float l = model(m).gradients(x, t, rows, grad);
            if (loss)
                *loss = l;
        });
    }
    int qcs_ext_mlp_train(qcs_ext_mlp *m, const float *x, const float *t, int rows, int epochs, int batch, float lr, uint64_t seed, float *loss)
    {
        return guard([&] {
            if (rows < 1 || !x || !t)
                throw std::runtime_error("bad training buffers");
            float l = model(m).train(x, t, rows, epochs, batch, lr, seed);
            if (loss)
                *loss = l;
        });
    }
    int qcs_ext_planner_num_features(int *count) { return guard([&] { put(count, qcore::kPlanFeatures); }); }
    int qcs_ext_planner_record(qcs_simulator *s, int count, uint64_t seed, float *features, float *log2_us)
    {
        return guard([&] { core(s).planner_record(count, seed, features, log2_us); });
    }
    int qcs_ext_set_planner(qcs_simulator *s, qcs_ext_mlp *m)
    {
        return guard([&] {
            if (m && m->core && m->core->num_features() != qcore::kPlanFeatures)
                throw std::runtime_error("the planner model must take 16 features");
            core(s).set_planner_model(m ? m->core : nullptr);
        });
    }
    int qcs_ext_planner_record_plans(qcs_simulator *s, int capacity, float *features, float *log2_us, int *count)
    {
        return guard([&] { put(count, core(s).planner_record_plans(capacity, features, log2_us)); });
    }
    int qcs_ext_planner_last_choice(qcs_simulator *s, int *plan)
    {
        return guard([&] {
            if (!plan)
                throw std::runtime_error("result pointer must not be NULL");
            core(s).planner_last_choice(plan);
        });
    }
    int qcs_ext_planner_stats(qcs_simulator *s, double *out)
    {
        return guard([&] {
            if (!out)
                throw std::runtime_error("result pointer must not be NULL");
            core(s).planner_stats(out);
        });
    }

    /* ---------------- batching form ---------------- */
    int qcs_ext_batch_create(qcs_ext_batch **out)
    {
        return guard([&] {
            if (!out)
                throw std::runtime_error("result pointer must not be NULL");
            *out = new qcs_ext_batch{new qcore::Batch()};
        });
    }
    int qcs_ext_batch_destroy(qcs_ext_batch *b)
    {
        return guard([&] {
            if (!b)
                return;
            delete b->core;
            delete b;
        });
    }
    int qcs_ext_batch_add_circuit(qcs_ext_batch *b, bit_num_t nq, int32_t ng, const char *const *names, const bit_num_t *targets, const bit_num_t *controls, const double *params)
    {
        return guard([&] {
            if (!b || !b->core)
                throw std::runtime_error("batch is NULL");
            if (ng < 0 || (ng > 0 && (!names || !targets || !controls || !params)))
                throw std::runtime_error("gate arrays must not be NULL");
            std::vector<qcore::BatchGate> gates(ng);
            for (int i = 0; i < ng; ++i)
            {
                if (!names[i])
                    throw std::runtime_error("gate name must not be NULL");
                gates[i].name = names[i];
                gates[i].target = targets[i];
                gates[i].control = controls[i];
                for (int j = 0; j < 4; ++j)
                    gates[i].params[j] = params[4 * i + j];
            }
            b->core->add_circuit(nq, gates);
        });
    }
    int qcs_ext_batch_set_tuning(qcs_ext_batch *b, int lds_max_qubits, int block_size, int block_qubits, int streams, int stage_bits)
    {
        return guard([&] {
            if (!b || !b->core)
                throw std::runtime_error("batch is NULL");
            qcore::BatchTuning t;
            t.lds_max_qubits = lds_max_qubits;
            t.block_size = block_size;
            t.block_qubits = block_qubits;
            t.streams = streams;
            t.stage_bits = stage_bits;
            b->core->set_tuning(t);
        });
    }
    int qcs_ext_batch_set_precision(qcs_ext_batch *b, int mode)
    {
        return guard([&] {
            if (!b || !b->core)
                throw std::runtime_error("batch is NULL");
            b->core->set_precision(mode);
        });
    }
    int qcs_ext_batch_autotune(qcs_ext_batch *b, int *tuning)
    {
        return guard([&] {
            if (!b || !b->core || !tuning)
                throw std::runtime_error("batch or result pointer is NULL");
            auto t = b->core->autotune();
            tuning[0] = t.lds_max_qubits;
            tuning[1] = t.block_size;
            tuning[2] = t.block_qubits;
            tuning[3] = t.streams;
            tuning[4] = t.stage_bits;
        });
    }
    int qcs_ext_batch_run(qcs_ext_batch *b, double *expz, uint64_t count, double *seconds)
    {
        return guard([&] {
            if (!b || !b->core)
                throw std::runtime_error("batch is NULL");
            if (expz && count != (uint64_t)b->core->total_qubits())
                throw std::runtime_error("output size must equal the total number of qubits");
            double s = b->core->run(expz);
            if (seconds)
                *seconds = s;
        });
    }
    int qcs_ext_batch_run_direct(qcs_ext_batch *b, double *expz, uint64_t count, double *seconds)
    {
        return guard([&] {
            if (!b || !b->core)
                throw std::runtime_error("batch is NULL");
            if (expz && count != (uint64_t)b->core->total_qubits())
                throw std::runtime_error("output size must equal the total number of qubits");
            double s = b->core->run_direct(expz);
            if (seconds)
                *seconds = s;
        });
    }
    int qcs_ext_batch_run_plain(qcs_ext_batch *b, double *expz, uint64_t count, double *seconds)
    {
        return guard([&] {
            if (!b || !b->core)
                throw std::runtime_error("batch is NULL");
            if (expz && count != (uint64_t)b->core->total_qubits())
                throw std::runtime_error("output size must equal the total number of qubits");
            double s = b->core->run_plain(expz);
            if (seconds)
                *seconds = s;
        });
    }
    int qcs_ext_batch_info(qcs_ext_batch *b, int *num_circuits, int *total_qubits, uint64_t *graph_nodes)
    {
        return guard([&] {
            if (!b || !b->core)
                throw std::runtime_error("batch is NULL");
            put(num_circuits, b->core->num_circuits());
            put(total_qubits, b->core->total_qubits());
            put(graph_nodes, b->core->graph_nodes());
        });
    }
    int qcs_ext_batch_traffic(qcs_ext_batch *b, double *hbm_bytes, uint64_t *kernel_launches)
    {
        return guard([&] {
            if (!b || !b->core)
                throw std::runtime_error("batch is NULL");
            put(hbm_bytes, b->core->hbm_bytes());
            put(kernel_launches, b->core->kernel_launches());
        });
    }
}
