/* A measurement-free circuit with asymmetric gates, used to compare saved state vectors
 * under different qubit mappings. */
#include <qcs.h>
#include <stdlib.h>

static bit_num_t n;

void circuit_init(qcs_simulator *sim)
{
    n = (bit_num_t)atoi(getenv("NUM_QUBITS"));
    qcs_simulator_set_num_qubits(sim, n);
    qcs_simulator_set_num_clbits(sim, 0);
}

void circuit_run(qcs_simulator *sim)
{
    for (bit_num_t q = 0; q < n; q++)
    {
        bit_num_t t[] = {q};
        qcs_simulator_gate_u3(sim, 0.3 + q, 0.7 * q, 1.1 - 0.2 * q, t, 1, NULL, 0, NULL, 0);
    }
    for (bit_num_t q = 0; q + 1 < n; q++)
    {
        bit_num_t t[] = {q, (bit_num_t)(q + 1)};
        qcs_simulator_gate_rzx(sim, 0.4 * (q + 1), t, 2, NULL, 0, NULL, 0);
    }
    bit_num_t t3[] = {0, 1, 2};
    bit_num_t c[] = {3};
    qcs_simulator_gate_rccx(sim, t3, 3, NULL, 0, c, 1);
}
