/* GHZ on NUM_QUBITS qubits, then every qubit measured into its clbit: the shape of the
 * reference example examples/standalone/ghz_circuit.c (H on q0, CNOT q0 -> qi). */
#include <qcs.h>
#include <stdio.h>
#include <stdlib.h>

static bit_num_t n;

void circuit_init(qcs_simulator *sim)
{
    const char *s = getenv("NUM_QUBITS");
    if (s == NULL || s[0] == '\0')
    {
        fprintf(stderr, "NUM_QUBITS is empty\n");
        exit(EXIT_FAILURE);
    }
    n = (bit_num_t)atoi(s);
    qcs_simulator_set_num_qubits(sim, n);
    qcs_simulator_set_num_clbits(sim, n);
}

void circuit_run(qcs_simulator *sim)
{
    bit_num_t t0[] = {0};
    qcs_simulator_gate_h(sim, t0, 1, NULL, 0, NULL, 0);
    for (bit_num_t q = 1; q < n; q++)
    {
        bit_num_t t[] = {q};
        bit_num_t c[] = {0};
        qcs_simulator_gate_x(sim, t, 1, NULL, 0, c, 1);
    }
    for (bit_num_t q = 0; q < n; q++)
    {
        bit_t m = 0;
        qcs_simulator_measure_to_clbit(sim, q, q, &m);
    }
}
