"""GPU parity of the qubit permutation, the in-HBM replacement of the reference's rank
exchange (src/qcs.cu:2206-2332): the LDS-tiled permutation kernel moves the data of
physical qubits, the mapping follows the data, and the logical state must not change at all
(bit-exact: the kernel only moves amplitudes).  Gates after a permutation then run on the
new physical positions and must match the oracle, which keeps its own mapping
(max|d amp| <= 1e-13).  Sizes: 20 and 21 qubits (256-512 tiles of 2^12 complex128 or 2^13
complex64 amplitudes), permutations that move qubits 0..2, need several passes, or move a
single pair; the window mode (global targets swapped in before each gate) at 20 qubits."""
import sys

import numpy as np
import pytest

import pyqcs
from pyqcs import circuits
from conftest import random_state

pytestmark = pytest.mark.gpu

AMP_TOL = 1e-13


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def perms(n, seed):
    rng = np.random.default_rng(seed)
    ident = list(range(n))
    out = {"random": [int(q) for q in rng.permutation(n)], "reversal": ident[::-1]}
    cyc = ident.copy()
    cyc[0], cyc[1], cyc[2] = ident[1], ident[2], ident[0]
    out["cycle012"] = cyc
    sw = ident.copy()
    sw[0], sw[n - 1] = sw[n - 1], sw[0]
    out["swap0_top"] = sw
    lowhigh = ident.copy()  # q0..q2 with three high qubits, plus a high pair
    for a, b in ((0, n - 3), (1, n - 1), (2, n - 5), (9, 14)):
        lowhigh[a], lowhigh[b] = lowhigh[b], lowhigh[a]
    out["low_high"] = lowhigh
    return out


@pytest.mark.parametrize("n", [20, 21])
def test_permute_moves_data_exactly(gpu, oracle, n):
    psi = random_state(n, 11 + n)
    circ = circuits.random_all_gates(n, 40, seed=n)
    for name, want in perms(n, n).items():
        say(f"permute n={n} {name}: {want}")
        outs = []
        for lib in (gpu, oracle):
            with pyqcs.Simulator(lib, n, seed=1) as sim:
                sim.set_statevector(psi)
                before = sim.statevector()
                sim.permute(want)
                assert sim.mapping() == want
                after = sim.statevector()
                assert np.array_equal(before, after), "the permutation changed the logical state"
                sim.run(circ)
                outs.append(sim.statevector())
        assert np.max(np.abs(outs[0] - outs[1])) <= AMP_TOL


@pytest.mark.parametrize("mode", [2, 3])
def test_permute_fused_states_exactly(gpu, mode):
    """complex128 (mode 2) and complex64 (mode 3) states of the precision form."""
    n = 20
    for name, want in perms(n, 5).items():
        say(f"permute fused mode={mode} {name}")
        with pyqcs.Simulator(gpu, n, seed=1, precision=mode, block_qubits=4) as sim:
            sim.set_statevector(random_state(n, 3))
            before = sim.statevector()
            sim.permute(want)
            assert sim.mapping() == want
            assert np.array_equal(before, sim.statevector())


def test_window_swap_ins_match_oracle(gpu, oracle):
    """Global targets (at or above the local window) are swapped in before each gate: one
    pair through the SWAP kernel, several (2-4 target gates) through the permutation
    kernel; mappings and states must equal the oracle's."""
    n = 20
    circ = circuits.random_all_gates(n, 150, seed=21)
    outs = []
    for lib in (gpu, oracle):
        with pyqcs.Simulator(lib, n, seed=1) as sim:
            sim.set_local_qubits(14)
            sim.set_statevector(random_state(n, 4))
            sim.run(circ)
            outs.append((sim.mapping(), sim.statevector()))
    assert outs[0][0] == outs[1][0]
    assert np.max(np.abs(outs[0][1] - outs[1][1])) <= AMP_TOL
